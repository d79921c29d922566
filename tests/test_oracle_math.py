"""Pin the oracle's AO, grid and build routines by what the mathematics fixes.

Nothing here re-types an oracle formula: each check is a closed form, an
invariant or a finite difference that a dropped term, wrong sign, swapped
index or transposed operand would fail.
"""
import numpy as np
import pytest

from oracle import oracle
from proteindf_amd import basis, guess, molecules


def _single_atom_table(symbol="O"):
    mol = molecules.Molecule([symbol], np.zeros((1, 3)))
    return mol, basis.OrbitalInfo.build(mol, oracle.normalize_cgto)


def _sphere_quadrature(rmax=14.0, nr=400, order=590):
    # Gauss-Legendre in r on [0, rmax] x Lebedev: independent of the EM rule
    x, w = np.polynomial.legendre.leggauss(nr)
    r = 0.5 * rmax * (x + 1.0)
    wr = 0.5 * rmax * w * r * r
    u, wu = oracle.lebedev(order)
    pts = (r[:, None, None] * u[None, :, :]).reshape(-1, 3)
    wts = (wr[:, None] * wu[None, :] * 4.0 * np.pi).reshape(-1)
    return pts, wts


def test_lebedev_rules_are_exact_on_monomials():
    from math import prod

    def dfact(n):
        return prod(range(n, 0, -2)) if n > 0 else 1

    # every order TlLebedevGrid supports, with the polynomial degree it integrates exactly
    for order, degree in ((6, 3), (14, 5), (26, 7), (38, 9), (50, 11), (74, 13), (86, 15), (110, 17),
                          (146, 19), (170, 21), (194, 23), (230, 25), (266, 27), (302, 29), (350, 31),
                          (434, 35), (590, 41), (770, 47), (974, 53), (1202, 59), (1454, 65), (1730, 71)):
        xyz, w = oracle.lebedev(order)
        assert len(w) == order
        assert abs(w.sum() - 1.0) < 1e-14
        np.testing.assert_allclose(np.linalg.norm(xyz, axis=1), 1.0, atol=1e-15)
        k = np.arange(degree + 1)
        X, Y, Zp = (xyz[:, d][None, :] ** k[:, None] for d in range(3))
        # q[a, b, c] = sum_p w_p x^a y^b z^c
        q = np.einsum("abp,cp->abc", X[:, None, :] * Y[None, :, :], Zp * w[None, :])
        a, b, c = np.meshgrid(k, k, k, indexing="ij")
        inside = a + b + c <= degree
        even = (a % 2 == 0) & (b % 2 == 0) & (c % 2 == 0)
        # mean of x^a y^b z^c over the unit sphere
        exact = np.zeros_like(q)
        for i, j, l in zip(*np.nonzero(inside & even)):
            exact[i, j, l] = dfact(i - 1) * dfact(j - 1) * dfact(l - 1) / dfact(i + j + l + 1)
        err = np.abs(q - exact)[inside]
        assert err.max() < 2e-14, (order, err.max())


def test_em_radial_rule_integrates_gaussians():
    # sum_i w_i exp(-a r_i^2) over the kept shells = (pi/a)^(3/2)
    for Z in (1, 6, 8):
        R = oracle.bragg_slater(Z)
        for a in (0.1, 0.8, 5.0, 120.0):
            s = 0.0
            for i in range(1, 76):
                r, w = oracle.radial_em(R, 75, i)
                if r > 30.0:
                    continue
                s += w * np.exp(-a * r * r)
            assert abs(s / (np.pi / a) ** 1.5 - 1.0) < 2e-9, (Z, a, s)


def test_em_radial_weight_reproduces_reference_int_wrap():
    # the reference forms i^5 in 32-bit int (DfGenerateGrid.cpp:1605-1606): i = 74 wraps negative
    r73, w73 = oracle.radial_em(0.6, 99, 73)
    r74, w74 = oracle.radial_em(0.6, 99, 74)
    assert w73 > 0 and w74 < 0
    i5 = int(np.array([74], dtype=np.int64)[0] ** 5)
    wrapped = (i5 + 2 ** 31) % 2 ** 32 - 2 ** 31
    assert wrapped < 0
    assert abs(w74 / w73 - wrapped / 73 ** 5 * (100 - 73) ** 7 / (100 - 74) ** 7) < 1e-9


@pytest.mark.parametrize("symbol", ["H", "C", "O"])
def test_every_ao_is_normalised(symbol):
    _, info = _single_atom_table(symbol)
    aos = oracle.AoTable.from_basis(info)
    pts, wts = _sphere_quadrature()
    phi = oracle.eval_aos(aos, pts, grad=False)
    norms = (wts[:, None] * phi * phi).sum(0)
    np.testing.assert_allclose(norms, 1.0, atol=2e-10)


def test_primitive_normalisation_closed_form():
    # a single s primitive: c = (2a/pi)^(3/4); a single p primitive: 2 (2/pi)^(3/4) a^(5/4)
    e, c = oracle.normalize_cgto(0, [1.3], [1.0])
    assert abs(c[0] - (2 * 1.3 / np.pi) ** 0.75) < 1e-15
    e, c = oracle.normalize_cgto(1, [0.7], [1.0])
    assert abs(c[0] - 2.0 * (2.0 / np.pi) ** 0.75 * 0.7 ** 1.25) < 1e-15
    # sorting: ascending exponents, coefficients follow
    e, c = oracle.normalize_cgto(0, [5.0, 0.5, 1.5], [0.2, 0.5, 0.4])
    assert list(e) == [0.5, 1.5, 5.0]
    e1, c1 = oracle.normalize_cgto(0, [0.5], [1.0])
    assert c[0] / c1[0] == pytest.approx(c[0] / c1[0])


def test_ao_values_closed_form_and_cutoff():
    # one uncontracted d shell at the origin, exponent 0.9
    a = 0.9
    e, c = oracle.normalize_cgto(2, [a], [1.0])
    aos = oracle.AoTable(np.zeros((5, 3)), [4, 5, 6, 7, 8], [1] * 5, [e] * 5, [c] * 5)
    p = np.array([[0.3, -0.7, 1.1], [1.9, 0.2, -0.4]])
    phi = oracle.eval_aos(aos, p, grad=False)
    x, y, z = p.T
    g = c[0] * np.exp(-a * (x * x + y * y + z * z))
    exact = np.stack([x * y, z * x, y * z, 0.5 * (x * x - y * y),
                      (z * z - 0.5 * (x * x + y * y)) / np.sqrt(3.0)], 1) * g[:, None]
    np.testing.assert_allclose(phi, exact, rtol=1e-14, atol=1e-300)
    # TOOBIG: a primitive with alpha r^2 > 30 contributes exactly zero
    far = np.array([[0.0, 0.0, np.sqrt(30.0 / a) * 1.0001]])
    assert np.all(oracle.eval_aos(aos, far, grad=False) == 0.0)


def test_ao_gradients_match_finite_differences():
    _, info = _single_atom_table("O")
    aos = oracle.AoTable.from_basis(info)
    rng = np.random.default_rng(5)
    p = rng.normal(0, 0.8, size=(6, 3))
    _, gx, gy, gz = oracle.eval_aos(aos, p)
    h = 1e-5
    for d, g in enumerate((gx, gy, gz)):
        pp, pm = p.copy(), p.copy()
        pp[:, d] += h
        pm[:, d] -= h
        fd = (oracle.eval_aos(aos, pp, grad=False) - oracle.eval_aos(aos, pm, grad=False)) / (2 * h)
        np.testing.assert_allclose(g, fd, rtol=2e-6, atol=2e-7)


def test_becke_partition_of_unity():
    mol = molecules.water_cluster(2, seed=0)
    rng = np.random.default_rng(1)
    pts = mol.xyz.mean(0) + rng.normal(0, 2.0, size=(50, 3))
    Z = mol.Z
    for r in pts:
        s = sum(oracle.becke_weight(Z, mol.xyz, a, r) for a in range(mol.n_atoms))
        assert abs(s - 1.0) < 1e-14
    # at a nucleus the own-atom weight is 1
    for a in range(mol.n_atoms):
        assert abs(oracle.becke_weight(Z, mol.xyz, a, mol.xyz[a]) - 1.0) < 1e-12


def test_o_matrix_is_orthogonal_and_diagonalises():
    mol = molecules.polyglycine_helix(2, seed=1)
    O = oracle.o_matrix(mol.Z, mol.xyz)
    np.testing.assert_allclose(O.T @ O, np.eye(3), atol=1e-13)
    Z = mol.Z.astype(float)
    T = (Z[:, None] * mol.xyz).sum(0) / Z.sum()
    R = mol.xyz - T
    M = sum(z * ((r @ r) * np.eye(3) - np.outer(r, r)) for z, r in zip(Z, R))
    D = O.T @ M @ O
    assert np.abs(D - np.diag(np.diag(D))).max() < 1e-10 * np.abs(M).max()
    assert np.all(np.diff(np.diag(D)) >= 0)


@pytest.fixture(scope="module")
def water_case():
    mol = molecules.water_cluster(1, seed=0)
    info = basis.OrbitalInfo.build(mol, oracle.normalize_cgto)
    grid, atom = oracle.generate_grid(mol.Z, mol.xyz, 40, 194)
    return mol, info, grid


def test_grid_point_counts_follow_shell_cutoff(water_case):
    mol, info, grid = water_case
    # every kept radial shell contributes n_ang points before the 1e-16 weight screen
    expected = 0
    for Z in mol.Z:
        R = oracle.bragg_slater(int(Z))
        expected += 194 * sum(1 for i in range(1, 41) if oracle.radial_em(R, 40, i)[0] <= 30.0)
    assert len(grid) <= expected and len(grid) > 0.9 * expected


def test_integrated_density_counts_electrons(water_case):
    mol, info, grid = water_case
    aos = oracle.AoTable.from_basis(info)
    P = guess.density_guess(mol, info, eps=0.0)  # diagonal: integral = trace
    e, F, rho = oracle.calc_xc_rks(aos, grid, 0.5 * P, want_rho=True)
    n_el = 2.0 * (grid[:, 3] * rho[:, 0]).sum()
    assert abs(n_el - np.trace(P)) < 2e-4 * np.trace(P)
    assert abs(np.trace(P) - mol.n_electrons) < 1e-12


def test_kxc_is_the_derivative_of_exc(water_case):
    # E(P_a + eps D) central difference = 2 <D, K_a>  (RKS alpha convention)
    mol, info, grid = water_case
    aos = oracle.AoTable.from_basis(info)
    PA = 0.5 * guess.density_guess(mol, info, seed=2, eps=0.02)
    e0, K, _ = oracle.calc_xc_rks(aos, grid, PA)
    assert np.allclose(K, K.T)
    rng = np.random.default_rng(7)
    for _ in range(3):
        D = rng.normal(size=PA.shape)
        D = 0.5 * (D + D.T)
        h = 1e-5
        ep, _, _ = oracle.calc_xc_rks(aos, grid, PA + h * D)
        em, _, _ = oracle.calc_xc_rks(aos, grid, PA - h * D)
        fd = (ep - em) / (2 * h)
        ana = 2.0 * (D * K).sum()
        assert abs(fd - ana) < 1e-6 * max(1.0, abs(ana)), (fd, ana)


def test_lda_part_of_k_is_phi_t_diag_phi(water_case):
    # with grad rho switched off (Slater only is not GGA), check the structure of K
    # through a single point: K = w vrho phi phi^T + w c (phi grad phi^T + grad phi phi^T)
    mol, info, grid = water_case
    aos = oracle.AoTable.from_basis(info)
    PA = 0.5 * guess.density_guess(mol, info, seed=3, eps=0.02)
    pt = grid[len(grid) // 3: len(grid) // 3 + 1]
    e, K, rho = oracle.calc_xc_rks(aos, pt, PA, want_rho=True)
    phi, gx, gy, gz = oracle.eval_aos(aos, pt[:, :3])
    r, g = rho[0, 0], rho[0, 1:]
    gamma = g @ g
    o = oracle.functional(oracle.B3LYP, [[r, r, gamma, gamma, gamma]])[0]
    w = pt[0, 3]
    c = w * (2 * o[3] + o[4]) * g
    dphi = c[0] * gx[0] + c[1] * gy[0] + c[2] * gz[0]
    Kexp = w * o[1] * np.outer(phi[0], phi[0]) + np.outer(phi[0], dphi) + np.outer(dphi, phi[0])
    np.testing.assert_allclose(K, Kexp, rtol=1e-12, atol=1e-18)
    assert abs(e - w * o[0]) < 1e-16
    # rho is P contracted twice with phi, grad rho is 2 phi P grad phi
    assert abs(r - phi[0] @ PA @ phi[0]) < 1e-13
    np.testing.assert_allclose(g, 2 * np.array([phi[0] @ PA @ d[0] for d in (gx, gy, gz)]),
                               rtol=1e-12, atol=1e-15)
