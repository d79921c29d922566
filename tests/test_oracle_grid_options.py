"""Pin the oracle's grid options (DfGenerateGrid's grid types) by what the
mathematics and the published rules fix:

* Gauss-Chebyshev radial rules (Becke, TA, KK mappings) integrate Gaussians
  to their closed form; KK ignores the atomic radius.
* Murray-Handy-Laming pruning and the SG-1 regions give the published orders
  at hand-picked radii; the order never falls with r and is always a
  supported Lebedev rule.
* Stratmann-Scuseria and Becke partitions sum to one over every atom, dummy
  atoms included; the own weight is one at a nucleus; a dummy atom has a
  Becke cell of its own but cuts no other cell.
* Whole grids (SG-1, pruned FINE, GC + Becke) integrate a known density to
  its electron count.
"""
import numpy as np
import pytest

from oracle import oracle
from proteindf_amd import basis, guess, molecules

LEB = [6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434, 590, 770, 974, 1202, 1454, 1730]


@pytest.mark.parametrize("mapping", [0, 1, 2], ids=["becke", "ta", "kk"])
def test_gc_radial_rules_integrate_gaussians(mapping):
    # sum_i w_i exp(-a r_i^2) = (pi/a)^(3/2): w_i carries 4 pi r^2.  Widths
    # are scaled by R (not for KK, whose mapping has no R) so every Gaussian
    # lies inside the mapped range
    for R in (0.25, 0.6, 1.2308):
        for aR2 in (0.05, 0.5, 5.0, 50.0):
            a = aR2 if mapping == 2 else aR2 / R ** 2
            s = 0.0
            for i in range(1, 100):
                r, w = oracle.radial_gc(R, 99, i, mapping)
                if r <= 30.0:
                    s += w * np.exp(-a * r * r)
            assert abs(s / (np.pi / a) ** 1.5 - 1.0) < 1e-10, (mapping, R, a, s)
    if mapping == 2:
        assert oracle.radial_gc(0.25, 50, 7, 2) == oracle.radial_gc(1.5, 50, 7, 2)


def test_gc_abscissae_are_increasing_and_positive():
    for mapping in (0, 1, 2):
        r = [oracle.radial_gc(0.7, 75, i, mapping)[0] for i in range(1, 76)]
        assert r[0] > 0
        # i runs from the outside in for the second-kind nodes
        assert np.all(np.diff(r) < 0) or np.all(np.diff(r) > 0)


def test_lebedev_floor():
    assert oracle.lebedev_floor(1) == 6
    for k, n in enumerate(LEB):
        assert oracle.lebedev_floor(n) == n
        if k + 1 < len(LEB):
            assert oracle.lebedev_floor(LEB[k + 1] - 1) == n
    assert oracle.lebedev_floor(10 ** 6) == 1730


def test_mhl_pruning_known_orders():
    # n = min(int(5 r / R_Bragg * n_ang), n_ang), then the largest rule <= n
    # H: R_Bragg = 0.25; O: 0.60
    assert oracle.pruned_order(1, 1, 0.01, 0.25, 302) == 50     # 60.4 -> 60 -> 50
    assert oracle.pruned_order(1, 1, 0.03, 0.25, 302) == 170    # 181.2 -> 170
    assert oracle.pruned_order(1, 8, 0.1, 0.6, 302) == 230      # 251.7 -> 230
    assert oracle.pruned_order(1, 8, 0.001, 0.6, 302) == 6      # 2.5 -> 6
    assert oracle.pruned_order(1, 8, 0.2, 0.6, 302) == 302      # saturated
    # no pruning: n_ang itself, floored to a supported rule
    assert oracle.pruned_order(0, 8, 1e-6, 0.6, 302) == 302
    assert oracle.pruned_order(0, 8, 1e-6, 0.6, 300) == 266
    for Z, RB in ((1, 0.25), (6, 0.70), (8, 0.60), (16, 1.00)):
        orders = [oracle.pruned_order(1, Z, r, RB, 590) for r in np.linspace(1e-4, 1.0, 400)]
        assert all(o in LEB for o in orders)
        assert np.all(np.diff(orders) >= 0)
        assert orders[-1] == 590


def test_sg1_regions():
    # Gill, Johnson, Pople (1993): regions of r / R with 6, 38, 86, 194, 86 points;
    # alpha = (0.25, 0.5, 1.0, 4.5) for H-He, (0.1667, 0.5, 0.9, 3.5) for Li-Ne,
    # (0.1, 0.4, 0.8, 2.5) for Na-Ar
    cases = {1: [(0.2, 6), (0.3, 38), (0.7, 86), (2.0, 194), (5.0, 86)],
             6: [(0.1, 6), (0.3, 38), (0.7, 86), (2.0, 194), (4.0, 86)],
             11: [(0.05, 6), (0.2, 38), (0.5, 86), (1.0, 194), (3.0, 86)]}
    for Z, rows in cases.items():
        R = oracle.sg1_radius(Z)
        for q, n in rows:
            assert oracle.pruned_order(2, Z, q * R, R, 194) == n, (Z, q)
    # the SG-1 radii: H 1.0000, C 1.2308, O 0.8791 (atomic units)
    assert (oracle.sg1_radius(1), oracle.sg1_radius(6), oracle.sg1_radius(8)) == (1.0, 1.2308, 0.8791)


def _water_with_dummy():
    mol = molecules.water_cluster(2, seed=0)
    Z = np.append(mol.Z, 0)
    xyz = np.vstack([mol.xyz, mol.xyz.mean(0) + [0.3, -0.2, 0.1]])
    return Z, xyz


def test_ss_partition_of_unity_and_nuclei():
    Z, xyz = _water_with_dummy()
    rng = np.random.default_rng(2)
    for r in xyz.mean(0) + rng.normal(0, 2.0, size=(60, 3)):
        s = sum(oracle.ss_weight(xyz, a, r) for a in range(len(Z)))
        assert abs(s - 1.0) < 1e-14
    for a in range(len(Z)):
        assert oracle.ss_weight(xyz, a, xyz[a]) == 1.0
    # with u_an < -0.64 for every other atom n the step is exactly 1
    near = xyz[0] + [0.01, 0.0, 0.0]
    assert oracle.ss_weight(xyz, 0, near) == 1.0


def test_becke_partition_with_a_dummy_atom():
    Z, xyz = _water_with_dummy()
    rng = np.random.default_rng(3)
    dummy = len(Z) - 1
    for r in xyz.mean(0) + rng.normal(0, 1.0, size=(40, 3)):
        for adj in (0, 1):
            ws = [oracle.becke_weight(Z, xyz, a, r, adj) for a in range(len(Z))]
            assert abs(sum(ws) - 1.0) < 1e-14
    # the dummy atom has a cell of its own (calcMultiCenterWeight_Becke sums
    # Ps_uij over every atom) but does not cut the real atoms' cells (Ps_uij
    # skips neighbours with Z = 0): near it the real atoms share what is left
    w_dummy = oracle.becke_weight(Z, xyz, dummy, xyz[dummy], 1)
    assert 0.0 < w_dummy < 1.0
    real = sum(oracle.becke_weight(Z, xyz, a, xyz[dummy], 1) for a in range(dummy))
    assert abs(real + w_dummy - 1.0) < 1e-14
    # at a real nucleus the dummy's cell is cut to zero
    assert oracle.becke_weight(Z, xyz, dummy, xyz[0], 1) == 0.0


@pytest.mark.parametrize("name,opts", [
    ("sg-1", dict(n_rad=50, n_ang=194, pruning=2, partitioning=1, size_adjust=0, radii=1)),
    ("fine", dict(n_rad=75, n_ang=302, pruning=1, partitioning=1, size_adjust=1)),
    ("user-gc-ta-becke", dict(n_rad=60, n_ang=194, radial=1, gc_mapping=1, pruning=1, partitioning=0)),
    ("user-gc-becke-noadj", dict(n_rad=60, n_ang=146, radial=1, gc_mapping=0, pruning=0, partitioning=0,
                                 size_adjust=0)),
])
def test_whole_grids_count_electrons(name, opts):
    mol = molecules.water_cluster(2, seed=1)
    info = basis.OrbitalInfo.build(mol, oracle.normalize_cgto)
    grid, atom = oracle.generate_grid_ex(mol.Z, mol.xyz, oracle.grid_opts(**opts))
    assert len(grid) > 0 and np.all(np.diff(atom) >= 0)
    P = guess.density_guess(mol, info, eps=0.0)  # diagonal: integral = trace
    _, _, rho = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), grid, 0.5 * P, want_rho=True)
    n_el = 2.0 * (grid[:, 3] * rho[:, 0]).sum()
    assert abs(n_el / np.trace(P) - 1.0) < 2e-4, (name, n_el)


def test_unpruned_options_reproduce_the_plain_generator():
    mol = molecules.water_cluster(1, seed=4)
    ref, atom_ref = oracle.generate_grid(mol.Z, mol.xyz, 40, 86)
    g, atom = oracle.generate_grid_ex(mol.Z, mol.xyz, oracle.grid_opts(n_rad=40, n_ang=86))
    np.testing.assert_array_equal(g, ref)
    np.testing.assert_array_equal(atom, atom_ref)


def test_pruned_shell_sizes():
    # one atom: the partition weight is 1, so every point of every kept shell stays
    mol = molecules.Molecule(["O"], np.zeros((1, 3)))
    for pruning, R in ((1, oracle.bragg_slater(8)), (2, oracle.sg1_radius(8))):
        o = oracle.grid_opts(n_rad=50, n_ang=194, pruning=pruning, radii=1 if pruning == 2 else 0,
                             partitioning=1, size_adjust=0)
        g, _ = oracle.generate_grid_ex(mol.Z, mol.xyz, o)
        expected = 0
        for i in range(1, 51):
            r, _ = oracle.radial_em(R, 50, i)
            if r <= 30.0:
                expected += oracle.pruned_order(pruning, 8, r, R, 194)
        assert len(g) == expected
        assert len(g) < 50 * 194
