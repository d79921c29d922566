"""Pins the oracle's XC nuclear gradient (energyGradient, DfCalcGridX.cpp:
1215-1415): AO second derivatives against finite differences of the AO
gradients, the Becke partition derivative against finite differences of the
weight, and the whole gradient against central finite differences of E_xc on
a grid that moves rigidly with its atoms (fixed orientation matrix)."""
import numpy as np
import pytest

from oracle import oracle
from proteindf_amd import calc_grid, guess, molecules


def water(n, seed):
    mol = molecules.water_cluster(n, seed)
    return mol, calc_grid.orbital_info(mol)


def test_ao_hessians_match_finite_differences_of_gradients():
    mol, info = water(1, 3)
    aos = oracle.AoTable.from_basis(info)
    rng = np.random.default_rng(0)
    pts = mol.xyz[rng.integers(0, 3, 40)] + rng.normal(0, 0.8, (40, 3))
    h = oracle.eval_ao_hessians(aos, pts)
    eps = 1e-5
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for q, (a, b) in enumerate(pairs):
        d = np.zeros(3)
        d[b] = eps
        gp = oracle.eval_aos(aos, pts + d)[1 + a]
        gm = oracle.eval_aos(aos, pts - d)[1 + a]
        fd = (gp - gm) / (2 * eps)
        np.testing.assert_allclose(h[:, :, q], fd, rtol=1e-6, atol=1e-7 * np.abs(fd).max())


@pytest.mark.parametrize("partition", [0, 1, 2], ids=["becke", "becke-noadj", "ss"])
def test_becke_weight_gradient_matches_finite_differences(partition):
    # the weight itself comes from the grid code (becke_weight / ss_weight),
    # its derivative from the gradient code: central differences tie them
    mol, _ = water(2, 4)
    Z, xyz = mol.Z, mol.xyz
    rng = np.random.default_rng(1)
    eps = 1e-6

    def weight(x, A, r):
        if partition == 2:
            return oracle.ss_weight(x, A, r)
        return oracle.becke_weight(Z, x, A, r, partition == 0)

    nonzero = 0
    for trial in range(12):
        A = int(rng.integers(0, len(Z)))
        r = xyz[A] + rng.normal(0, 1.2, 3)
        om, g = oracle.becke_weight_gradient(Z, xyz, A, r, partition)
        assert abs(om - weight(xyz, A, r)) < 1e-15
        assert np.all(g[A] == 0.0)
        nonzero += np.any(g != 0.0)
        for B in range(len(Z)):
            if B == A:
                continue
            for d in range(3):
                xp, xm = xyz.copy(), xyz.copy()
                xp[B, d] += eps
                xm[B, d] -= eps
                fd = (weight(xp, A, r) - weight(xm, A, r)) / (2 * eps)
                assert abs(g[B, d] - fd) <= 1e-7 * max(1.0, abs(fd)), (trial, B, d, g[B, d], fd)
    assert nonzero >= 6  # points in the switching regions, not only inside a cell


def _energy(mol, PA, O, n_rad, n_ang):
    info = calc_grid.orbital_info(mol)
    grid, atom = oracle.generate_grid(mol.Z, mol.xyz, n_rad, n_ang, O=O)
    e, _, _ = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), grid, PA)
    return e


@pytest.fixture(scope="module")
def dimer_gradient():
    mol, info = water(2, 5)
    PA = 0.5 * guess.density_guess(mol, info, seed=3, eps=0.05)
    O = oracle.o_matrix(mol.Z, mol.xyz)
    grid, atom = oracle.generate_grid(mol.Z, mol.xyz, 40, 86, O=O)
    e, gf, gw = oracle.xc_gradient_rks(oracle.AoTable.from_basis(info), info.ao_atom, grid, atom,
                                       mol.Z, mol.xyz, PA)
    return mol, info, PA, O, e, gf, gw


def test_gradient_energy_equals_build_energy(dimer_gradient):
    mol, info, PA, O, e, gf, gw = dimer_gradient
    assert abs(e - _energy(mol, PA, O, 40, 86)) < 1e-12


def test_gradient_terms_are_translation_invariant(dimer_gradient):
    _, _, _, _, _, gf, gw = dimer_gradient
    np.testing.assert_allclose(gf.sum(0), 0.0, atol=1e-12)
    np.testing.assert_allclose(gw.sum(0), 0.0, atol=1e-12)


def test_gradient_matches_central_differences(dimer_gradient):
    mol, info, PA, O, e, gf, gw = dimer_gradient
    g = gf + gw
    h = 1e-4
    rng = np.random.default_rng(7)
    for A in rng.choice(mol.n_atoms, 4, replace=False):
        for d in range(3):
            mp = molecules.Molecule(mol.symbols, mol.xyz.copy(), mol.name)
            mm = molecules.Molecule(mol.symbols, mol.xyz.copy(), mol.name)
            mp.xyz[A, d] += h
            mm.xyz[A, d] -= h
            fd = (_energy(mp, PA, O, 40, 86) - _energy(mm, PA, O, 40, 86)) / (2 * h)
            # measured agreement ~1e-8 (the FD truncation error at h = 1e-4)
            assert abs(g[A, d] - fd) <= 5e-8 + 1e-6 * abs(fd), (A, d, g[A, d], fd, gf[A, d], gw[A, d])
