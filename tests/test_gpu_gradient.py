"""GPU checks of the XC nuclear gradient (pdfxc_gradient_rks) against the
oracle's energyGradient restatement (which finite differences pin, see
test_oracle_gradient.py) and against central differences of the GPU E_xc."""
import numpy as np
import pytest

from oracle import oracle
from proteindf_amd import calc_grid, engine, guess, molecules

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _case(n, seed, n_rad, n_ang):
    mol = molecules.water_cluster(n, seed)
    info = calc_grid.orbital_info(mol)
    PA = 0.5 * guess.density_guess(mol, info, seed=seed, eps=0.05)
    O = oracle.o_matrix(mol.Z, mol.xyz)
    grid, atom = oracle.generate_grid(mol.Z, mol.xyz, n_rad, n_ang, O=O)
    return mol, info, PA, O, grid, atom


@pytest.mark.parametrize("n,seed,n_rad,n_ang", [(2, 5, 40, 86), (4, 2, 50, 194)])
def test_gradient_matches_oracle(n, seed, n_rad, n_ang):
    mol, info, PA, O, grid, atom = _case(n, seed, n_rad, n_ang)
    ctx = engine.XcContext(info)
    ctx.set_grid(grid)
    E, gf, gw = ctx.gradient_rks(mol.Z, mol.xyz, guess.pack_lower(PA), point_atom=atom)
    ctx.close()
    e_ref, gf_ref, gw_ref = oracle.xc_gradient_rks(oracle.AoTable.from_basis(info), info.ao_atom, grid, atom,
                                                   mol.Z, mol.xyz, PA)
    say(f"(H2O){n} {len(grid)} pts: |dE| {abs(E - e_ref):.2e} max|dgf| {np.abs(gf - gf_ref).max():.2e} "
        f"max|dgw| {np.abs(gw - gw_ref).max():.2e} max|gf| {np.abs(gf_ref).max():.3f} max|gw| {np.abs(gw_ref).max():.3f}")
    assert abs(E - e_ref) <= 1e-10
    np.testing.assert_allclose(gf, gf_ref, rtol=0, atol=1e-10)
    np.testing.assert_allclose(gw, gw_ref, rtol=0, atol=1e-10)
    np.testing.assert_allclose(gf.sum(0), 0.0, atol=1e-11)
    np.testing.assert_allclose(gw.sum(0), 0.0, atol=1e-11)


def test_gradient_on_generated_grid_uses_its_atom_assignment():
    mol, info, PA, O, grid, atom = _case(2, 6, 40, 86)
    ctx = engine.XcContext(info)
    ctx.generate_grid(mol.Z, mol.xyz, 40, 86)
    E, gf, gw = ctx.gradient_rks(mol.Z, mol.xyz, guess.pack_lower(PA))
    ctx.set_grid(ctx.get_grid())  # a user grid has no atom assignment
    with pytest.raises(engine.PdfxcError) as ei:
        ctx.gradient_rks(mol.Z, mol.xyz, guess.pack_lower(PA))
    assert ei.value.code == 1
    ctx.close()
    e_ref, gf_ref, gw_ref = oracle.xc_gradient_rks(oracle.AoTable.from_basis(info), info.ao_atom, grid, atom,
                                                   mol.Z, mol.xyz, PA)
    assert abs(E - e_ref) <= 1e-9
    np.testing.assert_allclose(gf + gw, gf_ref + gw_ref, rtol=0, atol=1e-9)


def test_gradient_matches_central_differences_of_gpu_energy():
    mol, info, PA, O, grid, atom = _case(2, 7, 40, 86)
    ctx = engine.XcContext(info)
    ctx.set_grid(grid)
    E, gf, gw = ctx.gradient_rks(mol.Z, mol.xyz, guess.pack_lower(PA), point_atom=atom)
    ctx.close()
    g = gf + gw
    h = 1e-4

    def energy(xyz):
        m = molecules.Molecule(mol.symbols, xyz, mol.name)
        inf = calc_grid.orbital_info(m)
        gr, _ = oracle.generate_grid(m.Z, m.xyz, 40, 86, O=O)
        c = engine.XcContext(inf)
        c.set_grid(gr)
        e, _, _ = c.build_rks(guess.pack_lower(PA))
        c.close()
        return e

    rng = np.random.default_rng(3)
    worst = 0.0
    for A in rng.choice(mol.n_atoms, 3, replace=False):
        for d in range(3):
            xp, xm = mol.xyz.copy(), mol.xyz.copy()
            xp[A, d] += h
            xm[A, d] -= h
            fd = (energy(xp) - energy(xm)) / (2 * h)
            worst = max(worst, abs(g[A, d] - fd))
            assert abs(g[A, d] - fd) <= 5e-8 + 1e-6 * abs(fd), (A, d, g[A, d], fd)
    say(f"gradient vs central differences: max |diff| {worst:.2e}")


# split precision: the functional term's contractions X = P phi and Z = P W
# on bf16 MFMA (hi+mid+lo, or hi+lo); against the oracle's FP64 gradient,
# bounds (|dE|, max|dg|) per mode at about 4x the measured (H2O)4 50 x 194
# errors (hi+mid+lo: |dg| 5-6e-9; hi+lo: |dE| 4e-5, |dg| 1.1e-6; the FP64-rho
# mode's energy is the FP64 one)
SPLIT_GRAD_BOUNDS = {engine.PDFXC_SPLIT_K_BF16X3: (0.0, 2e-8), engine.PDFXC_SPLIT_BF16X3: (2e-6, 2e-8),
                     engine.PDFXC_SPLIT_BF16: (2e-4, 5e-6)}


@pytest.mark.parametrize("prec", sorted(SPLIT_GRAD_BOUNDS))
def test_split_gradient_matches_oracle(prec):
    mol, info, PA, O, grid, atom = _case(4, 2, 50, 194)
    ctx = engine.XcContext(info, engine.default_opts(precision=prec))
    ctx.set_grid(grid)
    E, gf, gw = ctx.gradient_rks(mol.Z, mol.xyz, guess.pack_lower(PA), point_atom=atom)
    ctx.close()
    e_ref, gf_ref, gw_ref = oracle.xc_gradient_rks(oracle.AoTable.from_basis(info), info.ao_atom, grid, atom,
                                                   mol.Z, mol.xyz, PA)
    e_tol, g_tol = SPLIT_GRAD_BOUNDS[prec]
    say(f"[gpu-case] split gradient precision {prec}: |dE| {abs(E - e_ref):.2e} "
        f"max|dgf| {np.abs(gf - gf_ref).max():.2e} max|dgw| {np.abs(gw - gw_ref).max():.2e}")
    assert abs(E - e_ref) <= e_tol + 1e-10
    np.testing.assert_allclose(gf, gf_ref, rtol=0, atol=g_tol)
    np.testing.assert_allclose(gw, gw_ref, rtol=0, atol=g_tol)
    np.testing.assert_allclose(gf.sum(0), 0.0, atol=10 * g_tol)
