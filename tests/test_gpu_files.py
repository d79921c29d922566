"""GPU checks of the file-level path: grid.mat written and read back by the
engine (points, atoms, rho columns), an update-mode SCF continued from the
file in a new context, and DfXCFunctional's fl_Work round (~Ppq / dP in,
Fxc out, grid.mat rewritten) against the oracle to the FP64 bar."""
import numpy as np
import pytest

from oracle import oracle
from proteindf_amd import calc_grid, engine, guess, molecules

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


def _case(seed):
    mol = molecules.water_cluster(2, seed)
    info = calc_grid.orbital_info(mol)
    P1 = guess.density_guess(mol, info, seed=seed, eps=0.05)
    P2 = guess.density_guess(mol, info, seed=seed + 1, eps=0.05)
    return mol, info, P1, P2


def test_grid_matrix_round_trip_and_update_continuation(tmp_path):
    mol, info, P1, P2 = _case(8)
    n = info.n_aos
    ctx1 = engine.XcContext(info)
    npts = ctx1.generate_grid(mol.Z, mol.xyz, 40, 86)
    say("grid.mat round trip", npts, "points")
    E1, K1, rho1 = ctx1.build_rks(guess.pack_lower(0.5 * P1), want_rho=True)
    path = tmp_path / "grid.mat"
    ctx1.save_grid_matrix(path, 4)
    g = engine.load_general_matrix(path)
    assert g.shape == (npts, 9)
    np.testing.assert_array_equal(g[:, :4], ctx1.get_grid())
    # the atom column is the generating atom: match points with the oracle's grid by position
    # (points at the 1e-16 weight screen may be kept by one side only)
    ref, atom_ref = oracle.generate_grid(mol.Z, mol.xyz, 40, 86)
    key = lambda p: tuple(np.round(p[:3], 9))
    ref_atom = {key(p): a for p, a in zip(ref, atom_ref)}
    matched = [(a, ref_atom[key(p)]) for p, a in zip(g, g[:, 4]) if key(p) in ref_atom]
    assert len(matched) > 0.999 * npts
    assert all(a == b for a, b in matched)
    np.testing.assert_array_equal(g[:, 5:], rho1)

    # a new context continues the update-mode SCF from the file
    ctx2 = engine.XcContext(info)
    assert ctx2.load_grid_matrix(path) == 9
    dP = guess.pack_lower(0.5 * (P2 - P1))
    Ea, Ka, rhoa = ctx1.build_rks_update(dP, want_rho=True)
    Eb, Kb, rhob = ctx2.build_rks_update(dP, want_rho=True)
    say("  update from file: dE", abs(Ea - Eb), "max dK", np.abs(Ka - Kb).max())
    np.testing.assert_array_equal(rhoa, rhob)  # the stored rho came back bit for bit
    assert Ea == Eb                            # fixed-order energy sum
    np.testing.assert_allclose(Ka, Kb, rtol=0, atol=1e-13)  # K sums by atomics
    e_ref, K_ref, _ = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), ctx1.get_grid(), 0.5 * P2)
    assert abs(Eb - e_ref) <= 1e-10
    assert np.abs(guess.unpack_lower(Kb, n) - K_ref).max() <= 1e-10

    # the atoms of the file drive the gradient's point motion
    Pg = guess.pack_lower(0.5 * P1)
    Eg1, gf1, gw1 = ctx1.gradient_rks(mol.Z, mol.xyz, Pg)
    Eg2, gf2, gw2 = ctx2.gradient_rks(mol.Z, mol.xyz, Pg)
    assert Eg1 == Eg2
    np.testing.assert_allclose(gf1, gf2, rtol=0, atol=1e-13)
    np.testing.assert_allclose(gw1, gw2, rtol=0, atol=1e-13)
    ctx1.close()
    ctx2.close()


@pytest.mark.parametrize("xc", ["B3LYP", "SVWN"])
def test_xc_step_through_work_files(tmp_path, xc):
    mol, info, P1, P2 = _case(9)
    n = info.n_aos
    work = calc_grid.WorkFiles(tmp_path)
    engine.save_symmetric_matrix(work.diff_density(1), guess.pack_lower(P1))
    engine.save_symmetric_matrix(work.diff_density(2), guess.pack_lower(P2 - P1))
    engine.save_symmetric_matrix(work.p_in(2), guess.pack_lower(P2))
    say("work files", xc)
    # iteration 1 generates the grid; iteration 2 reads grid.mat (points, atoms, rho) in a new object
    cg1 = calc_grid.DfCalcGridX(mol, n_rad=40, n_ang=86, info=info)
    calc_grid.DfXCFunctional(cg1, 20, xc).buildXcMatrixFromFiles(work, 1, update_xc=True)
    grid = cg1.grid
    cg1.ctx.close()
    ncols = engine.matrix_file_info(work.grid_matrix())[2]
    assert ncols == (6 if xc == "SVWN" else 9)
    cg2 = calc_grid.DfCalcGridX(mol, grid=work.grid_matrix(), info=info)
    xcf = calc_grid.DfXCFunctional(cg2, 20, xc)
    xcf.buildXcMatrixFromFiles(work, 2, update_xc=True)
    F_update = engine.load_symmetric_matrix(work.fxc(2))
    E_update = xcf.XC_energy_
    # the same iteration without update mode reads ~Ppq.rks2.mat
    xcf.buildXcMatrixFromFiles(work, 2, update_xc=False)
    F_full = engine.load_symmetric_matrix(work.fxc(2))
    cg2.ctx.close()
    fid = {"B3LYP": oracle.B3LYP, "SVWN": oracle.SVWN}[xc]
    e_ref, K_ref, rho_ref = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), grid, 0.5 * P2, fid=fid,
                                               want_rho=True)
    say("  E update", E_update, "full", xcf.XC_energy_, "ref", e_ref)
    assert abs(E_update - e_ref) <= 1e-10 and abs(xcf.XC_energy_ - e_ref) <= 1e-10
    assert np.abs(guess.unpack_lower(F_update, n) - K_ref).max() <= 1e-10
    assert np.abs(guess.unpack_lower(F_full, n) - K_ref).max() <= 1e-10
    assert abs(xcf.electrons_on_grid - 2.0 * (grid[:, 3] * rho_ref[:, 0]).sum()) <= 1e-9


def test_density_field_matches_oracle():
    # TlDensField / gridDensity: rho of P at arbitrary points, the caller's grid untouched
    mol, info, P1, P2 = _case(10)
    cg = calc_grid.DfCalcGridX(mol, n_rad=30, n_ang=86, info=info)
    grid0 = cg.grid
    rng = np.random.default_rng(4)
    pts = mol.xyz[rng.integers(0, mol.n_atoms, 3000)] + rng.normal(0, 1.2, size=(3000, 3))
    pts[:30] = mol.xyz.mean(0) + rng.normal(0, 12.0, size=(30, 3))
    say("density field", len(pts), "points")
    rho = calc_grid.TlDensField(cg).makeDensFld(guess.pack_lower(P1), pts)
    rho2, grad = cg.ctx.density_at_points(guess.pack_lower(P1), pts, want_grad=True)
    np.testing.assert_array_equal(rho, rho2)
    ref = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), np.c_[pts, np.ones(len(pts))], P1, want_rho=True)[2]
    np.testing.assert_allclose(rho, ref[:, 0], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(grad, ref[:, 1:], rtol=1e-11, atol=1e-14)
    np.testing.assert_array_equal(cg.grid, grid0)
    cg.ctx.close()
