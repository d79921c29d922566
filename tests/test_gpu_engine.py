"""GPU parity of the HIP engine against the CPU oracle, through the C ABI.

Tolerances: the FP64 path must agree with the oracle's restatement of the
reference to |dE_xc| <= 1e-10 Eh and max|dK| <= 1e-10 (the north-star FP64
bar); pointwise quantities (AO values, functional, weights) to 1e-13
relative.  Every case prints its name and sizes before it launches work.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle
from proteindf_amd import basis, calc_grid, engine, guess, molecules

pytestmark = pytest.mark.gpu

E_TOL = 1e-10
K_TOL = 1e-10
KATS = json.loads((Path(__file__).parent / "golden" / "functional_kats.json").read_text())["cases"]


def say(*a):
    print("[gpu-case]", *a, flush=True)


def test_mfma_accumulator_layout():
    say("mfma layout selftest")
    D = engine.selftest_mfma()
    i = np.arange(16)[:, None]
    k = np.arange(4)[None, :]
    A = i + 100.0 * k                      # 16 x 4
    B = (np.arange(16)[None, :] - 7.0 * np.arange(4)[:, None])  # 4 x 16
    np.testing.assert_array_equal(D, A @ B)


def closed_shell_kats():
    out = []
    for c in KATS:
        if c["functional"] != "DfFunctional_B3LYP":
            continue
        i = c["inputs"]
        if i["rhoA"] == i["rhoB"] and i["gAA"] == i["gAB"] == i["gBB"]:
            out.append(c)
    return out


@pytest.mark.parametrize("case", closed_shell_kats(), ids=lambda c: c["test"])
def test_functional_kernel_reference_kats(case):
    say("functional KAT", case["test"])
    i = case["inputs"]
    o = engine.eval_functional([i["rhoA"]], [i["gAA"]])[0]
    got = {"zk": o[0], "vrhoA": o[1], "vrhoB": o[1], "vgAA": o[2], "vgBB": o[2], "vgAB": o[3]}
    for name in case["checked"]:
        assert abs(got[name] - case["expected"][name]) <= case["eps"], name


def test_functional_kernel_matches_oracle():
    rng = np.random.default_rng(11)
    rho = 10.0 ** rng.uniform(-15.5, 3.0, size=4000)
    gam = np.where(rng.uniform(size=4000) < 0.05, 0.0, 10.0 ** rng.uniform(-20, 4, size=4000))
    say("functional vs oracle", len(rho))
    o = engine.eval_functional(rho, gam)
    ref = oracle.functional(oracle.B3LYP, np.stack([rho, rho, gam, gam, gam], 1))
    for col, rcol in ((0, 0), (1, 1), (2, 3), (3, 4)):
        np.testing.assert_allclose(o[:, col], ref[:, rcol], rtol=2e-13, atol=1e-300)


def water(n, seed=0):
    mol = molecules.water_cluster(n, seed)
    info = calc_grid.orbital_info(mol)
    return mol, info


def test_ao_values_match_oracle():
    mol, info = water(2)
    ctx = engine.XcContext(info)
    rng = np.random.default_rng(3)
    pts = mol.xyz[rng.integers(0, mol.n_atoms, 600)] + rng.normal(0, 1.5, size=(600, 3))
    pts[:20] = mol.xyz.mean(0) + rng.normal(0, 15.0, size=(20, 3))  # far points hit the cutoff
    say("AO eval", info.n_aos, "AOs", len(pts), "points")
    phi, gx, gy, gz = ctx.eval_aos(pts)
    rphi, rgx, rgy, rgz = oracle.eval_aos(oracle.AoTable.from_basis(info), pts)
    for a, b in ((phi, rphi), (gx, rgx), (gy, rgy), (gz, rgz)):
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-15)
    ctx.close()


@pytest.mark.parametrize("n_rad,n_ang", [(40, 194), (75, 302)])
def test_grid_generation_matches_oracle(n_rad, n_ang):
    mol, info = water(2, seed=1)
    ctx = engine.XcContext(info)
    say("grid", n_rad, n_ang)
    n = ctx.generate_grid(mol.Z, mol.xyz, n_rad, n_ang)
    g = ctx.get_grid()
    ref, _ = oracle.generate_grid(mol.Z, mol.xyz, n_rad, n_ang)
    assert n == len(g)
    # Becke weights near the 1e-16 screen are differences 1 - f3(nu) at nu -> 1,
    # where FMA contraction changes the last bits: match points by position, and
    # a point kept by only one side must carry a negligible weight
    key = lambda p: tuple(np.round(p[:3], 9))
    ga = {key(p): p[3] for p in g}
    ra = {key(p): p[3] for p in ref}
    common = sorted(set(ga) & set(ra))
    assert len(common) > 0.999 * max(len(ga), len(ra))
    for k in set(ga) ^ set(ra):
        assert abs(ga.get(k, ra.get(k))) < 1e-13
    wg = np.array([ga[k] for k in common])
    wr = np.array([ra[k] for k in common])
    np.testing.assert_allclose(wg, wr, rtol=1e-12, atol=1e-13)
    # the kept points are in the reference's order: atom, radial shell, Lebedev point
    np.testing.assert_allclose(g[:50, :3], ref[:50, :3], rtol=0, atol=1e-13)
    ctx.close()


def _compare_build(mol, info, grid, PA, ao_screen=0.0, etol=E_TOL, ktol=K_TOL):
    n = info.n_aos
    ctx = engine.XcContext(info, engine.default_opts(ao_screen=ao_screen))
    ctx.set_grid(grid)
    st = ctx.stats()
    say("build", mol.name, n, "AOs", len(grid), "points", st.n_blocks, "blocks",
        "mean N_b", round(st.mean_nao, 1), "screen", ao_screen)
    E, K, rho = ctx.build_rks(guess.pack_lower(PA), want_rho=True)
    e_ref, K_ref, rho_ref = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), grid, PA, want_rho=True)
    Kd = guess.unpack_lower(K, n)
    dE = abs(E - e_ref)
    dK = np.abs(Kd - K_ref).max()
    say("  E", E, "ref", e_ref, "dE", dE, "max dK", dK)
    assert dE <= etol
    assert dK <= ktol
    np.testing.assert_allclose(rho, rho_ref, rtol=1e-9, atol=1e-13)
    ctx.close()
    return E, Kd


def test_build_matches_oracle_water_dimer():
    mol, info = water(2, seed=2)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 40, 194)
    PA = 0.5 * guess.density_guess(mol, info, seed=4, eps=0.05)
    _compare_build(mol, info, grid, PA)


def test_build_matches_oracle_h2o8_full_grid():
    # the (H2O)_8 / 6-31G(d) / 75 x 302 workload, whole grid
    mol, info = water(8, seed=0)
    assert info.n_aos == 144
    ctx = engine.XcContext(info)
    ctx.generate_grid(mol.Z, mol.xyz, 75, 302)
    grid = ctx.get_grid()
    ctx.close()
    PA = 0.5 * guess.density_guess(mol, info, seed=0, eps=0.05)
    _compare_build(mol, info, grid, PA)


def test_small_and_ragged_grids():
    mol, info = water(1, seed=5)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 40, 194)
    PA = 0.5 * guess.density_guess(mol, info, seed=1, eps=0.05)
    rng = np.random.default_rng(0)
    for npts in (1, 5, 127, 128, 129, 1000):
        sub = grid[np.sort(rng.choice(len(grid), npts, replace=False))]
        _compare_build(mol, info, sub, PA)


def test_zero_density_gives_zero():
    mol, info = water(1, seed=6)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 30, 86)
    ctx = engine.XcContext(info)
    ctx.set_grid(grid)
    n = info.n_aos
    say("zero density")
    E, K, _ = ctx.build_rks(np.zeros(n * (n + 1) // 2))
    assert E == 0.0 and np.all(K == 0.0)
    ctx.close()


def test_accumulates_into_caller_matrix():
    mol, info = water(1, seed=7)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 30, 86)
    ctx = engine.XcContext(info)
    ctx.set_grid(grid)
    PA = guess.pack_lower(0.5 * guess.density_guess(mol, info, eps=0.02))
    say("accumulate")
    E1, K1, _ = ctx.build_rks(PA)
    K2 = np.full_like(K1, 0.25)
    E2, K2, _ = ctx.build_rks(PA, K2)
    assert E1 == E2
    np.testing.assert_allclose(K2 - 0.25, K1, rtol=0, atol=1e-14)
    ctx.close()


def test_errors_are_reported_not_faulted():
    mol, info = water(1, seed=8)
    ctx = engine.XcContext(info)
    n = info.n_aos
    with pytest.raises(engine.PdfxcError) as ei:
        ctx.build_rks(np.zeros(n * (n + 1) // 2))
    assert ei.value.code == 3
    with pytest.raises(engine.PdfxcError) as ei:
        ctx.generate_grid(mol.Z, mol.xyz, 30, 100)
    assert ei.value.code == 4
    ctx.close()


def test_rejected_grid_leaves_no_grid():
    # a grid that set_grid refuses drops the previous one: the next build is
    # refused with PDFXC_ERR_NO_GRID before any kernel runs
    mol, info = water(1, seed=8)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 20, 50)
    ctx = engine.XcContext(info)
    n = info.n_aos
    PA = guess.pack_lower(0.5 * guess.density_guess(mol, info, seed=1, eps=0.05))
    ctx.set_grid(grid)
    E0 = ctx.build_rks(PA)[0]
    bad = grid.copy()
    bad[len(bad) // 2, 3] = np.nan
    with pytest.raises(engine.PdfxcError) as ei:
        ctx.set_grid(bad)
    assert ei.value.code == 1
    with pytest.raises(engine.PdfxcError) as ei:
        ctx.build_rks(np.zeros(n * (n + 1) // 2))
    assert ei.value.code == 3
    # a shape the blocks cannot take is refused and the grid stays usable
    ctx.set_grid(grid)
    with pytest.raises(engine.PdfxcError):
        ctx.set_tuning(split_k_patches=64)
    assert ctx.build_rks(PA)[0] == E0
    ctx.close()


def test_kxc_is_derivative_of_exc_on_gpu():
    mol, info = water(2, seed=9)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 40, 194)
    ctx = engine.XcContext(info)
    ctx.set_grid(grid)
    n = info.n_aos
    PA = 0.5 * guess.density_guess(mol, info, seed=3, eps=0.05)
    E0, K, _ = ctx.build_rks(guess.pack_lower(PA))
    K = guess.unpack_lower(K, n)
    rng = np.random.default_rng(12)
    say("FD check")
    for _ in range(4):
        D = rng.normal(size=(n, n))
        D = 0.5 * (D + D.T)
        h = 1e-5
        Ep, _, _ = ctx.build_rks(guess.pack_lower(PA + h * D))
        Em, _, _ = ctx.build_rks(guess.pack_lower(PA - h * D))
        fd = (Ep - Em) / (2 * h)
        ana = 2.0 * (D * K).sum()
        assert abs(fd - ana) < 1e-7 * max(1.0, abs(ana)), (fd, ana)
    ctx.close()


def test_screened_build_error_is_small():
    mol, info = water(8, seed=0)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 40, 194)
    PA = guess.pack_lower(0.5 * guess.density_guess(mol, info, seed=0, eps=0.05))
    out = {}
    for scr in (0.0, 1e-12):
        ctx = engine.XcContext(info, engine.default_opts(ao_screen=scr))
        ctx.set_grid(grid)
        say("screening", scr, "mean N_b", ctx.stats().mean_nao)
        out[scr] = ctx.build_rks(PA)[:2]
        ctx.close()
    assert abs(out[0.0][0] - out[1e-12][0]) < 1e-9
    assert np.abs(out[0.0][1] - out[1e-12][1]).max() < 1e-9


def test_reference_style_dfxcfunctional_electron_count():
    # DfXCFunctional::buildXcMatrix + checkGridAccuracy on the (H2O)_8 workload:
    # with a diagonal guess the integrated density is the electron count
    mol, _ = water(8, seed=0)
    cg = calc_grid.DfCalcGridX(mol, n_rad=75, n_ang=302)
    P = guess.density_guess(mol, cg.info, eps=0.0)
    say("DfXCFunctional (H2O)_8")
    xc = calc_grid.DfXCFunctional(cg, mol.n_electrons)
    F = xc.buildXcMatrix(guess.pack_lower(P))
    assert abs(xc.electrons_on_grid - mol.n_electrons) < 1e-4 * mol.n_electrons
    assert np.isfinite(xc.XC_energy_) and xc.XC_energy_ < 0
    assert np.all(np.isfinite(F))


def test_update_mode_accumulates_rho_from_density_differences():
    # xc-update = yes: rho[P2] = rho[P1] + rho[P2 - P1] (calcRho_GGA_part adds,
    # DfCalcGridX.cpp:1925-1961); K and E then follow from the accumulated rho
    mol, info = water(2, seed=9)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 40, 194)
    P1 = 0.5 * guess.density_guess(mol, info, seed=1, eps=0.05)
    P2 = 0.5 * guess.density_guess(mol, info, seed=2, eps=0.08)
    ctx = engine.XcContext(info)
    ctx.set_grid(grid)
    E0, K0, rho0 = ctx.build_rks_update(guess.pack_lower(P1), want_rho=True)  # stored rho starts at zero
    E1, K1, rho1 = ctx.build_rks(guess.pack_lower(P1), want_rho=True)
    assert abs(E0 - E1) <= 1e-12 * abs(E1)
    np.testing.assert_allclose(rho0, rho1, rtol=1e-14, atol=1e-18)
    Eu, Ku, rhou = ctx.build_rks_update(guess.pack_lower(P2 - P1), want_rho=True)
    E2, K2, rho2 = ctx.build_rks(guess.pack_lower(P2), want_rho=True)
    ctx.close()
    say("update mode: |dE|", abs(Eu - E2), "max|dK|", np.abs(Ku - K2).max())
    assert abs(Eu - E2) <= 1e-11
    np.testing.assert_allclose(Ku, K2, rtol=0, atol=1e-12)
    np.testing.assert_allclose(rhou, rho2, rtol=1e-12, atol=1e-16)
    e_ref, K_ref, _ = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), grid, P2)
    assert abs(Eu - e_ref) <= 1e-10
