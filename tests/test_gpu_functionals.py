"""GPU parity of the functional family (HF, HFS, SVWN, HFB, BLYP, B3LYP).

The device functionals are checked pointwise against the oracle (whose
components the reference's own known-answer tests pin, test_oracle_functionals.py)
and against those known answers directly where a closed-shell case isolates a
component (Slater through HFS, Becke88 through HFB).  Whole builds and the
nuclear gradient of every grid functional are checked against the oracle's
restatement of calcXCIntegForFockAndEnergy / energyGradient to the FP64 bar
(|dE_xc| <= 1e-10 Eh, max|dK| <= 1e-10).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle
from proteindf_amd import calc_grid, engine, guess, molecules

pytestmark = pytest.mark.gpu

KATS = json.loads((Path(__file__).parent / "golden" / "functional_kats.json").read_text())["cases"]

# engine functional -> the oracle's id for the same functional
ORACLE_ID = {engine.PDFXC_HFS: oracle.HFS, engine.PDFXC_SVWN: oracle.SVWN,
             engine.PDFXC_HFB: oracle.BECKE88, engine.PDFXC_BLYP: oracle.BLYP,
             engine.PDFXC_B3LYP: oracle.B3LYP}
GRID_FUNCTIONALS = sorted(ORACLE_ID)
NAMES = {v: k for k, v in engine.XC_TYPES.items() if k != "B88LYP"}


def say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.mark.parametrize("fid", GRID_FUNCTIONALS, ids=lambda f: NAMES[f])
def test_functional_kernel_matches_oracle(fid):
    rng = np.random.default_rng(100 + fid)
    n = 4000
    rho = 10.0 ** rng.uniform(-15.5, 3.0, size=n)
    gam = np.where(rng.uniform(size=n) < 0.05, 0.0, 10.0 ** rng.uniform(-20, 4, size=n))
    gam[:8] = [0.0, 1e-17, 1e-16, 2e-16, 1e-30, 0.0, 5e-17, 1e-15]  # around B88's gamma guard
    say("functional vs oracle", NAMES[fid], n)
    o = engine.eval_functional(rho, gam, fid)
    ref = oracle.functional(ORACLE_ID[fid], np.stack([rho, rho, gam, gam, gam], 1))
    # The device's log/atan/pow differ from the host's in the last bit, and
    # the reference formulas amplify that where they cancel: VWN's eps_c(x)
    # at low density (t1 + t2 ~ 1/x^2, x = rs^1/2), and B88 + LYP sums near
    # zero (e ~ 1e-6 from terms ~ 1e-2).  With the host's libm the same
    # source agrees with the oracle to the last bit for HFS, SVWN and HFB.
    # So: 2e-13 relative on 97% of the points, 1e-11 relative on all (the
    # largest seen is 4e-12; a wrong constant or term shows as >= 1e-8).
    bad = []
    for col, rcol in ((0, 0), (1, 1), (2, 3), (3, 4)):
        d = np.abs(o[:, col] - ref[:, rcol])
        r = d / np.maximum(np.abs(ref[:, rcol]), 1e-300)
        r[d == 0] = 0.0
        say(f"  col {col}: max|d| {d.max():.2e} max rel {r.max():.2e} beyond 2e-13: {(r > 2e-13).sum()}")
        if (r > 1e-11).any() or (r > 2e-13).mean() > 0.03:
            i = int(np.argmax(r))
            bad.append((col, rho[i], gam[i], o[i, col], ref[i, rcol]))
    assert not bad, bad
    if fid in (engine.PDFXC_HFS, engine.PDFXC_SVWN):
        assert not o[:, 2:].any()  # LDA: no gamma dependence


def _closed_shell(functional):
    out = []
    for c in KATS:
        i = c["inputs"]
        if c["functional"] == functional and i["rhoA"] == i["rhoB"] and i["gAA"] == i["gBB"]:
            out.append(c)
    return out


# Slater has no gamma inputs; B88 does not depend on gAB (its derivative is 0)
KAT_CASES = ([(engine.PDFXC_HFS, c) for c in _closed_shell("DfFunctional_Slater")] +
             [(engine.PDFXC_HFB, c) for c in _closed_shell("DfFunctional_Becke88")])


@pytest.mark.parametrize("fid,case", KAT_CASES, ids=[f"{NAMES[f]}.{c['test']}" for f, c in KAT_CASES])
def test_functional_kernel_reference_kats(fid, case):
    say("functional KAT", NAMES[fid], case["test"])
    i = case["inputs"]
    o = engine.eval_functional([i["rhoA"]], [i["gAA"]], fid)[0]
    got = {"zk": o[0], "vrhoA": o[1], "vrhoB": o[1], "vgAA": o[2], "vgBB": o[2], "vgAB": o[3]}
    for name in case["checked"]:
        assert abs(got[name] - case["expected"][name]) <= case["eps"], name


def test_kat_cases_present():
    assert len(_closed_shell("DfFunctional_Slater")) >= 1
    assert len(_closed_shell("DfFunctional_Becke88")) >= 1


def _water_dimer(seed):
    mol = molecules.water_cluster(2, seed)
    info = calc_grid.orbital_info(mol)
    PA = 0.5 * guess.density_guess(mol, info, seed=seed, eps=0.05)
    return mol, info, PA


@pytest.mark.parametrize("fid", GRID_FUNCTIONALS, ids=lambda f: NAMES[f])
def test_build_matches_oracle(fid):
    mol, info, PA = _water_dimer(3)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 40, 194)
    n = info.n_aos
    ctx = engine.XcContext(info)
    ctx.set_grid(grid)
    say("build", NAMES[fid], mol.name, n, "AOs", len(grid), "points")
    E, K, rho = ctx.build_rks(guess.pack_lower(PA), want_rho=True, functional=fid)
    ctx.close()
    e_ref, K_ref, rho_ref = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), grid, PA,
                                               fid=ORACLE_ID[fid], want_rho=True)
    Kd = guess.unpack_lower(K, n)
    say("  E", E, "ref", e_ref, "dE", abs(E - e_ref), "max dK", np.abs(Kd - K_ref).max())
    assert abs(E - e_ref) <= 1e-10
    assert np.abs(Kd - K_ref).max() <= 1e-10
    np.testing.assert_allclose(rho, rho_ref, rtol=1e-9, atol=1e-13)


@pytest.mark.parametrize("fid", [engine.PDFXC_SVWN, engine.PDFXC_BLYP], ids=lambda f: NAMES[f])
def test_split_build_within_bounds(fid):
    # the split-bf16 (3 slices) build of the other functionals keeps the B3LYP error bounds
    mol, info, PA = _water_dimer(4)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 40, 194)
    out = {}
    for prec in ("fp64", "bf16x3"):
        ctx = engine.XcContext(info, engine.default_opts(precision=engine.PRECISIONS[prec]))
        ctx.set_grid(grid)
        say("split build", NAMES[fid], prec)
        E, K, _ = ctx.build_rks(guess.pack_lower(PA), functional=fid)
        ctx.close()
        out[prec] = (E, K)
    dE = abs(out["bf16x3"][0] - out["fp64"][0])
    dK = np.abs(out["bf16x3"][1] - out["fp64"][1]).max()
    say("  dE", dE, "max dK", dK)
    assert dE <= 5e-6 and dK <= 5e-7


def test_hf_has_no_grid_term():
    mol, info, PA = _water_dimer(5)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 30, 86)
    ctx = engine.XcContext(info)
    ctx.set_grid(grid)
    say("HF build")
    K0 = np.arange(info.n_aos * (info.n_aos + 1) // 2, dtype=np.float64)
    E, K, _ = ctx.build_rks(guess.pack_lower(PA), K0.copy(), functional=engine.PDFXC_HF)
    assert E == 0.0
    np.testing.assert_array_equal(K, K0)  # accumulated into: untouched
    E, gf, gw = ctx.gradient_rks(mol.Z, mol.xyz, guess.pack_lower(PA),
                                 point_atom=np.zeros(len(grid), np.int32), functional=engine.PDFXC_HF)
    assert E == 0.0 and not gf.any() and not gw.any()
    ctx.close()
    assert not engine.eval_functional([0.3], [0.1], engine.PDFXC_HF).any()


@pytest.mark.parametrize("fid", [engine.PDFXC_HFS, engine.PDFXC_SVWN, engine.PDFXC_HFB, engine.PDFXC_BLYP],
                         ids=lambda f: NAMES[f])
def test_gradient_matches_oracle(fid):
    mol = molecules.water_cluster(2, 5)
    info = calc_grid.orbital_info(mol)
    PA = 0.5 * guess.density_guess(mol, info, seed=5, eps=0.05)
    O = oracle.o_matrix(mol.Z, mol.xyz)
    grid, atom = oracle.generate_grid(mol.Z, mol.xyz, 40, 86, O=O)
    ctx = engine.XcContext(info)
    ctx.set_grid(grid)
    say("gradient", NAMES[fid], len(grid), "points")
    E, gf, gw = ctx.gradient_rks(mol.Z, mol.xyz, guess.pack_lower(PA), point_atom=atom, functional=fid)
    ctx.close()
    e_ref, gf_ref, gw_ref = oracle.xc_gradient_rks(oracle.AoTable.from_basis(info), info.ao_atom, grid, atom,
                                                   mol.Z, mol.xyz, PA, fid=ORACLE_ID[fid])
    say("  |dE|", abs(E - e_ref), "max|dgf|", np.abs(gf - gf_ref).max(), "max|dgw|", np.abs(gw - gw_ref).max())
    assert abs(E - e_ref) <= 1e-10
    np.testing.assert_allclose(gf, gf_ref, rtol=0, atol=1e-10)
    np.testing.assert_allclose(gw, gw_ref, rtol=0, atol=1e-10)


def test_xc_functional_by_name():
    mol, info, PA = _water_dimer(6)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 30, 86)
    cg = calc_grid.DfCalcGridX(mol, grid=grid, info=info)
    say("DfXCFunctional by name")
    P = guess.pack_lower(2.0 * PA)
    got = {}
    for name in ("SVWN", "BLYP", "B88LYP", "HF", "B3LYP~"):
        xc = calc_grid.DfXCFunctional(cg, 20, name)
        F = xc.buildXcMatrix(P)
        got[name] = (xc.XC_energy_, F)
    assert got["HF"][0] == 0.0 and not got["HF"][1].any()
    assert got["BLYP"][0] == got["B88LYP"][0]
    e_ref = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), grid, PA, fid=oracle.SVWN)[0]
    assert abs(got["SVWN"][0] - e_ref) <= 1e-10
    cg.ctx.close()
