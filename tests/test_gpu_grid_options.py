"""GPU parity of pdfxc_generate_grid_ex (every grid type of DfGenerateGrid)
against the oracle's restatement (pinned in test_oracle_grid_options.py),
builds on those grids, dummy atoms, the gradient's partition rules, and a
build driven by the reference's own pdfparam.mpac."""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle
from proteindf_amd import calc_grid, engine, guess, molecules, pdfparam

pytestmark = pytest.mark.gpu

OPTION_SETS = {
    "sg-1": dict(n_rad=50, n_ang=194, pruning=2, partitioning=1, size_adjust=0, radii=1),
    "fine": dict(n_rad=75, n_ang=302, pruning=1, partitioning=1, size_adjust=1),
    "coarse-becke": dict(n_rad=35, n_ang=110, pruning=1, partitioning=0, size_adjust=1),
    "user-gc-ta": dict(n_rad=60, n_ang=266, radial=1, gc_mapping=1, pruning=1, partitioning=0),
    "user-gc-becke-noadj": dict(n_rad=50, n_ang=146, radial=1, gc_mapping=0, pruning=0, partitioning=0,
                                size_adjust=0),
    "user-gc-kk-ss": dict(n_rad=50, n_ang=350, radial=1, gc_mapping=2, pruning=1, partitioning=1),
}


def say(*a):
    print("[gpu-case]", *a, flush=True)


def _engine_opts(d):
    o = oracle.grid_opts(**d)
    return engine.GridOpts(**{f: getattr(o, f) for f, _ in engine.GridOpts._fields_})


def _compare_grids(g, ref):
    # match points by position: weights at the 1e-16 screen may differ in the last bits
    key = lambda p: tuple(np.round(p[:3], 9))
    ga = {key(p): p[3] for p in g}
    ra = {key(p): p[3] for p in ref}
    common = sorted(set(ga) & set(ra))
    assert len(common) > 0.999 * max(len(ga), len(ra))
    for k in set(ga) ^ set(ra):
        assert abs(ga.get(k, ra.get(k))) < 1e-13
    np.testing.assert_allclose([ga[k] for k in common], [ra[k] for k in common], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("name", sorted(OPTION_SETS))
def test_grid_options_match_oracle(name):
    mol = molecules.water_cluster(2, seed=1)
    info = calc_grid.orbital_info(mol)
    opts = _engine_opts(OPTION_SETS[name])
    say("grid", name, opts.as_dict())
    ctx = engine.XcContext(info)
    n = ctx.generate_grid_ex(mol.Z, mol.xyz, opts)
    g = ctx.get_grid()
    ctx.close()
    ref, _ = oracle.generate_grid_ex(mol.Z, mol.xyz, oracle.grid_opts(**OPTION_SETS[name]))
    say("  points", n, "oracle", len(ref))
    assert n == len(g)
    _compare_grids(g, ref)


def test_dummy_atom_grid_matches_oracle():
    mol = molecules.water_cluster(2, seed=0)
    Z = np.append(mol.Z, 0)
    xyz = np.vstack([mol.xyz, mol.xyz.mean(0) + [0.3, -0.2, 0.1]])
    info = calc_grid.orbital_info(mol)
    for name in ("sg-1", "coarse-becke"):
        say("dummy atom grid", name)
        ctx = engine.XcContext(info)
        ctx.generate_grid_ex(Z, xyz, _engine_opts(OPTION_SETS[name]))
        g = ctx.get_grid()
        ctx.close()
        ref, _ = oracle.generate_grid_ex(Z, xyz, oracle.grid_opts(**OPTION_SETS[name]))
        _compare_grids(g, ref)


@pytest.mark.parametrize("name", ["sg-1", "user-gc-ta"])
def test_build_on_pruned_grid_matches_oracle(name):
    mol = molecules.water_cluster(3, seed=2)
    info = calc_grid.orbital_info(mol)
    PA = 0.5 * guess.density_guess(mol, info, seed=2, eps=0.05)
    grid, _ = oracle.generate_grid_ex(mol.Z, mol.xyz, oracle.grid_opts(**OPTION_SETS[name]))
    cg = calc_grid.DfCalcGridX(mol, grid=_engine_opts(OPTION_SETS[name]), info=info)
    n = info.n_aos
    F = np.zeros(n * (n + 1) // 2)
    E = cg.calcXCIntegForFockAndEnergy(guess.pack_lower(PA), calc_grid.DfFunctional_B3LYP(), F)
    cg.ctx.close()
    e_ref, K_ref, _ = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), grid, PA)
    dK = np.abs(guess.unpack_lower(F, n) - K_ref).max()
    say(name, len(grid), "points: dE", abs(E - e_ref), "max dK", dK)
    assert abs(E - e_ref) <= 1e-10 and dK <= 1e-10


@pytest.mark.parametrize("name,partition", [("sg-1", 2), ("fine", 2), ("user-gc-becke-noadj", 1), ("coarse-becke", 0)])
def test_gradient_on_option_grids_matches_oracle(name, partition):
    mol = molecules.water_cluster(2, seed=6)
    info = calc_grid.orbital_info(mol)
    PA = 0.5 * guess.density_guess(mol, info, seed=6, eps=0.05)
    ctx = engine.XcContext(info)
    ctx.generate_grid_ex(mol.Z, mol.xyz, _engine_opts(OPTION_SETS[name]))
    say("gradient on the", name, "grid")
    E, gf, gw = ctx.gradient_rks(mol.Z, mol.xyz, guess.pack_lower(PA))
    ctx.close()
    grid, atom = oracle.generate_grid_ex(mol.Z, mol.xyz, oracle.grid_opts(**OPTION_SETS[name]))
    e_ref, gf_ref, gw_ref = oracle.xc_gradient_rks(oracle.AoTable.from_basis(info), info.ao_atom, grid, atom,
                                                   mol.Z, mol.xyz, PA, partition=partition)
    say(f"  |dE| {abs(E - e_ref):.2e} max|dgf| {np.abs(gf - gf_ref).max():.2e} max|dgw| {np.abs(gw - gw_ref).max():.2e}"
        f" max|gw| {np.abs(gw_ref).max():.3f}")
    assert abs(E - e_ref) <= 1e-9
    np.testing.assert_allclose(gf, gf_ref, rtol=0, atol=1e-9)
    np.testing.assert_allclose(gw, gw_ref, rtol=0, atol=1e-9)
    assert np.abs(gw_ref).max() > 1e-3  # the weight term is there


def test_gradient_refuses_dummy_atoms():
    mol = molecules.water_cluster(2, seed=6)
    info = calc_grid.orbital_info(mol)
    Z = np.append(mol.Z, 0)
    xyz = np.vstack([mol.xyz, mol.xyz.mean(0) + [0.3, -0.2, 0.1]])
    ctx = engine.XcContext(info)
    ctx.generate_grid_ex(Z, xyz, _engine_opts(OPTION_SETS["sg-1"]))
    with pytest.raises(engine.PdfxcError) as ei:
        ctx.gradient_rks(Z, xyz, guess.pack_lower(0.5 * guess.density_guess(mol, info, seed=6, eps=0.05)))
    assert ei.value.code == engine.ERROR_CODES["UNSUPPORTED"]
    ctx.close()


@pytest.mark.parametrize("name", ["user-gc-becke-noadj", "sg-1"])
def test_gradient_matches_central_differences_on_option_grids(name):
    mol = molecules.water_cluster(2, seed=7)
    info = calc_grid.orbital_info(mol)
    PA = 0.5 * guess.density_guess(mol, info, seed=7, eps=0.05)
    o = OPTION_SETS[name]
    O = oracle.o_matrix(mol.Z, mol.xyz)
    ctx = engine.XcContext(info)
    ctx.generate_grid_ex(mol.Z, mol.xyz, _engine_opts(o))
    E, gf, gw = ctx.gradient_rks(mol.Z, mol.xyz, guess.pack_lower(PA))
    ctx.close()
    g = gf + gw
    h = 1e-4

    def energy(xyz):
        m = molecules.Molecule(mol.symbols, xyz, mol.name)
        gr, _ = oracle.generate_grid_ex(m.Z, m.xyz, oracle.grid_opts(**o), O=O)
        c = engine.XcContext(calc_grid.orbital_info(m))
        c.set_grid(gr)
        e, _, _ = c.build_rks(guess.pack_lower(PA))
        c.close()
        return e

    rng = np.random.default_rng(4)
    worst = 0.0
    for A in rng.choice(mol.n_atoms, 2, replace=False):
        for d in range(3):
            xp, xm = mol.xyz.copy(), mol.xyz.copy()
            xp[A, d] += h
            xm[A, d] -= h
            fd = (energy(xp) - energy(xm)) / (2 * h)
            worst = max(worst, abs(g[A, d] - fd))
            assert abs(g[A, d] - fd) <= 5e-8 + 1e-6 * abs(fd), (A, d, g[A, d], fd)
    say(f"gradient ({name} grid) vs central differences: max |diff| {worst:.2e}")


def test_build_from_the_reference_pdfparam():
    p = pdfparam.load(Path(__file__).parent / "golden" / "H2O.mpac")
    say("pdfparam H2O.mpac:", p.grid)
    cg = calc_grid.DfCalcGridX.from_pdfparam(p)
    info = p.orbital_info(oracle.normalize_cgto)
    assert cg.m_nNumOfAOs == info.n_aos == 24
    mol = p.molecule
    grid, _ = oracle.generate_grid_ex(mol.Z, mol.xyz, oracle.grid_opts(**p.grid.to_opts().as_dict()))
    assert len(grid) == len(cg.grid)
    PA = 0.5 * guess.density_guess(mol, info, seed=1, eps=0.05)
    n = info.n_aos
    # the file's functional is HF: no grid term
    xc = calc_grid.DfXCFunctional(cg, p.electrons(), p.xc_functional, method=p.method)
    assert not xc.buildXcMatrix(guess.pack_lower(2 * PA)).any() and xc.XC_energy_ == 0.0
    for fn in ("SVWN", "B3LYP"):
        xc = calc_grid.DfXCFunctional(cg, p.electrons(), fn, method=p.method)
        F = xc.buildXcMatrix(guess.pack_lower(2 * PA))
        e_ref, K_ref, _ = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), grid, PA,
                                             fid=oracle.FUNCTIONAL_IDS[calc_grid.FUNCTIONAL_FOR[
                                                 engine.XC_TYPES[fn]].__name__])
        dK = np.abs(guess.unpack_lower(F, n) - K_ref).max()
        say(fn, "dE", abs(xc.XC_energy_ - e_ref), "max dK", dK)
        assert abs(xc.XC_energy_ - e_ref) <= 1e-10 and dK <= 1e-10
    cg.ctx.close()


def test_option_errors_and_order_floor():
    mol = molecules.water_cluster(1, seed=3)
    info = calc_grid.orbital_info(mol)
    ctx = engine.XcContext(info)
    say("grid option errors")
    for bad in (dict(radial=5), dict(gc_mapping=7), dict(pruning=3), dict(partitioning=2), dict(radii=4),
                dict(n_rad=0)):
        with pytest.raises(engine.PdfxcError) as ei:
            ctx.generate_grid_ex(mol.Z, mol.xyz, _engine_opts({**OPTION_SETS["coarse-becke"], **bad}))
        assert ei.value.code == engine.ERROR_CODES["ARGUMENT"], bad
    with pytest.raises(engine.PdfxcError) as ei:
        ctx.generate_grid_ex(np.array([8, 1, 1, 19]), np.vstack([mol.xyz, [5.0, 5.0, 5.0]]),
                             _engine_opts(OPTION_SETS["coarse-becke"]))
    assert ei.value.code == engine.ERROR_CODES["UNSUPPORTED"]  # no radius past Ar
    # an unsupported n_ang is floored to the largest rule below it, as the reference does
    o = dict(n_rad=30, n_ang=300, pruning=0, partitioning=0)
    n = ctx.generate_grid_ex(mol.Z, mol.xyz, _engine_opts(o))
    ref, _ = oracle.generate_grid_ex(mol.Z, mol.xyz, oracle.grid_opts(**{**o, "n_ang": 266}))
    ctx.close()
    assert n == len(ref)
