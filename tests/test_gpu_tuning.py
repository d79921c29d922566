"""GPU checks of the launch-shape tuning: every shape gives the same build
(E to 1e-12, K to 1e-12: only summation orders change) in FP64 and split
precision, the autotuner returns one of its candidates with its measured
time and leaves the context on it, and out-of-range shapes are refused."""
import ctypes

import numpy as np
import pytest

from oracle import oracle
from proteindf_amd import calc_grid, engine, guess, molecules

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


def _case():
    mol = molecules.water_cluster(4, 21)
    info = calc_grid.orbital_info(mol)
    PA = 0.5 * guess.density_guess(mol, info, seed=21, eps=0.05)
    grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 40, 194)
    return mol, info, PA, grid


@pytest.mark.parametrize("prec", ["fp64", "bf16x3", "bf16x3k"])
def test_every_shape_gives_the_same_build(prec):
    mol, info, PA, grid = _case()
    ctx = engine.XcContext(info, engine.default_opts(precision=engine.PRECISIONS[prec]))
    ctx.set_grid(grid)
    P = guess.pack_lower(PA)
    E0, K0, _ = ctx.build_rks(P)
    shapes = [dict(rho_shift=1), dict(rho_shift=3, k_tiles=8), dict(k_tiles=24), dict(split_rho_shift=1),
              dict(split_rho_shift=2, split_k_patches=4), dict(split_k_patches=8),
              # 12 waves of two patches and 8 waves of four, full and partial items
              dict(split_k_patches=24), dict(split_k_patches=21), dict(split_k_patches=32),
              dict(split_k_patches=29),
              # split K work units of 1, 2 and 8 blocks (the default is 4)
              dict(split_k_group=1), dict(split_k_group=2), dict(split_k_group=8),
              dict(split_k_group=8, split_k_patches=16), dict(split_k_group=1, split_k_patches=32)]
    for sh in shapes:
        ctx.set_tuning(**sh)
        E, K, _ = ctx.build_rks(P)
        say(prec, sh, "dE", abs(E - E0), "max dK", np.abs(K - K0).max())
        assert abs(E - E0) <= 1e-12 * abs(E0)
        np.testing.assert_allclose(K, K0, rtol=0, atol=1e-12)
    for bad in (dict(k_tiles=200), dict(split_k_patches=33), dict(split_k_patches=3), dict(split_k_group=3),
                dict(split_k_group=16), dict(split_k_group=0), dict(persistent=4), dict(persistent=-1)):
        with pytest.raises(engine.PdfxcError) as ei:
            ctx.set_tuning(**bad)
        assert ei.value.code == engine.ERROR_CODES["ARGUMENT"]
    ctx.close()


def test_autotune_keeps_a_candidate_and_builds_correctly():
    mol, info, PA, grid = _case()
    ctx = engine.XcContext(info)
    ctx.set_grid(grid)
    n = info.n_aos
    Pd = np.ascontiguousarray(PA)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    ptr = []
    for nbytes in (Pd.nbytes, Pd.nbytes, 8):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
        ptr.append(p.value)
    assert hip.hipMemcpy(ptr[0], Pd.ctypes.data, Pd.nbytes, 1) == 0
    t = ctx.autotune(*ptr, reps=2)
    say("autotune chose", t.as_dict())
    assert t.ms > 0 and t.rho_shift in (0, 1) and t.k_tiles in (96, 48, 24)
    for p in ptr:
        hip.hipFree(ctypes.c_void_p(p))
    E, K, _ = ctx.build_rks(guess.pack_lower(PA))
    ctx.close()
    e_ref, K_ref, _ = oracle.calc_xc_rks(oracle.AoTable.from_basis(info), grid, PA)
    assert abs(E - e_ref) <= 1e-10
    assert np.abs(guess.unpack_lower(K, n) - K_ref).max() <= 1e-10


@pytest.mark.parametrize("prec", ["bf16x3k", "bf16x2"])
def test_split_k_shapes_on_wide_blocks(prec):
    # blocks of more than 160 AO columns: items of 16 patches split such a
    # block, items of 24 and 32 (12 waves of two patches, 8 waves of four)
    # take it whole or nearly; every shape sums the same slab products
    mol = molecules.polyglycine_helix(3, seed=5)
    info = calc_grid.orbital_info(mol)
    PA = 0.5 * guess.density_guess(mol, info, seed=5, eps=0.05)
    ctx = engine.XcContext(info, engine.default_opts(precision=engine.PRECISIONS[prec]))
    ctx.generate_grid(mol.Z, mol.xyz, 40, 146)
    st = ctx.stats()
    say(prec, "n_aos", info.n_aos, "max N_b", st.max_nao, "mean N_b", round(st.mean_nao, 1))
    assert st.max_nao > 160
    P = guess.pack_lower(PA)
    E0, K0, _ = ctx.build_rks(P)
    for kp, kg in ((24, 4), (21, 4), (32, 4), (29, 4), (16, 4), (24, 1), (16, 2), (32, 8), (24, 8)):
        ctx.set_tuning(split_k_patches=kp, split_k_group=kg)
        E, K, _ = ctx.build_rks(P)
        say(prec, "split_k_patches", kp, "split_k_group", kg, "dE", abs(E - E0), "max dK", np.abs(K - K0).max())
        assert abs(E - E0) <= 1e-12 * abs(E0)
        np.testing.assert_allclose(K, K0, rtol=0, atol=1e-12)
    ctx.close()
    ref = engine.XcContext(info, engine.default_opts())
    ref.generate_grid(mol.Z, mol.xyz, 40, 146)
    Ef, Kf, _ = ref.build_rks(P)
    ref.close()
    e_tol, k_tol = {"bf16x3k": (1e-10, 1e-7), "bf16x2": (2e-4, 2e-5)}[prec]
    say(prec, "vs FP64: dE", abs(E0 - Ef), "max dK", np.abs(K0 - Kf).max())
    assert abs(E0 - Ef) <= e_tol and np.abs(K0 - Kf).max() <= k_tol


@pytest.mark.parametrize("prec", ["fp64", "bf16x3", "bf16x3k"])
def test_persistent_and_per_item_launches_agree(prec):
    """(H2O)8 on a 75 x 302 grid has more rho and K work items than the
    device keeps workgroups resident, so persistent launches (workgroups take
    items from a device counter) engage: with the rho kernels, the K kernels
    or both persistent, the build equals the one-workgroup-per-item build, the
    energy bit for bit (every item's arithmetic is the same) and K up to the
    order of its atomic adds, on repeated builds too (every build resets the
    counters)."""
    mol = molecules.water_cluster(8, 0)
    info = calc_grid.orbital_info(mol)
    PA = 0.5 * guess.density_guess(mol, info, seed=0, eps=0.05)
    ctx = engine.XcContext(info, engine.default_opts(precision=engine.PRECISIONS[prec]))
    n = ctx.generate_grid(mol.Z, mol.xyz, 75, 302)
    say(prec, "(H2O)8", n, "points", ctx.stats().n_blocks, "blocks")
    P = guess.pack_lower(PA)
    ctx.set_tuning(persistent=0)
    E0, K0, _ = ctx.build_rks(P)
    for mask in (1, 2, 3):
        ctx.set_tuning(persistent=mask)
        for rep in range(2):
            E, K, _ = ctx.build_rks(P)
            say(prec, "persistent", mask, "build", rep, "dE", E - E0, "max dK", np.abs(K - K0).max())
            assert E == E0
            np.testing.assert_allclose(K, K0, rtol=0, atol=1e-13)
    ctx.close()
