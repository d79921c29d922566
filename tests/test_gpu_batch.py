"""GPU checks of batched builds (pdfxc_batch_*): several molecules' blocks in
one graph-captured launch sequence give each molecule the result of its own
context -- E_xc bit for bit (same slots, same fixed-order sums), K to 1e-13
(the order of FP64 atomic adds differs) -- in FP64 and split precision, and
every replay of the captured graph gives the same result."""
import numpy as np
import pytest

from oracle import oracle
from proteindf_amd import calc_grid, engine, guess, molecules

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


def _molecules():
    out = []
    for n, seed in ((2, 11), (4, 12), (3, 13)):
        mol = molecules.water_cluster(n, seed)
        info = calc_grid.orbital_info(mol)
        P = 0.5 * guess.density_guess(mol, info, seed=seed, eps=0.05)
        out.append((mol, info, P))
    return out


@pytest.mark.parametrize("prec", ["fp64", "bf16x3", "bf16x3k", "bfp", "f16k"])
def test_batch_equals_individual_builds(prec):
    mols = _molecules()
    ctxs, single = [], []
    for k, (mol, info, P) in enumerate(mols):
        ctx = engine.XcContext(info, engine.default_opts(precision=engine.PRECISIONS[prec]))
        if k == 1:
            grid, _ = oracle.generate_grid(mol.Z, mol.xyz, 30, 86)
            ctx.set_grid(grid)
        else:
            ctx.generate_grid(mol.Z, mol.xyz, 40, 86)
        # each molecule its own launch shapes: the batch keeps every item's
        # workgroup shape (split K: 16 x 1, 12 x 2, 8 x 4 waves x patches)
        ctx.set_tuning(k_tiles=(96, 48, 24)[k], split_k_patches=(16, 24, 32)[k])
        E, K, _ = ctx.build_rks(guess.pack_lower(P))
        single.append((E, K))
        ctxs.append(ctx)
    say("batch", prec, [c.n_pts for c in ctxs], "points")
    batch = engine.XcBatch(ctxs)
    st = batch.stats()
    assert st.n_points == sum(c.n_pts for c in ctxs)
    Ps = [guess.pack_lower(P) for _, _, P in mols]
    first = None
    for rep in range(3):  # capture, then two replays
        E, Ks = batch.build_rks(Ps)
        say(f"  replay {rep}: E {E}")
        for i, (Ei, Ki) in enumerate(single):
            assert E[i] == Ei, (rep, i, E[i], Ei)
            np.testing.assert_allclose(Ks[i], Ki, rtol=0, atol=1e-13)
        if first is None:
            first = (E.copy(), [k.copy() for k in Ks])
        else:
            np.testing.assert_array_equal(E, first[0])
    batch.close()
    for c in ctxs:
        c.close()


def test_batch_functionals_and_hf():
    mols = _molecules()[:2]
    ctxs = []
    for mol, info, P in mols:
        ctx = engine.XcContext(info)
        ctx.generate_grid(mol.Z, mol.xyz, 30, 86)
        ctxs.append(ctx)
    batch = engine.XcBatch(ctxs)
    Ps = [guess.pack_lower(P) for _, _, P in mols]
    for fid in (engine.PDFXC_SVWN, engine.PDFXC_BLYP):
        E, Ks = batch.build_rks(Ps, functional=fid)
        for i, ctx in enumerate(ctxs):
            Ei, Ki, _ = ctx.build_rks(Ps[i], functional=fid)
            assert E[i] == Ei
            np.testing.assert_allclose(Ks[i], Ki, rtol=0, atol=1e-13)
    E, Ks = batch.build_rks(Ps, functional=engine.PDFXC_HF)
    assert not E.any() and not any(k.any() for k in Ks)
    batch.close()
    for c in ctxs:
        c.close()
