"""The adjoint at (Gly)20 scale: K_xc = dE_xc/dP_alpha (RKS: dE/deps along a
symmetric direction D is 2 <D, K>) against central differences of the GPU
energy along 32 random symmetric directions (entries N(0, 1), step 1e-5), in
FP64 and in split precision (there with a step of 1e-3: the split energy
carries ~1e-6 Eh of rounding noise, ~5e-4 of the derivative at that step;
tolerance 5e-3 relative with a floor of 1).  Workload: (Gly)20 alpha-helix (143 atoms, 1258 AOs, seed 4
jitter), 6-31G(d), B3LYP, 75 x 302 grid generated on the device."""
import numpy as np
import pytest

from proteindf_amd import calc_grid, engine, guess, molecules

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def gly20():
    mol = molecules.polyglycine_helix(20, seed=4)
    info = calc_grid.orbital_info(mol)
    g = engine.XcContext(info, engine.default_opts(ao_screen=1e-12))
    g.generate_grid(mol.Z, mol.xyz, 75, 302)
    grid = g.get_grid()
    g.close()
    PA = 0.5 * guess.density_guess(mol, info, seed=0, eps=0.05)
    return mol, info, grid, PA


@pytest.mark.parametrize("prec,h,tol", [("fp64", 1e-5, 1e-6), ("bf16x3", 1e-3, 5e-3)])
def test_kxc_is_the_derivative_of_exc(gly20, prec, h, tol):
    mol, info, grid, PA = gly20
    n = info.n_aos
    ctx = engine.XcContext(info, engine.default_opts(ao_screen=1e-12, precision=engine.PRECISIONS[prec]))
    ctx.set_grid(grid)
    say(prec, n, "AOs", len(grid), "points")
    E0, K, _ = ctx.build_rks(guess.pack_lower(PA))
    Kd = guess.unpack_lower(K, n)
    rng = np.random.default_rng(17)
    worst = 0.0
    for k in range(32):
        D = rng.normal(size=(n, n))
        D = 0.5 * (D + D.T)
        ep = ctx.build_rks(guess.pack_lower(PA + h * D))[0]
        em = ctx.build_rks(guess.pack_lower(PA - h * D))[0]
        fd = (ep - em) / (2 * h)
        ana = 2.0 * float((D * Kd).sum())
        rel = abs(fd - ana) / max(abs(ana), 1.0)
        worst = max(worst, rel)
        assert rel <= tol, (k, fd, ana)
    ctx.close()
    say(f"  {prec}: 32 directions, worst relative |fd - 2<D,K>| {worst:.2e}")
