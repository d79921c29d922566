"""pdfxc_eval_functional_noalloc (staging in the library's fixed device array,
no allocation) against pdfxc_eval_functional (allocates its staging per call):
two calls in a row, the second with fewer points, equal bit for bit (same
kernel on the same inputs), and sizes beyond the array refused."""
import numpy as np
import pytest

from proteindf_amd import engine

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    rho = 10.0 ** rng.uniform(-12, 2, n)
    gamma = rho ** (8.0 / 3.0) * 10.0 ** rng.uniform(-3, 2, n)
    return rho, gamma


@pytest.mark.parametrize("functional", [engine.PDFXC_B3LYP, engine.PDFXC_BLYP, engine.PDFXC_SVWN])
def test_noalloc_equals_allocating_entry_point(functional):
    print(f"[gpu-case] eval_functional_noalloc functional {functional}", flush=True)
    for n, seed in ((engine.NOALLOC_MAX_POINTS, 1), (1000, 2), (1, 3)):
        rho, gamma = _inputs(n, seed)
        ref = engine.eval_functional(rho, gamma, functional)
        got = engine.eval_functional_noalloc(rho, gamma, functional)
        assert got.shape == (n, 4)
        np.testing.assert_array_equal(got, ref)
        assert np.isfinite(got).all()


def test_noalloc_refuses_sizes_beyond_its_array():
    n = engine.NOALLOC_MAX_POINTS + 1
    rho, gamma = _inputs(n, 4)
    with pytest.raises(engine.PdfxcError) as e:
        engine.eval_functional_noalloc(rho, gamma)
    assert e.value.code == engine.ERROR_CODES["ARGUMENT"]
    assert "65536" in str(e.value)
    # still usable afterwards
    rho, gamma = _inputs(10, 5)
    np.testing.assert_array_equal(engine.eval_functional_noalloc(rho, gamma), engine.eval_functional(rho, gamma))
