"""The allocation-free entry point gs_nl_response_noalloc against the
allocating gs_nl_response: the same kernels on the same input, so the results
agree bit for bit, on two calls in a row (the second smaller, on the scratch
the first left behind), in every precision; sizes beyond the fixed scratch
are rejected."""
import numpy as np
import pytest

from conftest import write_cfg

pytestmark = pytest.mark.gpu

gsa = pytest.importorskip("gsolver_amd")

HOT = dict(ENERGY="20e-6", DIAMETER_FWHM="4e-5", N_T="1024", N_R="64", R_MAX="0.2e-3")
MAX_SAMPLES = 262144


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_noalloc_nl_response_matches_allocating_entry_point(tmp_path, prec):
    p = gsa.Plan(write_cfg(tmp_path / "c.txt", **HOT), precision=prec)
    f = p.get()
    rng = np.random.default_rng(5)
    for n_lines in (48, 7):
        rows = rng.choice(p.R, n_lines, replace=False)
        x = f[rows] * (1 + 0.01 * rng.standard_normal((n_lines, p.T)))
        ref = p.nl_response(x)
        got = p.nl_response_noalloc(x)
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref), (prec, n_lines)
    with pytest.raises(gsa.GsError, match="exceeds the fixed scratch") as ei:
        p.nl_response_noalloc(np.zeros((MAX_SAMPLES // p.T + 1, p.T), complex))
    assert ei.value.code == -1
    p.close()
