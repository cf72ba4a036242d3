"""The GPU march against the reference's own shipped linear-regime dumps
(bin/fil_SiO2_1d.txt -> fil_SiO2__{f,fa,mI,pm,s}.udf, 128 planes)."""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, SIO2_NO_FULL, rel_l2

pytestmark = pytest.mark.gpu
gsa = pytest.importorskip("gsolver_amd")


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 2e-5), ("mixed", 2e-5)])
def test_linear_fixture_all_planes(sio2_dir, prec, tol):
    p = gsa.Plan(str(sio2_dir / "fil_SiO2_1d.txt"), SIO2_NO_FULL, precision=prec, workdir=str(sio2_dir))
    np.testing.assert_array_equal(np.fromfile(sio2_dir / "omega.bin"), np.fromfile(os.path.join(GOLDEN, "omega.bin")))
    assert p.march(str(sio2_dir)) == 0
    ref = np.load(os.path.join(GOLDEN, "linear_sio2_udf.npz"))
    for suf in ("f", "fa", "mI", "pm", "s"):
        hdr, data = oracle.load_udf(str(sio2_dir / ("fil_SiO2__%s.udf" % suf)))
        np.testing.assert_array_equal(hdr, ref["hdr_" + suf])
        worst = 0.0
        for k, z in enumerate(ref["slabs"]):
            b = ref["udf_" + suf][k]
            e = rel_l2(data[z], b) if np.linalg.norm(b) > 0 else float(np.abs(data[z]).max())
            worst = max(worst, e)
        print(prec, suf, worst)
        assert worst < (tol if suf != "s" else max(tol, 1e-9)), (suf, worst)
