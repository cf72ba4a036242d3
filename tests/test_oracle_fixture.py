"""The oracle against the reference's own shipped data: omega.bin and
wavenum.bin (process_input.cpp:118-119) and the linear-regime dumps
bin/fil_SiO2__{f,fa,mI,pm,s}.udf of the 128-plane march of
bin/fil_SiO2_1d.txt.  CPU only."""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, SIO2_NO_FULL


def test_omega_wavenum_bins_exact(sio2_dir):
    sim = oracle.Sim(str(sio2_dir / "fil_SiO2_1d.txt"), SIO2_NO_FULL, workdir=str(sio2_dir), nthreads=4)
    om = np.fromfile(sio2_dir / "omega.bin")
    wk = np.fromfile(sio2_dir / "wavenum.bin", dtype=np.complex128)
    np.testing.assert_array_equal(om, np.fromfile(os.path.join(GOLDEN, "omega.bin")))
    ref = np.fromfile(os.path.join(GOLDEN, "wavenum.bin"), dtype=np.complex128)
    assert np.abs(wk - ref).max() <= 1e-15 * np.abs(ref).max()
    assert sim.R == 512 and sim.T == 512 and sim.NZ == 128


def _compare_slabs(outdir, nslab, tol):
    ref = np.load(os.path.join(GOLDEN, "linear_sio2_udf.npz"))
    slabs = list(ref["slabs"])
    worst = {}
    for suf in ("f", "fa", "mI", "pm", "s"):
        hdr, data = oracle.load_udf(os.path.join(outdir, "fil_SiO2__%s.udf" % suf))
        np.testing.assert_array_equal(hdr, ref["hdr_" + suf])
        for k, z in enumerate(slabs):
            if z >= nslab:
                break
            a, b = data[z], ref["udf_" + suf][k]
            nb = np.linalg.norm(b)
            err = np.linalg.norm(a - b) / nb if nb > 0 else np.abs(a).max()
            worst[suf] = max(worst.get(suf, 0), err)
            assert err < tol, (suf, z, err)
    return worst


def test_linear_march_first_planes(sio2_dir):
    """First 9 planes (8 accepted steps) of the 128-plane linear march."""
    sim = oracle.Sim(str(sio2_dir / "fil_SiO2_1d.txt"), SIO2_NO_FULL, nthreads=8)
    rc = sim.run(str(sio2_dir), max_steps=8)
    assert rc == 1
    worst = _compare_slabs(str(sio2_dir), 9, 1e-10)
    assert worst["f"] < 1e-12 and worst["fa"] < 1e-12


@pytest.mark.slow
def test_linear_march_all_planes(sio2_dir):
    sim = oracle.Sim(str(sio2_dir / "fil_SiO2_1d.txt"), SIO2_NO_FULL, nthreads=8)
    assert sim.run(str(sio2_dir)) == 0
    _compare_slabs(str(sio2_dir), 128, 1e-9)
