"""The mixed linear step's unbiased FP64 -> FP32 rounding and its hash
(round_unbiased / mix_bits, gsolver_amd/csrc/gs_dev.h), compiled for the host
with hipcc and checked by tests/host/round_unbiased_main.cpp: neighbours only,
IEEE specials and exact values unchanged, zero mean error at several
positions between neighbours, uniform hash bits.  CPU only."""
import os
import subprocess

import pytest

from conftest import ROOT

HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_round_unbiased_on_host(tmp_path):
    exe = str(tmp_path / "rut")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-host-only",
                        "-I" + os.path.join(ROOT, "gsolver_amd", "csrc"), "-o", exe,
                        os.path.join(ROOT, "tests", "host", "round_unbiased_main.cpp")],
                       capture_output=True, text=True, timeout=600, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:])
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-3000:]
