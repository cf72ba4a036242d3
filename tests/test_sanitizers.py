"""Race and memory-error checks on the CPU side (no reference counterpart:
the reference has none): the oracle under AddressSanitizer + UBSan and under
ThreadSanitizer (its line- and column-parallel loops run on two threads), and
the library's host set-up code (gs_host.cpp, gs_special.cpp: config parser,
tables, QDHT plan, initial field) under AddressSanitizer + UBSan.  Each is
compiled with gcc / g++ into the test's temporary directory and run on a small
ionizing configuration; any sanitizer report fails the test."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT, write_cfg

ORACLE = os.path.join(ROOT, "oracle")
CSRC = os.path.join(ROOT, "gsolver_amd", "csrc")
SMALL = dict(ENERGY="26e-6", DIAMETER_FWHM="4e-5", N_T="256", N_R="16", R_MAX="0.2e-3")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="needs gcc / g++")


def _run(cmd, cwd, env=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, env=env, timeout=600)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_oracle_under_sanitizer(tmp_path, san):
    srcs = [os.path.join(ORACLE, f) for f in ("o_fft.c", "o_special.c", "o_setup.c", "o_march.c", "o_adjoint.c",
                                              "o_sancheck.c")]
    exe = str(tmp_path / "sancheck")
    rc, out = _run(["gcc", "-O1", "-g", "-std=gnu11", "-fno-omit-frame-pointer", "-fsanitize=" + san,
                    "-fno-sanitize-recover=all", "-DORACLE_REAL=double", "-I" + ORACLE, "-o", exe] + srcs
                   + ["-lm", "-lpthread"], str(tmp_path))
    assert rc == 0, out
    cfg = write_cfg(tmp_path / "c.txt", **SMALL)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", TSAN_OPTIONS="halt_on_error=1")
    rc, out = _run([exe, cfg], str(tmp_path), env)
    print(out[-2000:])
    assert rc == 0, out[-4000:]
    assert "ok" in out and "Sanitizer" not in out and "runtime error" not in out


def test_host_setup_under_sanitizer(tmp_path):
    exe = str(tmp_path / "hostcheck")
    srcs = [os.path.join(ROOT, "tests", "sanitize", "host_setup_main.cpp"), os.path.join(CSRC, "gs_host.cpp"),
            os.path.join(CSRC, "gs_special.cpp")]
    rc, out = _run(["g++", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-I" + CSRC, "-o", exe] + srcs + ["-lm"], str(tmp_path))
    assert rc == 0, out
    cfg = write_cfg(tmp_path / "c.txt", **dict(SMALL, DUMPS_N="2", DUMPTYPE0="flux", DUMP_FILENAME0="f_%s.udf",
                                               DUMPTYPE1="maxI", DUMP_FILENAME1="m_%s.udf"))
    rc, out = _run([exe, cfg, "N_Z", "3"], str(tmp_path), dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    print(out[-2000:])
    assert rc == 0, out[-4000:]
    assert "missing key:" in out and "Sanitizer" not in out and "runtime error" not in out
