"""Build libgsolver_amd.so and the gsolver_amd executable in-tree with hipcc
for gfx950.  Usage: python -m gsolver_amd.build [-j N]"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]
SOURCES = ["gs_line.hip", "gs_adjoint.hip", "gs_qdht.hip", "gs_host.cpp", "gs_special.cpp", "gs_plan.cpp"]
LIB = os.path.join(HERE, "libgsolver_amd.so")
EXE = os.path.join(HERE, "gsolver_amd")


def _compile(src):
    out = os.path.join(OBJ, src + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    return out


def build(jobs=4, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs, check=True)
    os.replace(tmp, LIB)
    main_o = _compile("gs_main.cpp")
    subprocess.run([HIPCC, "--offload-arch=" + ARCH, "-o", EXE, main_o, "-L" + HERE, "-lgsolver_amd",
                    "-Wl,-rpath,$ORIGIN"], check=True)
    if verbose:
        print("built", LIB, EXE)
    return LIB


if __name__ == "__main__":
    j = int(sys.argv[sys.argv.index("-j") + 1]) if "-j" in sys.argv else 4
    build(j, verbose=True)
