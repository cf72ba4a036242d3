"""Write the reference-held fixtures used by the parity tests.

Reads data files shipped in the reference's bin/ directory (configs, the
SiO2 .rawdisp table, omega.bin / wavenum.bin and the linear-regime dumps
fil_SiO2__{f,fa,mI,pm,s}.udf) and stores them, or a subsample of their
z-slabs, under tests/golden/.  Run once in a checkout that has the
reference mounted; the outputs are committed so tests never need it:

    python tests/golden/make_fixtures.py /root/reference/bin
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SLABS = sorted(set(list(range(0, 17)) + list(range(24, 128, 8)) + [127]))


def load_udf(path):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[7:31], dtype=np.int32).copy()
    fs, cplx, tN, xN, yN, zN = (int(v) for v in hdr)
    dt = np.complex128 if cplx else np.float64
    data = np.frombuffer(raw[31:], dtype=dt).reshape(zN, yN * xN * tN)
    return raw[:7], hdr, data


def main(src):
    for name in ("fil_SiO2_1d.txt", "SiO2.rawdisp", "omega.bin", "wavenum.bin"):
        dst = os.path.join(HERE, name)
        shutil.copyfile(os.path.join(src, name), dst)
        os.chmod(dst, 0o644)
    out = {"slabs": np.array(SLABS)}
    for suf in ("f", "fa", "mI", "pm", "s"):
        tag, hdr, data = load_udf(os.path.join(src, "fil_SiO2__%s.udf" % suf))
        assert tag == b"UDF1.0\0"
        out["hdr_" + suf] = hdr
        out["udf_" + suf] = data[SLABS]
    np.savez_compressed(os.path.join(HERE, "linear_sio2_udf.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/bin")
