import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU-only check, excluded from the default CPU suite")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture
def sio2_dir(tmp_path):
    """A working directory holding the linear-regime SiO2 config and its .rawdisp,
    with the 512 MB 'full' dump redirected to a second maxI file."""
    import shutil
    for f in ("fil_SiO2_1d.txt", "SiO2.rawdisp"):
        shutil.copyfile(os.path.join(GOLDEN, f), tmp_path / f)
    return tmp_path


SIO2_NO_FULL = ["DUMPTYPE4", "maxI", "DUMP_FILENAME4", "fil_SiO2_%s_mI2.udf"]


def write_cfg(path, **kw):
    """Write a reference-format config (whitespace NAME VALUE pairs) with the
    fused-silica SiO2.rawdisp next to it; keyword arguments override keys."""
    import shutil
    base = dict(MAX_TOLERANCE="1e-3", LAMBDA_V="0.8e-6", N_T="256", T_MIN="-100e-15", T_MAX="100e-15",
                DISPERSION_FILE="SiO2.rawdisp", N_R="64", R_MAX="0.3e-3", N_Z="3", Z_MIN="0", Z_MAX="2e-4",
                ZNET_TYPE="e", PULSE_SHAPE="gg", DURATION_FWHM="30e-15", DIAMETER_FWHM="8e-5",
                WAVEFRONT_RADIUS="1e30", ENERGY="1e-6", NONLIN_REFRINDEX="2.5e-20", RAMAN_FRACTION="0.18",
                TAU_RAMAN="32e-15", OMEGA_RAMAN="8.2938e13", NEUTRAL_DENSITY="2.1e28", RECOMBINATION_TAU="150e-15",
                AMBIENT_CARRIERS="0", COLLISION_TAU="23e-15", IONIZATION_POTENTIAL="9", DUMPS_N="0")
    base.update({k: str(v) for k, v in kw.items()})
    d = os.path.dirname(str(path))
    if not os.path.exists(os.path.join(d, "SiO2.rawdisp")):
        shutil.copyfile(os.path.join(GOLDEN, "SiO2.rawdisp"), os.path.join(d, "SiO2.rawdisp"))
    with open(str(path), "w") as f:
        for k, v in base.items():
            f.write("%s %s\n" % (k, v))
    return str(path)


def rel_l2(a, b):
    import numpy as np
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))
