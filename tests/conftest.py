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
