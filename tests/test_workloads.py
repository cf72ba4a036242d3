"""The benchmark workloads' inputs (gsolver_amd/workloads.py): the fused-silica
Sellmeier .rawdisp writer and the configs built on it.  CPU only: the
library's host set-up (gs_setup_table) reads the files without a GPU."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

gsa = pytest.importorskip("gsolver_amd")
from gsolver_amd import workloads as W   # noqa: E402

C_LIGHT = 299792458.0


def read_rawdisp(path):
    """type-3 .rawdisp (process_input.cpp:687-693, m/refractive_index_raw_write.m:5-16):
    int32 3, int32 N, double omega[N], double n[2N] as (re, im)"""
    b = open(path, "rb").read()
    kind, n = np.frombuffer(b[:8], np.int32)
    om = np.frombuffer(b[8:8 + 8 * n], np.float64)
    ri = np.frombuffer(b[8 + 8 * n:8 + 24 * n], np.float64).reshape(n, 2)
    assert len(b) == 8 + 24 * n
    return int(kind), om, ri[:, 0] + 1j * ri[:, 1]


def sellmeier_sio2(lam_um):
    """Malitson's three-term fit for fused silica, JOSA 55, 1205 (1965)"""
    l2 = np.asarray(lam_um) ** 2
    terms = [(0.6961663, 0.0684043), (0.4079426, 0.1162414), (0.8974794, 9.896161)]
    return np.sqrt(1 + sum(b * l2 / (l2 - c * c) for b, c in terms))


def test_sellmeier_index_matches_published_values():
    """n of fused silica at the helium d line (587.56 nm), 800 nm and 1550 nm:
    1.45846, 1.45332, 1.44402 (Malitson 1965)."""
    got = W.sellmeier_n([0.5875618, 0.8, 1.55])
    np.testing.assert_allclose(got, [1.45846, 1.45332, 1.44402], atol=2e-5)
    lam = np.linspace(0.25, 4.0, 777)
    assert np.abs(W.sellmeier_n(lam) - sellmeier_sio2(lam)).max() < 1e-15


def test_rawdisp_file_layout(tmp_path):
    p = W.write_sellmeier_rawdisp(str(tmp_path / "s.rawdisp"))
    kind, om, n = read_rawdisp(p)
    assert kind == 3 and len(om) == 20001
    assert np.all(np.diff(om) > 0)
    lam = 2 * np.pi * C_LIGHT / om * 1e6
    assert abs(lam[0] - 4.0) < 1e-12 and abs(lam[-1] - 0.25) < 1e-12
    assert np.abs(n.real - sellmeier_sio2(lam)).max() < 1e-14
    assert np.all(n.imag == 0)


@pytest.mark.parametrize("name", ["c0", "c1"])
def test_workload_wavenumber_read_back(tmp_path, name):
    """The set-up's k(omega) of the workload is conj(n(omega)) omega / c
    (process_input.cpp:585-623) with n the Sellmeier index at 2 pi c / omega inside
    the table's range and 1 outside it; the table's linear interpolation over its
    20001 points stays within 1e-7 of the formula."""
    cfg = W.write_config(str(tmp_path), name)
    keys = W.workload_keys(name)
    T = int(keys["N_T"])
    om = gsa.setup_table(cfg, "omega", T)
    k = gsa.setup_table(cfg, "wavenum", T, True)
    _, tab_om, _ = read_rawdisp(os.path.join(str(tmp_path), "fused_silica_sellmeier.rawdisp"))
    inside = (om >= tab_om[0]) & (om <= tab_om[-1])
    assert inside.sum() > T // 8
    want = np.where(inside, sellmeier_sio2(2 * np.pi * C_LIGHT / np.abs(om) * 1e6 * inside + ~inside), 1.0) * om / C_LIGHT
    rel = np.abs(k - want) / np.abs(want)
    assert rel.max() < 1e-7, rel.max()
    assert np.all(k.imag == 0)
    # the group index at the carrier, from the set-up's HO table: n_g = n0 + w0 dn/dw
    w0 = 2 * np.pi * C_LIGHT / 0.8e-6
    n0 = sellmeier_sio2(0.8)
    dn = (sellmeier_sio2(2 * np.pi * C_LIGHT / (w0 * 1.0001) * 1e6) - sellmeier_sio2(2 * np.pi * C_LIGHT / (w0 * 0.9999) * 1e6)) / (0.0002 * w0)
    ng = n0 + w0 * dn
    assert abs(ng - 1.4671) < 1e-3


def test_sellmeier_table_against_reference_sio2_rawdisp(tmp_path):
    """The reference ships its own fused-silica table (bin/SiO2.rawdisp, a
    golden fixture here).  Over the Sellmeier table's 0.25-4 um range the two
    indices differ by 2.9e-5 at the median and by 3.3e-3 at most (near the
    ends of the range, where the reference table carries absorption); this
    build's workloads use the Sellmeier medium the benchmark configurations
    name, not the reference's table."""
    _, om, n = read_rawdisp(os.path.join(GOLDEN, "SiO2.rawdisp"))
    lam = 2 * np.pi * C_LIGHT / om * 1e6
    m = (lam > 0.25) & (lam < 4.0)
    d = np.abs(n.real[m] - W.sellmeier_n(lam[m]))
    print("reference SiO2.rawdisp vs Sellmeier: median %.2e, max %.2e over %d points" % (np.median(d), d.max(), m.sum()))
    assert np.median(d) < 1e-4
    assert d.max() < 5e-3
    near = m & (lam > 0.7) & (lam < 0.9)
    assert np.abs(n.real[near] - W.sellmeier_n(lam[near])).max() < 1e-4


def test_batch_configs_draw_the_scan(tmp_path):
    """The 64-run energy/width scan: energies log-uniform in 0.5-10 uJ, w0 in
    40-150 um (written as DIAMETER_FWHM = w0 sqrt(2 ln 2)), grids from the two
    sets, seeded (the same seed gives the same files), each to z_max with
    planes at 0 and z_max; all three grid sizes of each axis occur."""
    import math
    from gsolver_amd import workloads
    a = workloads.batch_configs(str(tmp_path / "a"), 64, 2e-4)
    b = workloads.batch_configs(str(tmp_path / "b"), 64, 2e-4)
    assert len(a) == 64
    nrs, nts, es, ws = set(), set(), [], []
    for fa, fb in zip(a, b):
        ta, tb = open(fa).read(), open(fb).read()
        assert ta == tb
        kv = dict(l.split(None, 1) for l in ta.splitlines() if l.strip() and not l.startswith("%"))
        nrs.add(int(kv["N_R"]))
        nts.add(int(kv["N_T"]))
        es.append(float(kv["ENERGY"]))
        ws.append(float(kv["DIAMETER_FWHM"]) / math.sqrt(2 * math.log(2)))
        assert int(kv["N_Z"]) == 2 and float(kv["Z_MAX"]) == 2e-4
    assert nrs == {128, 256, 512} and nts == {512, 1024, 2048}
    assert 0.5e-6 <= min(es) and max(es) <= 10e-6 and max(es) / min(es) > 5
    assert 40e-6 <= min(ws) and max(ws) <= 150e-6 and max(ws) - min(ws) > 50e-6
