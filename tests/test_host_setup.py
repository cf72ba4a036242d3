"""The product's host-side set-up (C++, gs_setup_table: no GPU needed)
against the oracle's restatement, and the C-ABI export list.  CPU only."""
import os
import re

import numpy as np
import pytest

import oracle
from conftest import ROOT, write_cfg, rel_l2

gsa = pytest.importorskip("gsolver_amd")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "gsolver_amd.h")).read()
    names = re.findall(r"^\s*(?:int|void|const char\*)\s+(gs_\w+)\s*\(", hdr, re.M)
    assert len(names) >= 20
    L = gsa.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("kw", [dict(N_R=64, N_T=256), dict(N_R=256, N_T=1024, R_MAX="0.5e-3", T_MIN="-200e-15",
                                                              T_MAX="200e-15")])
def test_setup_tables_match_oracle(tmp_path, kw):
    cfg = write_cfg(tmp_path / "c.txt", **kw)
    o = oracle.Sim(cfg, nthreads=4)
    R, T = o.R, o.T
    for name, n, cplx, tol in [("omega", T, False, 0), ("wavenum", T, True, 1e-15), ("ho", T, False, 1e-15),
                               ("kerr", T, False, 1e-15), ("plasma_func", T, True, 1e-15),
                               ("raman_function", T, True, 1e-13), ("qdht_x", R, False, 1e-14),
                               ("qdht_m1", R, False, 1e-14), ("znet", o.NZ, False, 0)]:
        got = gsa.setup_table(cfg, name, n, cplx)
        e = rel_l2(got, o.get(name))
        assert e <= tol, (name, e)
    C = gsa.setup_table(cfg, "qdht_c", R * R).reshape(R, R)
    assert rel_l2(C, o.get("qdht_c")) < 2e-14
    ion = gsa.setup_table(cfg, "ionrate_ln", 8192)
    assert np.abs(ion - o.get("ionrate_ln")).max() < 1e-12


def test_setup_errors_carry_reference_messages(tmp_path):
    cfg = write_cfg(tmp_path / "c.txt")
    with pytest.raises(gsa.GsError, match="parameter N_Z not found"):
        txt = open(cfg).read().replace("N_Z 3\n", "")
        open(cfg, "w").write(txt)
        gsa.setup_table(cfg, "omega", 256)
    cfg2 = write_cfg(tmp_path / "d.txt", NOISE_LEVEL="0.1")
    with pytest.raises(gsa.GsError, match="NOISE_MAP_LEVEL"):
        gsa.setup_table(cfg2, "omega", 256)


def test_overrides_take_precedence(tmp_path):
    cfg = write_cfg(tmp_path / "c.txt")
    a = gsa.setup_table(cfg, "omega", 256)
    b = gsa.setup_table(cfg, "omega", 256, overrides=["T_MAX", "200e-15", "T_MIN", "-200e-15"])
    assert abs((b[1] - b[0]) / (a[1] - a[0]) - 0.5) < 1e-12


def test_noise_map_is_seeded_and_matches_oracle(tmp_path):
    """The NOISE_MAP_LEVEL/NOISE_SEED amplitude map is identical in both set-ups
    (checked through the oracle's initial field, whose spectrum depends on it)."""
    cfg = write_cfg(tmp_path / "c.txt", NOISE_MAP_LEVEL="0.01", NOISE_SEED="12345")
    o1 = oracle.Sim(cfg)
    o2 = oracle.Sim(cfg)
    assert np.array_equal(o1.get("field"), o2.get("field"))
    cfg0 = write_cfg(tmp_path / "d.txt")
    f0 = oracle.Sim(cfg0).get("field")
    d = rel_l2(o1.get("field"), f0)
    assert 1e-3 < d < 2e-2


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message and status 2."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_roofline_kernels_are_in_the_committed_profile():
    """The kernel names bench.py attaches HBM traffic and VALU counts to are
    the rocprofv3 names of the committed profile (narrow split-BF16 and FP32
    tiles with their K groups, the line kernel), so `traffic` is filled."""
    import importlib
    import sys
    sys.path.insert(0, ROOT)
    bench = importlib.import_module("bench")
    names = [bench.t32_kernel(6, False, 2), bench.t32_kernel(3, True, 2), "k_qdht_f32<false, 32, 2>",
             "k_nl_rk4<PolMix>", bench.t32_kernel(6, False, 2) + " @all-columns"]
    for n in names:
        t = bench.measured_traffic(n)
        assert t is not None and t > 0, n
    assert bench.measured_valu_insts("k_nl_rk4<PolMix>") > 0


def test_record_steps_needs_exactly_one_stop():
    """record_steps stops either after n_steps or at z_end: giving both or
    neither is an error before the plan is touched."""
    from gsolver_amd.adjoint import record_steps
    with pytest.raises(ValueError):
        record_steps(None)
    with pytest.raises(ValueError):
        record_steps(None, n_steps=3, z_end=1e-3)


def test_phase_march_backward_after_later_forward_raises():
    """A plan keeps the checkpoints of its last forward only: the backward of an
    earlier PhaseMarch on the same plan raises (checked with a stand-in plan
    object; no GPU involved)."""
    torch = pytest.importorskip("torch")
    from gsolver_amd.adjoint import PhaseMarch

    class FakePlan:
        R, T = 2, 4

        def put(self, a, buf):
            self.a = np.asarray(a)

        def qdht(self, buf):
            pass

        def set_scalar(self, name, v):
            pass

        def adjoint_forward(self, rk, lin):
            pass

        def get(self, buf):
            return self.a

        def adjoint_backward(self, g, gb2, tie_b2=True):
            return np.asarray(g), np.zeros_like(g)

    p = FakePlan()
    f0 = np.ones((2, 4), dtype=np.complex128)
    phi1 = torch.zeros(4, dtype=torch.float64, requires_grad=True)
    out1 = PhaseMarch.apply(phi1, p, f0, [1e-6], [1e-6])
    phi2 = torch.zeros(4, dtype=torch.float64, requires_grad=True)
    out2 = PhaseMarch.apply(phi2, p, f0, [1e-6], [1e-6])
    (out2.abs() ** 2).sum().backward()          # the latest forward: fine
    assert phi2.grad is not None
    with pytest.raises(RuntimeError, match="another forward"):
        (out1.abs() ** 2).sum().backward()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible here")
def test_plan_without_gpu_fails_loudly(tmp_path):
    """No CPU fallback: without a visible GPU, creating a plan raises (after
    the host set-up has parsed the config) instead of computing anything."""
    cfg = write_cfg(tmp_path / "c.txt", N_R=64, N_T=256)
    with pytest.raises(gsa.GsError, match="no GPU visible"):
        gsa.Plan(cfg, precision="mixed")
