"""The benchmark's own workloads (gsolver_amd/workloads.py) on the HIP path
against the oracle, through the C-ABI:

- the 1024 x 4096 production grid (fused-silica Sellmeier medium, 5 uJ 40 fs
  w0 = 100 um pulse with the seeded 1 % amplitude noise map): one QDHT and
  one fixed z-step (FIELD and BIGBUFFER2) in every precision;
- the 256 x 1024 workload (1 uJ 50 fs w0 = 60 um, 0.5 mm window): an
  adaptive march of ~150 steps with flux, plasma_max and field_axis dumps at
  10, 20 and 30 um, the same step count as the oracle;
- the 512 x 2048 reverse-mode workload: three recorded steps of the
  transposed march;
- the energy / width scan over uneven grids: eight of its seeded runs marched
  together equal the same runs marched alone, bit for bit, and one of them
  matches the oracle.

Tolerances (relative L2): FP64 1e-11 for one step, FP32 and mixed 2e-5; the
adaptive march 1e-4 (FP64 1e-10)."""
import os

import numpy as np
import pytest

import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

gsa = pytest.importorskip("gsolver_amd")
from gsolver_amd import workloads as W   # noqa: E402

NTH = min(16, os.cpu_count() or 1)
STEP_TOL = {"fp64": 1e-11, "fp32": 2e-5, "mixed": 2e-5}


@pytest.fixture(scope="module")
def c1(tmp_path_factory):
    """The production workload, its initial fields, and the FP64 oracle's
    QDHT of FIELD and one fixed step at the config's initial ZSTEP."""
    d = str(tmp_path_factory.mktemp("c1"))
    cfg = W.write_config(d, "c1")
    o = oracle.Sim(cfg, nthreads=NTH)
    f0 = o.get("field")
    out = dict(cfg=cfg, field0=f0, qdht=o.qdht(f0), dz=o.scalar("ZSTEP"))
    o.fixed_step(out["dz"])
    out["field1"], out["b2_1"] = o.get("field"), o.get("b2")
    o.close()
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_production_grid_qdht_vs_oracle(c1, prec):
    """N_R = 1024 (8 x 32 tiles of the mixed GEMM, XCD-ordered) on all 4096 columns."""
    p = gsa.Plan(c1["cfg"], precision=prec)
    assert (p.R, p.T) == (1024, 4096)
    assert rel_l2(p.get(), c1["field0"]) < (1e-15 if prec == "fp64" else 1e-7)
    p.qdht(gsa.FIELD)
    e = rel_l2(p.get(), c1["qdht"])
    print("c1 qdht", prec, e)
    assert e < {"fp64": 1e-12, "fp32": 1e-5, "mixed": 1e-5}[prec]
    p.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_production_grid_fixed_step_vs_oracle(c1, prec):
    p = gsa.Plan(c1["cfg"], precision=prec)
    assert abs(p.scalar("ZSTEP") / c1["dz"] - 1) < 1e-12
    p.fixed_step(c1["dz"])
    ef, eb = rel_l2(p.get(), c1["field1"]), rel_l2(p.get(gsa.B2), c1["b2_1"])
    print("c1 step", prec, ef, eb)
    assert ef < STEP_TOL[prec] and eb < STEP_TOL[prec]
    p.close()


C0_KEYS = dict(N_Z=4, Z_MAX="3e-5", DUMPS_N="3", DUMPTYPE0="flux", DUMP_FILENAME0="c0%s_f.udf",
               DUMPTYPE1="plasma_max", DUMP_FILENAME1="c0%s_pm.udf", DUMPTYPE2="field_axis",
               DUMP_FILENAME2="c0%s_fa.udf")


@pytest.fixture(scope="module")
def c0(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("c0"))
    cfg = W.write_config(d, "c0", **C0_KEYS)
    od = os.path.join(d, "oracle")
    os.makedirs(od)
    o = oracle.Sim(cfg, nthreads=NTH)
    o.run(od)
    out = dict(cfg=cfg, dir=d, steps=o.scalar("STEP_N"), rejects=o.scalar("n_reject"), z=o.scalar("CURRENT_Z"),
               field=o.get("field"), files={k: oracle.load_udf(os.path.join(od, "c0_%s.udf" % k))[1]
                                            for k in ("f", "pm", "fa")})
    o.close()
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_c0_adaptive_march_vs_oracle(c0, prec, tmp_path):
    """The reference controller from z = 0 to 30 um (about 150 accepted steps,
    ionizing: the plasma_max dump is non-zero): same step count, and the dumps
    at every plane and the final field within tolerance."""
    p = gsa.Plan(c0["cfg"], precision=prec)
    out = str(tmp_path)
    assert p.march(out) == 0
    assert p.scalar("STEP_N") == c0["steps"] and p.scalar("n_reject") == c0["rejects"]
    assert c0["steps"] >= 100
    assert abs(p.scalar("CURRENT_Z") - c0["z"]) < 1e-15
    tol = 1e-10 if prec == "fp64" else 1e-4
    ef = rel_l2(p.get(), c0["field"])
    print("c0 march", prec, "steps", c0["steps"], "field", ef)
    assert ef < tol
    for k in ("f", "pm", "fa"):
        got, ref = oracle.load_udf(os.path.join(out, "c0_%s.udf" % k))[1], c0["files"][k]
        assert got.shape == ref.shape
        for z in range(ref.shape[0]):
            e = rel_l2(got[z], ref[z])
            print("   dump", k, "plane", z, e)
            assert e < tol, (k, z, e)
    assert np.abs(c0["files"]["pm"][-1]).max() > 0
    p.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_reverse_mode_workload_vs_oracle(tmp_path, prec):
    """512 x 2048 (the gradient workload's grid and medium): three steps at the
    controller's first step sizes, cotangents on the final FIELD and BIGBUFFER2."""
    cfg = W.write_config(str(tmp_path), "c3", ENERGY="0.3e-6")
    p, o = gsa.Plan(cfg, precision=prec), oracle.Sim(cfg, nthreads=NTH)
    dz = p.scalar("ZSTEP")
    rk = np.array([dz, 0.9 * dz, 1.1 * dz])
    lin = np.array([0.9 * dz, 1.1 * dz, dz])
    rng = np.random.default_rng(3)
    gF = rng.standard_normal((p.R, p.T)) + 1j * rng.standard_normal((p.R, p.T))
    gB = rng.standard_normal((p.R, p.T)) + 1j * rng.standard_normal((p.R, p.T))
    g0_ref, b0_ref = o.adjoint_march(rk, lin, gF, gB)
    p.adjoint_forward(rk, lin)
    assert rel_l2(p.get(), o.get("field")) < STEP_TOL[prec]
    g0, b0 = p.adjoint_backward(gF, gB, tie_b2=False)
    e1, e2 = rel_l2(g0, g0_ref), rel_l2(b0, b0_ref)
    print("c3 adjoint", prec, e1, e2)
    tol = {"fp64": 1e-10, "fp32": 2e-4, "mixed": 2e-4}[prec]
    assert e1 < tol and e2 < tol
    p.close()
    o.close()


def test_energy_width_scan_runs(tmp_path):
    """Eight runs of the seeded scan (uneven grids), 2 um each in mixed
    precision: gs_march_many gives every run's field bit for bit as its single
    march, and the smallest run agrees with the FP64 oracle to 1e-4."""
    d = str(tmp_path)
    cfgs = W.batch_configs(d, nruns=8, z_max=2e-6)
    grids = []
    single = []
    for c in cfgs:
        p = gsa.Plan(c, precision="mixed")
        grids.append(p.R * p.T)
        assert p.march(d) == 0
        single.append(p.get())
        p.close()
    assert len(set(grids)) > 2
    plans = [gsa.Plan(c, precision="mixed") for c in cfgs]
    gsa.march_many(plans, [d] * len(plans))
    for p, f in zip(plans, single):
        assert np.array_equal(p.get(), f)
        p.close()
    i = int(np.argmin(grids))
    o = oracle.Sim(cfgs[i], nthreads=NTH)
    od = os.path.join(d, "oracle_run")
    os.makedirs(od)
    o.run(od)
    e = rel_l2(single[i], o.get("field"))
    print("scan run", i, "grid", grids[i], "vs oracle", e)
    assert e < 1e-4
    o.close()
