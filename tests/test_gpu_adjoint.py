"""Reverse mode on the GPU (gs_adjoint_forward / gs_adjoint_backward) against
the oracle's transposed march (itself pinned by finite differences), against
central finite differences of the GPU's own FP64 forward, and through the
torch.autograd wrapper."""
import numpy as np
import pytest

import oracle
from conftest import write_cfg, rel_l2

pytestmark = pytest.mark.gpu

gsa = pytest.importorskip("gsolver_amd")

HOT = dict(ENERGY="20e-6", DIAMETER_FWHM="4e-5", N_T="512", N_R="64", R_MAX="0.2e-3")
DZ_RK = np.array([2e-6, 1.5e-6, 2.5e-6])
DZ_LIN = np.array([1.5e-6, 2.5e-6, 2e-6])


def crand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rdot(a, b):
    return float(np.real(np.vdot(a, b)))


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
@pytest.mark.parametrize("nr", [64, 100])
def test_adjoint_march_vs_oracle(tmp_path, prec, nr):
    """Three recorded steps with distinct NL / linear step sizes, cotangents on
    the final FIELD and BIGBUFFER2; N_R = 100 leaves padding rows in the GEMM tiles."""
    cfg = write_cfg(tmp_path / "c.txt", **dict(HOT, N_R=str(nr)))
    p, o = gsa.Plan(cfg, precision=prec), oracle.Sim(cfg, nthreads=8)
    rng = np.random.default_rng(7)
    gF, gB = crand(rng, (p.R, p.T)), crand(rng, (p.R, p.T))
    g0_ref, b0_ref = o.adjoint_march(DZ_RK, DZ_LIN, gF, gB)
    p.adjoint_forward(DZ_RK, DZ_LIN)
    assert rel_l2(p.get(), o.get("field")) < {"fp64": 1e-11, "fp32": 2e-5, "mixed": 2e-5}[prec]
    g0, b0 = p.adjoint_backward(gF, gB, tie_b2=False)
    e1, e2 = rel_l2(g0, g0_ref), rel_l2(b0, b0_ref)
    print("adjoint", prec, nr, e1, e2)
    tol = {"fp64": 1e-10, "fp32": 2e-4, "mixed": 2e-4}[prec]
    assert e1 < tol and e2 < tol
    # a second backward with other cotangents reuses the recorded steps
    gF2 = crand(rng, (p.R, p.T))
    g0b, b0b = p.adjoint_backward(gF2, None, tie_b2=False)
    o2 = oracle.Sim(cfg, nthreads=8)
    g0r2, b0r2 = o2.adjoint_march(DZ_RK, DZ_LIN, gF2, None)
    assert rel_l2(g0b, g0r2) < tol and rel_l2(b0b, b0r2) < tol


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_adjoint_segmented_matches_single_segment(tmp_path, prec):
    """A checkpoint budget below one field per step makes the backward recompute
    segments (7 steps: K = 4, segments of 4 and 3) from FIELD / BIGBUFFER2
    snapshots; the replay runs the same kernels in the same order, so the
    cotangents are identical to the one-segment run."""
    cfg = write_cfg(tmp_path / "c.txt", **HOT)
    rk = np.array([2e-6, 1.5e-6, 2.5e-6, 1e-6, 2e-6, 1.8e-6, 2.2e-6])
    lin = np.roll(rk, -1)
    rng = np.random.default_rng(11)
    gF, gB = crand(rng, (64, 512)), crand(rng, (64, 512))
    out = []
    for budget in (None, 1.0):
        p = gsa.Plan(cfg, precision=prec)
        if budget is not None:
            p.set_scalar("ADJOINT_CHECKPOINT_BYTES", budget)
        p.adjoint_forward(rk, lin)
        out.append((p.scalar("ADJOINT_SEGMENT"), p.get()) + p.adjoint_backward(gF, gB, tie_b2=True))
        p.close()
    assert out[0][0] == 7 and out[1][0] == 4
    for a, b in zip(out[0][1:], out[1][1:]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_adjoint_forward_is_the_march(tmp_path, prec):
    """gs_adjoint_forward replays recorded steps with the march's own
    transforms (column skipping, splits, tile shapes, rounding seeds): from
    the same initial state it ends on the FIELD and BIGBUFFER2 of the
    controller's march bit for bit (N_R = 256)."""
    from gsolver_amd.adjoint import record_steps
    cfg = write_cfg(tmp_path / "c.txt", **dict(HOT, N_R="256"))
    rec = gsa.Plan(cfg, precision=prec)
    assert rec.scalar("QDHT_SKIP_COLUMNS") == 1
    dz_rk, dz_lin = record_steps(rec, n_steps=6)
    f_rec, b_rec = rec.get(), rec.get(gsa.B2)
    rec.close()
    p = gsa.Plan(cfg, precision=prec)
    p.adjoint_forward(dz_rk, dz_lin)
    assert np.array_equal(p.get(), f_rec)
    assert np.array_equal(p.get(gsa.B2), b_rec)
    p.close()


@pytest.mark.parametrize("case", ["no_raman", "plasma_only", "kerr_only"])
def test_adjoint_physics_variants_vs_oracle(tmp_path, case):
    """The transposed march with parts of the RHS switched off (FP64, 3 steps):
    no Raman response (RAMAN_FRACTION 0: y = I), Kerr negligible (the plasma
    and loss transposes alone), and a field below the ionization threshold
    (Kerr and Raman alone, plasma terms skipped)."""
    kw = dict(HOT)
    if case == "no_raman":
        kw["RAMAN_FRACTION"] = "0"
    elif case == "plasma_only":
        kw["NONLIN_REFRINDEX"] = "1e-30"
    else:
        kw["ENERGY"] = "1e-7"
    cfg = write_cfg(tmp_path / "c.txt", **kw)
    p, o = gsa.Plan(cfg, precision="fp64"), oracle.Sim(cfg, nthreads=8)
    rng = np.random.default_rng(12)
    gF, gB = crand(rng, (p.R, p.T)), crand(rng, (p.R, p.T))
    g0_ref, b0_ref = o.adjoint_march(DZ_RK, DZ_LIN, gF, gB)
    p.adjoint_forward(DZ_RK, DZ_LIN)
    g0, b0 = p.adjoint_backward(gF, gB, tie_b2=False)
    e1, e2 = rel_l2(g0, g0_ref), rel_l2(b0, b0_ref)
    print("adjoint", case, e1, e2)
    assert e1 < 1e-10 and e2 < 1e-10


def test_adjoint_tied_b2(tmp_path):
    """tie_b2 folds the initial BIGBUFFER2 = QDHT(FIELD) cotangent into FIELD's:
    gfield0(tied) = gfield0 + Q^T gb2_0, Q = diag(m1) C diag(1/m1)."""
    cfg = write_cfg(tmp_path / "c.txt", **HOT)
    p = gsa.Plan(cfg, precision="fp64")
    rng = np.random.default_rng(8)
    gF = crand(rng, (p.R, p.T))
    p.adjoint_forward(DZ_RK, DZ_LIN)
    g0, b0 = p.adjoint_backward(gF, None, tie_b2=False)
    g0t, _ = p.adjoint_backward(gF, None, tie_b2=True)
    m1, Cm = p.table("qdht_m1"), p.table("qdht_c")
    Q = m1[:, None] * Cm / m1[None, :]
    assert rel_l2(g0t, g0 + Q.T @ b0) < 1e-12


def test_adjoint_directional_derivative_fp64(tmp_path):
    """Re<g0, v> against central differences of the GPU FP64 forward, at a size
    beyond the oracle tests (N_R = 128, N_T = 1024), cotangent of the on-axis
    spectral energy L = sum_w |FIELD[0, w]|^2 plus a random linear functional."""
    cfg = write_cfg(tmp_path / "c.txt", **dict(HOT, N_T="1024", N_R="128", T_MIN="-150e-15", T_MAX="150e-15"))
    p = gsa.Plan(cfg, precision="fp64")
    F0, B0 = p.get(gsa.FIELD), p.get(gsa.B2)
    rng = np.random.default_rng(9)
    lam = 1e-3 * crand(rng, F0.shape) * np.abs(F0).max()

    def loss(F):
        return float(np.sum(np.abs(F[0]) ** 2)) + rdot(lam, F)

    def run(F, B):
        p.put(F, gsa.FIELD)
        p.put(B, gsa.B2)
        p.adjoint_forward(DZ_RK, DZ_LIN)
        return p.get(gsa.FIELD)

    F = run(F0, B0)
    gF = lam.copy()
    gF[0] += 2 * F[0]
    g0, b0 = p.adjoint_backward(gF, None, tie_b2=False)
    worst = 0.0
    for _ in range(3):
        vF, vB = crand(rng, F0.shape), crand(rng, F0.shape)
        vF *= np.linalg.norm(F0) / np.linalg.norm(vF)
        vB *= np.linalg.norm(B0) / np.linalg.norm(vB)
        # W(I) is piecewise ln-ln linear: a probe step crosses table nodes in
        # proportion to its size, so the step is kept small
        e = 1e-7
        d = (loss(run(F0 + e * vF, B0 + e * vB)) - loss(run(F0 - e * vF, B0 - e * vB))) / (2 * e)
        a = rdot(g0, vF) + rdot(b0, vB)
        worst = max(worst, abs(a - d) / abs(d))
    print("directional", worst)
    assert worst < 1e-6


def test_phase_march_autograd(tmp_path):
    """torch.autograd through PhaseMarch: d/dphi of the on-axis L = sum_t |E(0, t)|^4
    (E = IFFT of the output line; a spectral phase reshapes it in time) along
    random phase directions, against central differences of the same FP64
    forward.  The recorded steps are those of the reference controller."""
    torch = pytest.importorskip("torch")
    from gsolver_amd.adjoint import PhaseMarch, record_steps
    cfg = write_cfg(tmp_path / "c.txt", **HOT)
    p = gsa.Plan(cfg, precision="fp64")
    field0 = p.get(gsa.FIELD)
    dz_rk, dz_lin = record_steps(p, 4)
    assert np.all(dz_rk > 0) and np.all(dz_lin > 0)
    rng = np.random.default_rng(10)
    phi0 = 0.1 * rng.standard_normal(p.T)

    def L(phi_t):
        out = PhaseMarch.apply(phi_t, p, field0, dz_rk, dz_lin)
        return (torch.fft.ifft(out[0]).abs() ** 4).sum()

    phi = torch.tensor(phi0, requires_grad=True)
    val = L(phi)
    val.backward()
    g = phi.grad.numpy().copy()
    worst = 0.0
    for _ in range(3):
        v = rng.standard_normal(p.T)
        e = 1e-5
        with torch.no_grad():
            d = (L(torch.tensor(phi0 + e * v)).item() - L(torch.tensor(phi0 - e * v)).item()) / (2 * e)
        worst = max(worst, abs(g @ v - d) / abs(d))
    print("phase", worst)
    assert worst < 1e-6


def test_phase_march_repeatable_mixed(tmp_path):
    """Mixed precision: every PhaseMarch forward starts from step count 0 (the
    seed of the unbiased rounding), so a second gradient of the same phi, and
    one with sqrt-segmented checkpoints, repeat the first bit for bit
    (N_R = 256)."""
    torch = pytest.importorskip("torch")
    from gsolver_amd.adjoint import PhaseMarch, record_steps
    cfg = write_cfg(tmp_path / "c.txt", **dict(HOT, N_R="256"))
    rec = gsa.Plan(cfg, precision="mixed")
    field0 = rec.get(gsa.FIELD)
    dz_rk, dz_lin = record_steps(rec, n_steps=9)
    rec.close()
    p = gsa.Plan(cfg, precision="mixed")
    rng = np.random.default_rng(3)
    phi0 = 0.1 * rng.standard_normal(p.T)
    grads = []
    for budget in (-1.0, -1.0, 1.0):
        p.set_scalar("ADJOINT_CHECKPOINT_BYTES", budget)
        phi = torch.tensor(phi0, requires_grad=True)
        (torch.fft.ifft(PhaseMarch.apply(phi, p, field0, dz_rk, dz_lin)[0]).abs() ** 4).sum().backward()
        grads.append(phi.grad.numpy().copy())
    assert p.scalar("ADJOINT_SEGMENT") < 9
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])
    p.close()
