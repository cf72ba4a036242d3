"""Gradients through the z-march with torch.autograd.

PhaseMarch maps the spectral phase phi(omega) applied to an input pulse to
the field at the end of a march with a recorded step-size sequence:

    A0(r, omega) = field0(r, omega) exp(i phi(omega)),  BIGBUFFER2 = QDHT(A0),
    field = march(A0)  (gs_adjoint_forward),

and its backward runs the transposed march on the GPU (gs_adjoint_backward):
dL/dphi(omega) = sum_r Im(g0(r, omega) conj(A0(r, omega))), g0 the cotangent
of A0.  The reference has no adjoint; step sizes are held fixed (the gradient
of the march for the recorded dz sequence).  Torch's complex-gradient
convention (dL/dRe z + i dL/dIm z for a real L) is the library's.
"""
import numpy as np
import torch

from . import B2, FIELD


def record_steps(plan, n_steps=None, z_end=None, max_steps=None):
    """Advance `plan` with the reference controller (gs_step) by n_steps
    accepted steps, or until CURRENT_Z reaches z_end (a ZNET plane: the last
    step is clamped onto it) or max_steps steps were taken, and return the
    (dz_rk, dz_lin) arrays that replay them: the NL step of an accepted attempt
    runs at the previous ZSTEP, the linear step at the new one."""
    if (n_steps is None) == (z_end is None):
        raise ValueError("give exactly one of n_steps and z_end")
    if z_end is not None:
        znet = plan.table("znet")
        if z_end > znet[-1]:
            raise ValueError("z_end %.6g lies beyond the last plane ZNET[-1] = %.6g" % (z_end, znet[-1]))
    rk, lin = [], []
    cap = max_steps if max_steps is not None else float("inf")
    while ((len(rk) < n_steps) if n_steps is not None
           else (plan.scalar("CURRENT_Z") < z_end * (1 - 1e-12) and len(rk) < cap)):
        plan.step()
        rk.append(plan.scalar("LAST_DZ_NL"))
        lin.append(plan.scalar("ZSTEP"))
        if not lin[-1] > 0:
            raise RuntimeError("record_steps: the controller returned a zero step at z = %.6g"
                               % plan.scalar("CURRENT_Z"))
    return np.array(rk), np.array(lin)


class PhaseMarch(torch.autograd.Function):
    """field = PhaseMarch.apply(phi, plan, field0, dz_rk, dz_lin)

    phi: real tensor [N_T]; field0: complex [N_R, N_T] (r, omega) numpy array;
    returns the complex128 field [N_R, N_T] after the march (CPU tensor).  The
    plan's state is overwritten: every forward marches from z = 0 and step
    count 0 (the step count seeds the mixed policy's rounding), so the same phi
    gives the same bits on every call.  The plan keeps the checkpoints of its last
    forward only: a backward after a later forward on the same plan raises."""

    @staticmethod
    def forward(ctx, phi, plan, field0, dz_rk, dz_lin):
        ph = phi.detach().to("cpu", torch.float64).numpy()
        A0 = np.asarray(field0, dtype=np.complex128) * np.exp(1j * ph)[None, :]
        plan.put(A0, FIELD)
        plan.put(A0, B2)
        plan.qdht(B2)
        plan.set_scalar("STEP_N", 0)
        plan.set_scalar("CURRENT_Z", 0.0)
        plan.adjoint_forward(dz_rk, dz_lin)
        plan._adjoint_generation = getattr(plan, "_adjoint_generation", 0) + 1
        ctx.plan = plan
        ctx.generation = plan._adjoint_generation
        ctx.A0 = A0
        ctx.phi_meta = (phi.dtype, phi.device)
        return torch.from_numpy(plan.get(FIELD))

    @staticmethod
    def backward(ctx, grad_field):
        if getattr(ctx.plan, "_adjoint_generation", None) != ctx.generation:
            raise RuntimeError("PhaseMarch.backward: the plan has run another forward since this one; "
                               "its checkpoints belong to that march (use one plan per pending graph)")
        g = grad_field.detach().to("cpu", torch.complex128).numpy()
        g0, _ = ctx.plan.adjoint_backward(g, None, tie_b2=True)
        gphi = np.imag(g0 * np.conj(ctx.A0)).sum(axis=0)
        dtype, device = ctx.phi_meta
        return torch.from_numpy(gphi).to(device=device, dtype=dtype), None, None, None, None
