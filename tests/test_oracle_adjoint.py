"""The oracle's reverse-mode derivative (o_adjoint.c) against central finite
differences of its own forward restatement: one RHS evaluation, the RK4 NL
step, and a three-step march with distinct NL / linear step sizes.  Each check
compares Re<g, v> (g the adjoint cotangent) with the directional derivative
of L = Re<lam, output> along a random direction v.  CPU only, small grids."""
import numpy as np
import pytest

import oracle
from conftest import write_cfg

# fields hot enough to ionize (loss gate, plasma and avalanche active)
HOT = dict(ENERGY="26e-6", DIAMETER_FWHM="4e-5", N_T="256", N_R="16", R_MAX="0.2e-3")


def crand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rdot(a, b):
    return float(np.real(np.vdot(a, b)))


def fd_check(f, x, g, lam, rng, rel_eps=1e-6, ndir=3):
    """max relative mismatch of Re<g, v> vs d/de Re<lam, f(x + e v)> (central)"""
    worst = 0.0
    for _ in range(ndir):
        v = crand(rng, x.shape)
        v *= np.linalg.norm(x) / np.linalg.norm(v)
        e = rel_eps
        d = (rdot(lam, f(x + e * v)) - rdot(lam, f(x - e * v))) / (2 * e)
        a = rdot(g, v)
        worst = max(worst, abs(a - d) / max(abs(d), 1e-300))
    return worst


def hot_line(sim):
    F = sim.get("field")
    return F[int(np.argmax(np.abs(F).max(axis=1)))]


@pytest.mark.parametrize("case", ["hot", "kerr_only", "no_raman", "plasma_only"])
def test_nl_response_vjp_matches_finite_differences(tmp_path, case):
    kw = dict(HOT)
    if case == "kerr_only":
        kw["ENERGY"] = "1e-7"          # far below the ionization threshold
    if case == "no_raman":
        kw["RAMAN_FRACTION"] = "0"
    if case == "plasma_only":
        kw["NONLIN_REFRINDEX"] = "1e-30"   # Kerr negligible: the plasma and loss adjoint alone
    s = oracle.Sim(write_cfg(tmp_path / "c.txt", **kw))
    a = hot_line(s)
    E = np.fft.ifft(a)
    if case in ("hot", "plasma_only"):
        rho, _ = s.plasma(E)
        assert rho.max() > 0 and (np.abs(E) ** 2).max() > 29.94   # plasma and loss terms active
    rng = np.random.default_rng(1)
    lam = crand(rng, a.shape)
    g = s.nl_response_vjp(a, lam)
    err = fd_check(s.nl_response, a, g, lam, rng)
    print(case, err)
    assert err < 1e-6


def test_nl_step_vjp_matches_finite_differences(tmp_path):
    s = oracle.Sim(write_cfg(tmp_path / "c.txt", **HOT), nthreads=4)
    A = s.get("field")
    dz = 2e-6

    def b1(x):
        s.put("field", x)
        s.set_scalar("ZSTEP", dz)
        s.nl_step()
        return s.get("b1")

    rng = np.random.default_rng(2)
    lam = crand(rng, A.shape)
    g = s.nl_step_vjp(A, lam, dz)
    err = fd_check(b1, A, g, lam, rng)
    print("nl_step", err)
    assert err < 1e-6


def test_adjoint_march_matches_finite_differences(tmp_path):
    """Three steps, NL at dz_rk[k] and linear step at dz_lin[k] (an accepted
    attempt after a step-size change), cotangents on both FIELD and BIG2."""
    cfg = write_cfg(tmp_path / "c.txt", **HOT)
    s = oracle.Sim(cfg, nthreads=4)
    F0, B0 = s.get("field"), s.get("b2")
    dz_rk = np.array([2e-6, 1.5e-6, 2.5e-6])
    dz_lin = np.array([1.5e-6, 2.5e-6, 2e-6])
    rng = np.random.default_rng(3)
    lamF, lamB = crand(rng, F0.shape), crand(rng, F0.shape)

    def run(F, B):
        s.put("field", F)
        s.put("b2", B)
        for a, b in zip(dz_rk, dz_lin):
            s.fixed_step2(a, b)
        return s.get("field"), s.get("b2")

    s.put("field", F0)
    s.put("b2", B0)
    g0, gb0 = s.adjoint_march(dz_rk, dz_lin, lamF, lamB)
    worst = 0.0
    for _ in range(3):
        vF, vB = crand(rng, F0.shape), crand(rng, F0.shape)
        vF *= np.linalg.norm(F0) / np.linalg.norm(vF)
        vB *= np.linalg.norm(B0) / np.linalg.norm(vB)
        e = 1e-6
        Fp, Bp = run(F0 + e * vF, B0 + e * vB)
        Fm, Bm = run(F0 - e * vF, B0 - e * vB)
        d = (rdot(lamF, Fp - Fm) + rdot(lamB, Bp - Bm)) / (2 * e)
        a = rdot(g0, vF) + rdot(gb0, vB)
        worst = max(worst, abs(a - d) / abs(d))
    print("march", worst)
    assert worst < 1e-6
