"""Pins the oracle's numerical building blocks against independent checks:
brute-force DFT, numpy.fft, Parseval, scipy.special golden vectors, closed
forms.  CPU only."""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 12, 16, 43, 64, 129, 516])
def test_fft_bruteforce(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    k = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(k, k) / n)
    np.testing.assert_allclose(oracle.fft(x, -1), F @ x, rtol=0, atol=1e-12 * max(1, np.abs(F @ x).max()))
    np.testing.assert_allclose(oracle.fft(x, +1), F.conj() @ x, rtol=0, atol=1e-12 * max(1, np.abs(F @ x).max()))


@pytest.mark.parametrize("n", [512, 1024, 4096, 516, 2048])
def test_fft_numpy_parseval(n):
    rng = np.random.default_rng(7)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    X = oracle.fft(x, -1)
    ref = np.fft.fft(x)
    assert np.abs(X - ref).max() / np.abs(ref).max() < 5e-15
    # FFTW backward is the unnormalised inverse
    assert np.abs(oracle.fft(X, +1) / n - x).max() < 1e-14
    assert abs(np.vdot(X, X).real / n - np.vdot(x, x).real) / np.vdot(x, x).real < 1e-13


def test_fft_fp32_build():
    x = np.exp(2j * np.pi * 3 * np.arange(1024) / 1024)
    X = oracle.fft(x, -1, precision="fp32")
    assert abs(X[3] - 1024) < 1e-3 and np.abs(np.delete(X, 3)).max() < 1e-3


def test_bessel_golden():
    g = np.load(os.path.join(GOLDEN, "special.npz"))
    L = oracle.lib()
    j0 = np.array([L.orc_bessel_j0(v) for v in g["x"]])
    j1 = np.array([L.orc_bessel_j1(v) for v in g["x"]])
    # absolute error relative to the envelope sqrt(2/(pi x)) (values cross zero)
    env = np.minimum(1.0, np.sqrt(2 / (np.pi * np.maximum(g["x"], 1e-300))))
    assert np.max(np.abs(j0 - g["j0"]) / env) < 5e-15
    assert np.max(np.abs(j1 - g["j1"]) / env) < 5e-15
    zeros = np.array([L.orc_bessel_zero_j0(k) for k in range(1, len(g["j0_zeros"]) + 1)])
    assert np.max(np.abs(zeros - g["j0_zeros"]) / g["j0_zeros"]) < 2e-15


def test_dawson_ellipke_golden():
    g = np.load(os.path.join(GOLDEN, "special.npz"))
    L = oracle.lib()
    d = np.array([L.orc_dawson(v) for v in g["xd"]])
    assert np.max(np.abs(d - g["dawson"]) / np.maximum(np.abs(g["dawson"]), 1e-300)) < 1e-13
    import ctypes as C
    K, E = C.c_double(), C.c_double()
    ks, es = [], []
    for m in g["m"]:
        assert L.orc_ellipke(m, C.byref(K), C.byref(E)) == 0
        ks.append(K.value)
        es.append(E.value)
    assert np.max(np.abs(np.array(ks) - g["K"]) / g["K"]) < 1e-13
    assert np.max(np.abs(np.array(es) - g["E"]) / g["E"]) < 1e-13
    assert L.orc_ellipke(1.5, C.byref(K), C.byref(E)) != 0


def test_sqrtho_taylor():
    L = oracle.lib()
    for x in (-0.25, -0.1, -1e-3, 0.0, 1e-4):
        assert abs(L.orc_sqrtho(x) - (np.sqrt(1 + x) - 1)) < 2e-9
    # 20 terms: at x=-0.25 the truncation is ~ c20 x^21 ~ 1e-15
    assert abs(L.orc_sqrtho(-0.25) - (np.sqrt(0.75) - 1)) < 1e-14


def test_raman_closed_form_matches_circular_fft():
    """y = IFFT(FFT(I) * (FFT(R) + 1 - fR)) / T  (solver_main.cpp:401-413,
    propagator.cu:696-703) equals the scan form s[n] = z s[n-1] + I[n]."""
    import ctypes as C
    T = 1024
    dt = 400e-15 / T
    fr, tau, Om = 0.18, 32e-15, 2 * np.pi * 13.2e12
    M = dt * (1 + Om**2 * tau**2) / tau / (Om * tau)
    rng = np.random.default_rng(3)
    Iv = rng.uniform(0, 1, T)
    t = dt * np.arange(T)
    R = fr * M * np.exp(-t / tau) * np.sin(Om * t)
    ref = np.fft.ifft(np.fft.fft(Iv) * (np.fft.fft(R) + (1 - fr))).real
    z = np.exp((-1 / tau + 1j * Om) * dt)
    y = np.empty(T)
    oracle.lib().orc_raman_circular(T, oracle._dptr(Iv), fr, M, z.real, z.imag, oracle._dptr(y))
    assert np.abs(y - ref).max() / np.abs(ref).max() < 1e-13
    # dropping the wrap term is measurably wrong when I is non-zero at the edges
    s = np.zeros(T, complex)
    acc = 0
    for n in range(T):
        acc = z * acc + Iv[n]
        s[n] = acc
    causal = (1 - fr) * Iv + fr * M * s.imag
    assert np.abs(causal - ref).max() / np.abs(ref).max() > 1e-4
