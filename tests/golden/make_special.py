"""Golden vectors for the special functions used by the QDHT plan and the
Keldysh rate table, written with mpmath at 30 significant digits (a library,
not the reference): zeros of J0, J0/J1 on the argument ranges the QDHT matrix
uses, the Dawson integral F(x) = sqrt(pi)/2 exp(-x^2) erfi(x) and the complete
elliptic integrals K(m), E(m).  (scipy.special.j0 is only good to ~1e-15
absolute above x ~ 2000, which is why it is not used here.)

    python tests/golden/make_special.py
"""
import os

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20261202)
    mpmath.mp.dps = 30
    f = lambda fn, xs: np.array([float(fn(mpmath.mpf(float(v)))) for v in xs])
    zeros = np.array([float(mpmath.besseljzero(0, k)) for k in range(1, 2050)])
    x = np.concatenate([np.linspace(0, 30, 601), rng.uniform(0, 3300, 2000), [1e-8, 0.5, 24.999, 25.0, 25.001]])
    m = np.concatenate([np.linspace(0, 1, 101)[:-1], rng.uniform(0, 1, 200), [1 - 1e-6, 1e-9]])
    xd = np.concatenate([np.linspace(-8, 8, 801), rng.uniform(0, 40, 400), [0.19999, 0.2, 1e-6]])
    np.savez_compressed(
        os.path.join(HERE, "special.npz"),
        j0_zeros=zeros,
        x=x, j0=f(lambda v: mpmath.besselj(0, v), x), j1=f(lambda v: mpmath.besselj(1, v), x),
        m=m, K=f(mpmath.ellipk, m), E=f(mpmath.ellipe, m),
        xd=xd, dawson=f(lambda v: mpmath.sqrt(mpmath.pi) / 2 * mpmath.exp(-v * v) * mpmath.erfi(v), xd),
    )


if __name__ == "__main__":
    main()
