// gs_special.cpp -- special functions for the plan set-up (host, FP64).
//
// The reference takes J0/J1/zeros of J0 from GSL (fhatha.cpp:188-196), K/E
// from its own AGM (extmath.cpp:154-171) and the Dawson integral from Cody's
// rational fit (extmath.cpp:174-294).  Here:
//   J0, J1   std::cyl_bessel_j (C++17 special math) below x = 25, where it is
//            good to ~1e-14; above, Hankel's asymptotic expansion truncated at
//            its smallest term, with the phase built from cos(x), sin(x) so the
//            argument reduction is exact (std::cyl_bessel_j loses ~1e-11
//            between x = 100 and 1000, which the QDHT matrix would inherit)
//   zeros    McMahon's asymptotic start + Newton on J0 (J0' = -J1)
//   Dawson   e^{-x^2} sum x^{2k+1}/(k!(2k+1)) (all terms positive) for |x| < 8,
//            the asymptotic series 1/(2x) sum (2k-1)!!/(2x^2)^k above
//   K, E     arithmetic-geometric mean, E from the Legendre sum of 2^n c_n^2
#include <cmath>

#include "gs_host.h"

namespace gs {

static void hankel(double nu, double x, double* P, double* Q)
{
    const double mu = 4.0 * nu * nu;
    double p = 1.0, q = 0.0, term = 1.0, last = 1e300;
    for (int k = 1; k < 80; ++k) {
        term *= (mu - (2.0 * k - 1) * (2.0 * k - 1)) / (8.0 * k * x);
        if (std::fabs(term) > last) break;
        last = std::fabs(term);
        const double sgn = ((k / 2) & 1) ? -1.0 : 1.0;
        if (k & 1) q += sgn * term; else p += sgn * term;
        if (last < 1e-19) break;
    }
    *P = p;
    *Q = q;
}

double bessel_j0(double x)
{
    x = std::fabs(x);
    if (x < 25.0) return std::cyl_bessel_j(0.0, x);
    double P, Q, c = std::cos(x), s = std::sin(x);
    hankel(0.0, x, &P, &Q);
    return std::sqrt(2.0 / (M_PI * x)) * (P * (c + s) - Q * (s - c)) * M_SQRT1_2;
}

double bessel_j1(double x)
{
    const double ax = std::fabs(x);
    double v;
    if (ax < 25.0) {
        v = std::cyl_bessel_j(1.0, ax);
    } else {
        double P, Q, c = std::cos(ax), s = std::sin(ax);
        hankel(1.0, ax, &P, &Q);
        v = std::sqrt(2.0 / (M_PI * ax)) * (P * (s - c) + Q * (s + c)) * M_SQRT1_2;
    }
    return x < 0 ? -v : v;
}

double bessel_j0_zero(int k)
{
    const double b = (k - 0.25) * M_PI;
    const double e = 1.0 / (8.0 * b);
    double x = b + e - (124.0 / 3.0) * e * e * e + (120928.0 / 15.0) * e * e * e * e * e;
    for (int it = 0; it < 10; ++it) {
        const double dx = bessel_j0(x) / bessel_j1(x);
        x += dx;
        if (std::fabs(dx) <= 1e-16 * x) break;
    }
    return x;
}

double dawson(double x)
{
    const double ax = std::fabs(x);
    double r;
    if (ax < 8.0) {
        const double x2 = ax * ax;
        double term = ax, sum = ax;   // k = 0 term: x / (0! * 1)
        for (int k = 1; k < 400; ++k) {
            term *= x2 / k;               // x^{2k+1}/k!
            const double t = term / (2.0 * k + 1.0);
            sum += t;
            if (t < 1e-18 * sum) break;
        }
        r = std::exp(-x2) * sum;
    } else {
        const double y = 1.0 / (2.0 * ax * ax);
        double term = 1.0, sum = 1.0, last = 1e300;
        for (int k = 1; k < 200; ++k) {
            term *= (2.0 * k - 1.0) * y;
            if (term > last) break;
            last = term;
            sum += term;
            if (term < 1e-18 * sum) break;
        }
        r = sum / (2.0 * ax);
    }
    return x < 0 ? -r : r;
}

void ellipke(double m, double* K, double* E)
{
    if (m < 0 || m > 1) throw Error(-4, "ellipke(): Complete elliptical intgrals of both kinds are defined for 0 < z < 1");
    double a = 1.0, b = std::sqrt(1.0 - m), c = 0.0, sum = m, w = 1.0;
    do {
        const double an = 0.5 * (a + b);
        c = 0.5 * (a - b);
        b = std::sqrt(a * b);
        a = an;
        w *= 2.0;
        sum += w * c * c;
    } while (c > 1e-8);
    *K = M_PI / (2.0 * a);
    *E = *K * (1.0 - 0.5 * sum);
}

// Keldysh rate for a solid, ionization.cpp:203-232 (single level); returns ln W.
double keldysh_rate_ln(const Setup& s, double lnI, double U, double reduced_mass)
{
    const double e = 1.60217733e-19, me = 9.1093897e-31, hbar = 1.0545727e-34, mu0 = 1.2566370615e-6;
    const double w0 = s.OMEGA0;
    const double vp = w0 / s.WAVENUMBER0.real();
    const double E = std::exp(lnI / 2) * std::sqrt(2 * vp / s.GROUP_VELOCITY * vp * mu0);
    const double m = me * reduced_mass;
    const double g = w0 * std::sqrt(m / e / e * U) / E;
    const double G = g * g / (1 + g * g), X = 1 / (1 + g * g);
    double Kg, Eg, Kx, Ex;
    ellipke(G, &Kg, &Eg);
    ellipke(X, &Kx, &Ex);
    const double alpha = M_PI * (Kg - Eg) / Ex, beta = M_PI * M_PI / 4.0 / Kx / Ex;
    const double x = 2 / M_PI * U / hbar / w0 * Ex / std::sqrt(G);
    const double nu = std::floor(x + 1) - x;
    int n = (int)std::ceil(10 / alpha);
    if (n > 1000) n = 1000;
    double Q = 0;
    for (int i = 0; i < n; ++i) Q += std::exp(-i * alpha) * dawson(std::sqrt(beta * (i + 2 * nu)));
    Q *= std::sqrt(M_PI / 2 / Kx);
    return std::log(2 * w0 / 9 / M_PI) + 1.5 * std::log(w0 * m / hbar) - 0.75 * std::log(G) + std::log(Q) -
           alpha * std::floor(x + 1);
}

}  // namespace gs
