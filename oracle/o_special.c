/* o_special.c -- special functions for the oracle (test infrastructure only).
 *
 * The reference takes J0, J1 and the zeros of J0 from GSL (fhatha.cpp:188-196),
 * the complete elliptic integrals from its own AGM (extmath.cpp:154-171) and
 * the Dawson integral from Cody's rational approximation (extmath.cpp:174-294).
 * These are written here from the mathematics instead:
 *   J0, J1 : Miller's backward recurrence normalised by J0 + 2 sum J_2k = 1 for
 *            |x| < 25; Hankel's asymptotic expansion above, with cos/sin(x) taken
 *            from libm so the phase is exact for large x.
 *   zeros  : McMahon's expansion refined by Newton on J0 (J0' = -J1).
 *   Dawson : Rybicki's sampling-theorem series (Rybicki 1989, Computers in
 *            Physics 3(2) 85) with h = 0.25; Maclaurin series near 0.
 *   K, E   : arithmetic-geometric mean with the Legendre sum for E, the same
 *            textbook iteration as extmath.cpp:154-171 and its 1e-8 stop on c_n.
 * Pinned against scipy.special golden vectors in tests/golden/special.npz.
 */
#include <math.h>
#include "o_internal.h"

static void miller_j01(double x, double* j0, double* j1)
{
    double ax = fabs(x);
    int m = 2 * ((int)(ax + 30.0 + 3.0 * sqrt(ax + 1.0)) / 2) + 20;
    double jp = 0.0, jc = 1e-300, sum = 0.0;
    double r0 = 0, r1 = 0;
    for (int k = m; k >= 1; --k) {
        double jm = (2.0 * k / ax) * jc - jp;   /* J_{k-1} */
        jp = jc;
        jc = jm;
        if (fabs(jc) > 1e250) { jc *= 1e-250; jp *= 1e-250; sum *= 1e-250; r1 *= 1e-250; }
        if (k - 1 == 1) r1 = jc;
        if ((k - 1) % 2 == 0 && k - 1 > 0) sum += 2.0 * jc;
    }
    r0 = jc;
    sum += r0;
    *j0 = r0 / sum;
    *j1 = (x < 0 ? -r1 : r1) / sum;
}

static void hankel_pq(double nu, double x, double* P, double* Q)
{
    /* a_k(nu) = prod_{i=1..k} (4 nu^2 - (2i-1)^2) / (k! 8^k) */
    double mu = 4.0 * nu * nu;
    double p = 1.0, q = 0.0, term = 1.0;
    double last = 1e300;
    for (int k = 1; k < 60; ++k) {
        term *= (mu - (2.0 * k - 1) * (2.0 * k - 1)) / (k * 8.0 * x);
        if (fabs(term) > last) break;
        last = fabs(term);
        /* k odd -> Q with sign (-1)^((k-1)/2); k even -> P with sign (-1)^(k/2) */
        if (k % 2) q += ((k / 2) % 2 ? -term : term);
        else p += ((k / 2) % 2 ? -term : term);
        if (last < 1e-18) break;
    }
    *P = p;
    *Q = q;
}

static void asym_j01(double x, double* j0, double* j1)
{
    double c = cos(x), s = sin(x);
    double amp = sqrt(2.0 / (M_PI * x));
    double P, Q;
    /* J0: w = x - pi/4 : cos w = (c+s)/sqrt2, sin w = (s-c)/sqrt2 */
    hankel_pq(0.0, x, &P, &Q);
    *j0 = amp * (P * (c + s) - Q * (s - c)) * M_SQRT1_2;
    /* J1: w = x - 3pi/4 : cos w = (s-c)/sqrt2, sin w = -(s+c)/sqrt2 */
    hankel_pq(1.0, x, &P, &Q);
    *j1 = amp * (P * (s - c) + Q * (s + c)) * M_SQRT1_2;
}

double o_j0(double x)
{
    double ax = fabs(x), a, b;
    if (ax == 0.0) return 1.0;
    if (ax < 25.0) miller_j01(ax, &a, &b);
    else asym_j01(ax, &a, &b);
    return a;
}

double o_j1(double x)
{
    double ax = fabs(x), a, b;
    if (ax == 0.0) return 0.0;
    if (ax < 25.0) miller_j01(ax, &a, &b);
    else asym_j01(ax, &a, &b);
    return x < 0 ? -b : b;
}

double o_j0_zero(int k)
{
    double beta = (k - 0.25) * M_PI;
    double b8 = 8.0 * beta;
    double x = beta + 1.0 / b8 - 124.0 / (3.0 * b8 * b8 * b8) + 120928.0 / (15.0 * pow(b8, 5));
    for (int it = 0; it < 8; ++it) {
        double j0 = o_j0(x), j1 = o_j1(x);
        double dx = j0 / j1;           /* x_{n+1} = x_n - J0/J0' = x_n + J0/J1 */
        x += dx;
        if (fabs(dx) < 1e-16 * x) break;
    }
    return x;
}

double o_dawson(double x)
{
    double ax = fabs(x);
    if (ax < 0.2) {
        double x2 = x * x, r = 1.0;
        for (int k = 12; k >= 1; --k) r = 1.0 - 2.0 * x2 / (2.0 * k + 1.0) * r;
        return x * r;
    }
    const double h = 0.25;
    int n0 = 2 * (int)floor(0.5 * ax / h + 0.5);
    double xp = ax - n0 * h;
    double sum = 0.0;
    for (int n = -31; n <= 31; n += 2) {
        double d = xp - n * h;
        sum += exp(-d * d) / (double)(n + n0);
    }
    double r = sum / sqrt(M_PI);
    return x < 0 ? -r : r;
}

int o_ellipke(double m, double* K, double* E)
{
    if (m < 0 || m > 1) return -1;
    double a = 1.0, b = sqrt(1.0 - m), c;
    double acc = m;        /* sum_n 2^n c_n^2 with c_0^2 = m */
    double p2 = 1.0;
    do {
        double an = 0.5 * (a + b);
        double bn = sqrt(a * b);
        c = 0.5 * (a - b);
        p2 *= 2.0;
        acc += p2 * c * c;
        a = an;
        b = bn;
    } while (c > 1e-8);
    *K = M_PI / (2.0 * a);
    *E = *K * (1.0 - 0.5 * acc);
    return 0;
}

double orc_bessel_j0(double x) { return o_j0(x); }
double orc_bessel_j1(double x) { return o_j1(x); }
double orc_bessel_zero_j0(int k) { return o_j0_zero(k); }
double orc_dawson(double x) { return o_dawson(x); }
int orc_ellipke(double m, double* K, double* E) { return o_ellipke(m, K, E); }
