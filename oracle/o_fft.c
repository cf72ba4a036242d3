/* o_fft.c -- mixed-radix FFT for the oracle (test infrastructure only).
 *
 * Stands in for the FFTW plans of the reference (solver_main.cpp:339-341):
 * FFTW_FORWARD = sum x[n] e^{-2 pi i k n / N}, FFTW_BACKWARD with e^{+...},
 * both unnormalised (fftwt_Nnormalize, solver.h:304-305, divides separately).
 * Any length is accepted (the shipped configs use 512 and 516 = 4*3*43):
 * recursive decimation in time over the factors 4, 2, 3, 5 and then any
 * remaining primes, which are done as direct DFTs.
 * Pinned by a brute-force DFT, numpy.fft and Parseval in tests/test_oracle_numerics.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "o_internal.h"

static void twiddle_table(cplx_t* t, int n, int sign)
{
    /* exact-symmetry evaluation of e^{sign 2 pi i j / n} in double, then rounded */
    for (int j = 0; j < n; ++j) {
        /* reduce j/n to an octant for accuracy */
        long jj = j;
        double c, s;
        long n8 = (long)n * 8;
        long a = jj * 8;                 /* angle = 2 pi a / n8 */
        if (a * 2 == n8) { c = -1.0; s = 0.0; }
        else if (a * 4 == n8) { c = 0.0; s = 1.0; }
        else if (a * 4 == 3 * n8) { c = 0.0; s = -1.0; }
        else if (a == 0) { c = 1.0; s = 0.0; }
        else {
            double ang = 2.0 * M_PI * (double)j / (double)n;
            c = cos(ang);
            s = sin(ang);
        }
        t[j] = (real_t)c + (real_t)(sign * s) * I;
    }
}

int o_fft_plan_init(o_fftplan* p, int n)
{
    memset(p, 0, sizeof(*p));
    if (n < 1) return -1;
    p->n = n;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nfac++] = 2; m /= 2; }
    while (m % 3 == 0) { p->fac[p->nfac++] = 3; m /= 3; }
    while (m % 5 == 0) { p->fac[p->nfac++] = 5; m /= 5; }
    for (int f = 7; m > 1; f += 2) {
        while (m % f == 0) { p->fac[p->nfac++] = f; m /= f; }
        if (f * f > m && m > 1) { p->fac[p->nfac++] = m; m = 1; }
    }
    if (p->nfac == 0) p->fac[p->nfac++] = 1;
    p->tw_fwd = (cplx_t*)malloc(sizeof(cplx_t) * n);
    p->tw_bwd = (cplx_t*)malloc(sizeof(cplx_t) * n);
    twiddle_table(p->tw_fwd, n, -1);
    twiddle_table(p->tw_bwd, n, +1);
    return 0;
}

void o_fft_plan_free(o_fftplan* p)
{
    free(p->tw_fwd);
    free(p->tw_bwd);
    memset(p, 0, sizeof(*p));
}

/* out[0..n) = DFT of in[0], in[is], in[2 is], ...  tw: table of N roots; ts = N/n */
static void rec(const cplx_t* in, cplx_t* out, int n, int is, const int* fac,
                const cplx_t* tw, int ts, int N, int sign)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = fac[0];
    int m = n / p;
    for (int q = 0; q < p; ++q) rec(in + (long)q * is, out + (long)q * m, m, is * p, fac + 1, tw, ts * p, N, sign);
    /* combine: X[k + m s] = sum_q w_n^{q (k + m s)} Y_q[k] */
    if (p == 2) {
        for (int k = 0; k < m; ++k) {
            cplx_t a = out[k];
            cplx_t b = out[k + m] * tw[(long)k * ts];
            out[k] = a + b;
            out[k + m] = a - b;
        }
    } else if (p == 4) {
        /* w_4^1 = e^{sign i pi/2} = sign*i */
        for (int k = 0; k < m; ++k) {
            cplx_t a0 = out[k];
            cplx_t a1 = out[k + m] * tw[(long)k * ts];
            cplx_t a2 = out[k + 2 * m] * tw[(long)2 * k * ts];
            cplx_t a3 = out[k + 3 * m] * tw[(long)3 * k * ts];
            cplx_t s02 = a0 + a2, d02 = a0 - a2;
            cplx_t s13 = a1 + a3, d13 = a1 - a3;
            cplx_t jd13 = (real_t)sign * I * d13;
            out[k] = s02 + s13;
            out[k + m] = d02 + jd13;
            out[k + 2 * m] = s02 - s13;
            out[k + 3 * m] = d02 - jd13;
        }
    } else {
        cplx_t* t = (cplx_t*)malloc(sizeof(cplx_t) * p);
        int pstep = N / p;   /* w_p^1 = tw[N/p] */
        for (int k = 0; k < m; ++k) {
            for (int q = 0; q < p; ++q) t[q] = out[k + q * m] * tw[((long)q * k * ts) % N];
            for (int s2 = 0; s2 < p; ++s2) {
                cplx_t acc = 0;
                for (int q = 0; q < p; ++q) acc += t[q] * tw[(long)((q * s2) % p) * pstep];
                out[k + s2 * m] = acc;
            }
        }
        free(t);
    }
}

void o_fft_exec(const o_fftplan* p, const cplx_t* in, cplx_t* out, int sign)
{
    rec(in, out, p->n, 1, p->fac, sign < 0 ? p->tw_fwd : p->tw_bwd, 1, p->n, sign < 0 ? -1 : 1);
}

int orc_fft(int n, const double* in, double* out, int sign)
{
    o_fftplan p;
    if (o_fft_plan_init(&p, n)) return -1;
    cplx_t* a = (cplx_t*)malloc(sizeof(cplx_t) * n);
    cplx_t* b = (cplx_t*)malloc(sizeof(cplx_t) * n);
    for (int i = 0; i < n; ++i) a[i] = (real_t)in[2 * i] + (real_t)in[2 * i + 1] * I;
    o_fft_exec(&p, a, b, sign);
    for (int i = 0; i < n; ++i) { out[2 * i] = creal(b[i]); out[2 * i + 1] = cimag(b[i]); }
    free(a);
    free(b);
    o_fft_plan_free(&p);
    return 0;
}
