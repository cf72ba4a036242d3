/* o_adjoint.c -- reverse-mode derivative of the fixed-step z-march of
 * o_march.c (test infrastructure only).
 *
 * The reference has no adjoint.  What is restated here is the exact transpose
 * of the forward scheme the oracle restates, with the step sizes held fixed:
 *   propagator.cu:629-783   calculate_NLresponse (FFTs, Kerr/Raman, plasma, loss)
 *   ionization.cpp:9-95     plasma Heun recurrence and the loss term
 *   ionization.cpp:98-147   photoionization table lookup (d/dI of the ln-ln interpolation)
 *   propagator.cu:376-425   classical RK4 without linear propagation between stages
 *   propagator.cu:570-625   calculate_Lstep
 *   fhatha.cpp:218-230      QDHT run_many (its transpose diag(1/m1) C^T diag(m1))
 * Pinned by central finite differences of the forward restatement
 * (tests/test_oracle_adjoint.py); the derivative is taken almost everywhere:
 * the jumps of the forward (the loss gate I > 29.93, W = 0 below 1e13 W/m^2 or
 * above N, the clamp rho <= N) contribute nothing.
 *
 * Cotangent convention for a complex value z of a real objective L:
 * g = dL/dRe z + i dL/dIm z (so L = |z|^2 has g = 2 z).  A complex-linear map
 * y = M x has g_x = M^H g_y.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>
#include "o_internal.h"

static void put_c(double* out, const cplx_t* a, size_t n)
{
    for (size_t i = 0; i < n; ++i) { out[2 * i] = creal(a[i]); out[2 * i + 1] = cimag(a[i]); }
}
static void get_c(cplx_t* a, const double* in, size_t n)
{
    for (size_t i = 0; i < n; ++i) a[i] = (real_t)in[2 * i] + (real_t)in[2 * i + 1] * I;
}

/* W(I) of photoionization_function (ionization.cpp:98-147, rho <= N) and dW/dI:
 * ln W is linear in ln I between table nodes, so dW/dI = W (b - a) / (step I). */
static double o_photo_d(orc_sim* s, double Iv, double* dW)
{
    *dW = 0;
    if (Iv < (O_ION_MIN_I / O_INTENSITY_DENOM)) return 0;
    double lnI = log(Iv) + O_INTENSITY_DENOM_LN;
    if (isnan(lnI)) o_throw(s, "float_type photoionization_function(...): Intensity is nan!");
    int n = (int)floor((lnI - O_ION_MIN_LN) / O_ION_LN_STEP);
    if (n > O_ION_N - 2) o_throw(s, "Attention!!!! Intensity is too big! Keep away, blast is possible.");
    if (n < 0) return 0;
    double lnIn = O_ION_MIN_LN + n * O_ION_LN_STEP;
    double a = s->IONIZATION_RATE_LN[n], b = s->IONIZATION_RATE_LN[n + 1];
    double W = exp(a + (lnI - lnIn) * (b - a) / O_ION_LN_STEP);
    *dW = W * (b - a) / O_ION_LN_STEP / Iv;
    return W;
}

typedef struct {
    real_t *Iv, *W, *dW, *c, *dc, *rho, *lf, *dlf, *y, *gI, *gW, *gc, *grho, *gy;
    cplx_t *E, *lt, *lr, *ly, *gE, *tmp, *tmp2;
} vjp_buf;

static void vjp_alloc(vjp_buf* v, int nt)
{
    real_t** r[] = {&v->Iv, &v->W, &v->dW, &v->c, &v->dc, &v->rho, &v->lf, &v->dlf, &v->y, &v->gI, &v->gW, &v->gc, &v->grho, &v->gy};
    for (unsigned i = 0; i < sizeof r / sizeof r[0]; ++i) *r[i] = (real_t*)calloc(nt, sizeof(real_t));
    cplx_t** c[] = {&v->E, &v->lt, &v->lr, &v->ly, &v->gE, &v->tmp, &v->tmp2};
    for (unsigned i = 0; i < sizeof c / sizeof c[0]; ++i) *c[i] = (cplx_t*)calloc(nt, sizeof(cplx_t));
}

static void vjp_free(vjp_buf* v)
{
    real_t* r[] = {v->Iv, v->W, v->dW, v->c, v->dc, v->rho, v->lf, v->dlf, v->y, v->gI, v->gW, v->gc, v->grho, v->gy};
    for (unsigned i = 0; i < sizeof r / sizeof r[0]; ++i) free(r[i]);
    cplx_t* c[] = {v->E, v->lt, v->lr, v->ly, v->gE, v->tmp, v->tmp2};
    for (unsigned i = 0; i < sizeof c / sizeof c[0]; ++i) free(c[i]);
}

/* Adjoint of o_nl_response (propagator.cu:629-783) at the spectrum `in`:
 * gacc += J^H lam, J the Jacobian of in -> N(in). */
static void nl_response_vjp(orc_sim* s, vjp_buf* v, const cplx_t* in, const cplx_t* lam, cplx_t* gacc)
{
    const int T = s->NT;
    const real_t Nd = (real_t)s->NEUTRAL_DENSITY, U = (real_t)s->IONIZATION_POTENTIAL;
    const real_t pc = (real_t)s->PONDEROMOTIVE_COEFFICIENT, sig = (real_t)s->AVALANCHE_CROSSSECTION;
    const real_t itau = (real_t)(1.0 / s->RECOMBINATION_TAU);
    const real_t dt = (real_t)(s->TSTEP * O_IONRATE_DENOM);
    const int kerr = s->NONLIN_REFRINDEX != 0, raman = s->RAMAN_FRACTION > 0.001;
    if (s->AMBIENT_CARRIER_DENSITY != 0) o_throw(s, "adjoint: AMBIENT_CARRIERS must be 0 (the forward draws rho0 at random)");

    /* ---- forward quantities */
    o_fft_exec(&s->fft, in, v->E, +1);
    for (int n = 0; n < T; ++n) v->E[n] /= (real_t)T;
    o_plasma_line(s, v->E, v->rho, 1, NULL);
    for (int n = 0; n < T; ++n) {
        const real_t re = creal(v->E[n]), im = cimag(v->E[n]);
        const real_t Iv = re * re + im * im;
        double dW;
        v->Iv[n] = Iv;
        v->W[n] = (real_t)o_photo_d(s, (double)re * re + (double)im * im, &dW);
        v->dW[n] = (real_t)dW;
        v->c[n] = sig / (U + pc * Iv) * Iv;
        v->dc[n] = sig * U / ((U + pc * Iv) * (U + pc * Iv));
        /* loss[n] = -W (U + p I)/I E for 1 <= n <= T-2 above the ln-threshold (ionization.cpp:43) */
        if (n >= 1 && n <= T - 2 && Iv > (real_t)O_ION_MIN_LN) {
            v->lf[n] = -(real_t)0.5 * v->W[n] * (U + pc * Iv) / Iv;
            v->dlf[n] = -(real_t)0.5 * (v->dW[n] * (U + pc * Iv) / Iv - v->W[n] * U / (Iv * Iv));
        } else {
            v->lf[n] = 0;
            v->dlf[n] = 0;
        }
    }
    if (kerr) {
        for (int n = 0; n < T; ++n) v->tmp[n] = v->Iv[n];
        if (raman) {
            o_fft_exec(&s->fft, v->tmp, v->tmp2, -1);
            for (int w = 0; w < T; ++w) v->tmp2[w] *= s->RAMAN_FUNCTION[w];
            o_fft_exec(&s->fft, v->tmp2, v->tmp, +1);
            for (int n = 0; n < T; ++n) v->tmp[n] /= (real_t)T;
        }
        for (int n = 0; n < T; ++n) v->y[n] = creal(v->tmp[n]);
    }

    /* ---- cotangents of the three time signals: N = F(loss) + PF F(rho E) - i K F(y E) */
    o_fft_exec(&s->fft, lam, v->lt, +1);
    for (int w = 0; w < T; ++w) v->tmp[w] = conj(s->PLASMA_FUNC[w]) * lam[w];
    o_fft_exec(&s->fft, v->tmp, v->lr, +1);
    if (kerr) {
        for (int w = 0; w < T; ++w) v->tmp[w] = I * s->KERR_PROFILE[w] * lam[w];
        o_fft_exec(&s->fft, v->tmp, v->ly, +1);
    }
    for (int n = 0; n < T; ++n) {
        const cplx_t e = v->E[n];
        cplx_t g = v->lf[n] * v->lt[n] + v->rho[n] * v->lr[n];
        v->gI[n] = creal(conj(e) * v->lt[n]) * v->dlf[n];
        v->grho[n] = creal(conj(e) * v->lr[n]);
        if (kerr) {
            g += v->y[n] * v->ly[n];
            v->gy[n] = creal(conj(e) * v->ly[n]);
        }
        v->gE[n] = g;
    }
    /* y = IFFT(FFT(I) RF)/T is a real circular convolution: its transpose is
     * the correlation, IFFT(FFT(gy) conj(RF))/T */
    if (kerr) {
        if (raman) {
            for (int n = 0; n < T; ++n) v->tmp[n] = v->gy[n];
            o_fft_exec(&s->fft, v->tmp, v->tmp2, -1);
            for (int w = 0; w < T; ++w) v->tmp2[w] *= conj(s->RAMAN_FUNCTION[w]);
            o_fft_exec(&s->fft, v->tmp2, v->tmp, +1);
            for (int n = 0; n < T; ++n) v->gI[n] += creal(v->tmp[n]) / (real_t)T;
        } else {
            for (int n = 0; n < T; ++n) v->gI[n] += v->gy[n];
        }
    }

    /* ---- plasma: reverse sweep over the Heun steps n = T-1 .. 1 (ionization.cpp:24-41).
     * Step n maps rho[n-1] with the rates of samples n-1 (zero for n = 1: E0 = 0)
     * and n to rho[n]; mu is the cotangent of rho[n] including later steps. */
    for (int n = 0; n < T; ++n) { v->gW[n] = 0; v->gc[n] = 0; }
    real_t carry = 0;
    for (int n = T - 1; n >= 1; --n) {
        const real_t mu = v->grho[n] + carry;
        const real_t x = v->rho[n - 1];
        const real_t Wp = n == 1 ? 0 : v->W[n - 1], cp = n == 1 ? 0 : v->c[n - 1];
        const real_t Wn = v->W[n], cn = v->c[n];
        const real_t W01 = x > Nd ? 0 : Wp;
        const real_t av0 = (x < 0 || x > Nd) ? 0 : cp * x * (1 - x / Nd);
        const real_t re0 = x > 0 ? x * itau : 0;
        const real_t k1 = dt * (W01 + av0 - re0);
        const real_t xm = x + k1;
        const real_t W11 = xm > Nd ? 0 : Wn;
        const real_t av1 = (xm < 0 || xm > Nd) ? 0 : cn * xm * (1 - xm / Nd);
        const real_t re1 = xm > 0 ? xm * itau : 0;
        const real_t k2 = dt * (W11 + av1 - re1);
        if (x + (real_t)0.5 * (k1 + k2) > Nd) { carry = 0; continue; }   /* clamped: rho[n] = N */
        const real_t dk1 = dt * (((x < 0 || x > Nd) ? 0 : cp * (1 - 2 * x / Nd)) - (x >= 0 ? itau : 0));
        const real_t q = dt * (((xm < 0 || xm > Nd) ? 0 : cn * (1 - 2 * xm / Nd)) - (xm >= 0 ? itau : 0));
        const real_t h = (real_t)0.5 * mu;
        v->gW[n] += h * dt * (xm <= Nd ? 1 : 0);
        v->gc[n] += h * dt * ((xm < 0 || xm > Nd) ? 0 : xm * (1 - xm / Nd));
        if (n >= 2) {
            v->gW[n - 1] += h * dt * (x <= Nd ? 1 : 0) * (1 + q);
            v->gc[n - 1] += h * dt * ((x < 0 || x > Nd) ? 0 : x * (1 - x / Nd)) * (1 + q);
        }
        carry = mu * (1 + (real_t)0.5 * (dk1 + q * (1 + dk1)));
    }
    for (int n = 0; n < T; ++n) {
        v->gI[n] += v->gW[n] * v->dW[n] + v->gc[n] * v->dc[n];
        v->gE[n] += 2 * v->gI[n] * v->E[n];
    }
    /* E = IFFT(in)/T: g_in = FFT(gE)/T */
    o_fft_exec(&s->fft, v->gE, v->tmp, -1);
    for (int w = 0; w < T; ++w) gacc[w] += v->tmp[w] / (real_t)T;
}

/* Adjoint of one line of the RK4 NL step (propagator.cu:376-425):
 * B1 = (k1 + 2 k2 + 2 k3 + k4)/6, k_j = N(X_j), X1 = A, X2 = A + dz/2 k1,
 * X3 = A + dz/2 k2, X4 = A + dz k3.  gA = J^H g. */
static void rk4_vjp(orc_sim* s, o_linebuf* b, vjp_buf* v, const cplx_t* A, const cplx_t* g, real_t dz, cplx_t* gA,
                    cplx_t* X2, cplx_t* X3, cplx_t* X4, cplx_t* lk, cplx_t* gx)
{
    const int T = s->NT;
    o_nl_response(s, b, A, b->out1);
    for (int t = 0; t < T; ++t) X2[t] = A[t] + dz * (real_t)0.5 * b->out1[t];
    o_nl_response(s, b, X2, b->out2);
    for (int t = 0; t < T; ++t) X3[t] = A[t] + dz * (real_t)0.5 * b->out2[t];
    o_nl_response(s, b, X3, b->out3);
    for (int t = 0; t < T; ++t) X4[t] = A[t] + dz * b->out3[t];

    for (int t = 0; t < T; ++t) { lk[t] = g[t] / (real_t)6; gx[t] = 0; gA[t] = 0; }
    nl_response_vjp(s, v, X4, lk, gx);
    for (int t = 0; t < T; ++t) { gA[t] += gx[t]; lk[t] = g[t] / (real_t)3 + dz * gx[t]; gx[t] = 0; }
    nl_response_vjp(s, v, X3, lk, gx);
    for (int t = 0; t < T; ++t) { gA[t] += gx[t]; lk[t] = g[t] / (real_t)3 + dz * (real_t)0.5 * gx[t]; gx[t] = 0; }
    nl_response_vjp(s, v, X2, lk, gx);
    for (int t = 0; t < T; ++t) { gA[t] += gx[t]; lk[t] = g[t] / (real_t)6 + dz * (real_t)0.5 * gx[t]; gx[t] = 0; }
    nl_response_vjp(s, v, A, lk, gx);
    for (int t = 0; t < T; ++t) gA[t] += gx[t];
}

/* ---- line-parallel RK4 adjoint over the R lines of [R][T] arrays */
typedef struct {
    const cplx_t* A;
    const cplx_t* g;
    cplx_t* gA;
    real_t dz;
} rk4_ctx;

static void rk4_vjp_lines(orc_sim* s, int tid, int r0, int r1, void* vctx)
{
    rk4_ctx* c = (rk4_ctx*)vctx;
    const int T = s->NT;
    vjp_buf v;
    vjp_alloc(&v, T);
    cplx_t* w = (cplx_t*)calloc((size_t)5 * T, sizeof(cplx_t));
    for (int r = r0; r < r1; ++r) {
        const size_t o = (size_t)r * T;
        rk4_vjp(s, &s->lb[tid], &v, c->A + o, c->g + o, c->dz, c->gA + o, w, w + T, w + 2 * T, w + 3 * T, w + 4 * T);
    }
    free(w);
    vjp_free(&v);
}

/* transpose of o_qdht_run_many: data <- diag(1/m1) C^T diag(m1) data per column */
static void qdht_t_cols(orc_sim* s, int tid, int c0, int c1, void* vctx)
{
    (void)tid;
    cplx_t* data = (cplx_t*)vctx;
    const int n = s->NR, T = s->NT;
    cplx_t* buf = (cplx_t*)malloc(sizeof(cplx_t) * n);
    for (int c = c0; c < c1; ++c) {
        for (int j = 0; j < n; ++j) data[(size_t)j * T + c] *= s->qm1_r[j];
        for (int i = 0; i < n; ++i) {
            cplx_t acc = 0;
            for (int j = 0; j < n; ++j) acc += s->qC_r[j + (size_t)n * i] * data[(size_t)j * T + c];
            buf[i] = acc;
        }
        for (int i = 0; i < n; ++i) data[(size_t)i * T + c] = buf[i] / s->qm1_r[i];
    }
    free(buf);
}

static void qdht_t(orc_sim* s, cplx_t* data) { o_parallel(s, s->NT, qdht_t_cols, data); }

/* transpose of lstep_rows (propagator.cu:570-625) at step dz: with gfk the
 * cotangent of FIELD_k and gb2 that of BIG2 after the step, gb2 becomes the
 * cotangent of BIG2 before the step and gb1 that of BIG1 (k-space). */
static void lstep_t(orc_sim* s, real_t dz, const cplx_t* gfk, cplx_t* gb2, cplx_t* gb1)
{
    const int T = s->NT;
    for (int nx = 0; nx < s->NR; ++nx) {
        const real_t kt = (real_t)(2 * M_PI * s->qx[nx] * s->qV / s->XMAX);
        const real_t kt2 = kt * kt;
        for (int nw = 0; nw < T; ++nw) {
            const size_t o = (size_t)nx * T + nw;
            const real_t w = s->OMEGA[nw];
            const cplx_t k = s->WAVENUMBER[nw];
            gb1[o] = 0;
            /* masked points: FIELD = 0, BIG2 untouched */
            if (w < (real_t)s->OMEGA_MIN || w > (real_t)s->OMEGA_MAX) continue;
            if (kt2 > (real_t)O_MAX_KT2 * creal(k * k)) continue;
            const real_t k1 = creal(k), k2 = cimag(k);
            const real_t sq = o_sqrtho(-kt2 / k1 / k1);
            const cplx_t P = exp(dz * (I * (-s->HO[nw] - sq * k1) + k2));
            const cplx_t gpre = conj(P) * (gb2[o] + gfk[o]);
            gb2[o] = gpre;
            gb1[o] = dz / (1 + sq) * gpre;
        }
    }
}

/* Fixed-step march from the current FIELD / BIG2 and its adjoint: n steps of
 * (NL RK4 at dz_rk[k], QDHT, Lstep at dz_lin[k], QDHT); then, from the
 * cotangents of the final FIELD (gfield) and BIG2 (gb2, may be NULL), the
 * cotangents of the initial FIELD and BIG2.  The state is left at the end of
 * the march.  [R][T] complex interleaved arrays. */
int orc_adjoint_march(orc_sim* s, int n, const double* dz_rk, const double* dz_lin, const double* gfield,
                      const double* gb2, double* gfield0, double* gb2_0)
{
    const size_t RT = (size_t)s->NR * s->NT;
    cplx_t* ck = (cplx_t*)malloc(sizeof(cplx_t) * RT * (size_t)(n > 0 ? n : 1));
    cplx_t* gA = (cplx_t*)calloc(RT, sizeof(cplx_t));
    cplx_t* gB = (cplx_t*)calloc(RT, sizeof(cplx_t));
    cplx_t* t1 = (cplx_t*)calloc(RT, sizeof(cplx_t));
    cplx_t* t2 = (cplx_t*)calloc(RT, sizeof(cplx_t));
    s->jb_armed = 1;
    if (setjmp(s->jb)) {
        s->jb_armed = 0;
        free(ck); free(gA); free(gB); free(t1); free(t2);
        return -1;
    }
    for (int k = 0; k < n; ++k) {
        memcpy(ck + (size_t)k * RT, s->FIELD, sizeof(cplx_t) * RT);
        o_fixed_step2(s, dz_rk[k], dz_lin[k]);
    }
    get_c(gA, gfield, RT);
    if (gb2) get_c(gB, gb2, RT);
    for (int k = n - 1; k >= 0; --k) {
        /* FIELD_next = Q FIELD_k */
        memcpy(t1, gA, sizeof(cplx_t) * RT);
        qdht_t(s, t1);
        lstep_t(s, (real_t)dz_lin[k], t1, gB, t2);
        /* BIG1_k = Q BIG1 */
        qdht_t(s, t2);
        rk4_ctx c = {ck + (size_t)k * RT, t2, gA, (real_t)dz_rk[k]};
        o_parallel(s, s->NR, rk4_vjp_lines, &c);
    }
    DISARM(s);
    put_c(gfield0, gA, RT);
    put_c(gb2_0, gB, RT);
    free(ck); free(gA); free(gB); free(t1); free(t2);
    return 0;
}

/* one line: gin = J^H lam of calculate_NLresponse at `in` (T complex each) */
int orc_nl_response_vjp(orc_sim* s, const double* in, const double* lam, double* gin)
{
    const int T = s->NT;
    cplx_t* a = (cplx_t*)malloc(sizeof(cplx_t) * T);
    cplx_t* l = (cplx_t*)malloc(sizeof(cplx_t) * T);
    cplx_t* g = (cplx_t*)calloc(T, sizeof(cplx_t));
    vjp_buf v;
    vjp_alloc(&v, T);
    get_c(a, in, T);
    get_c(l, lam, T);
    s->jb_armed = 1;
    if (setjmp(s->jb)) { s->jb_armed = 0; free(a); free(l); free(g); vjp_free(&v); return -1; }
    nl_response_vjp(s, &v, a, l, g);
    DISARM(s);
    put_c(gin, g, T);
    free(a); free(l); free(g);
    vjp_free(&v);
    return 0;
}

/* [R][T]: gA = J^H g of the RK4 NL step at FIELD = A, step dz */
int orc_nl_step_vjp(orc_sim* s, const double* A, const double* g, double dz, double* gA)
{
    const size_t RT = (size_t)s->NR * s->NT;
    cplx_t* a = (cplx_t*)malloc(sizeof(cplx_t) * RT);
    cplx_t* gg = (cplx_t*)malloc(sizeof(cplx_t) * RT);
    cplx_t* out = (cplx_t*)calloc(RT, sizeof(cplx_t));
    get_c(a, A, RT);
    get_c(gg, g, RT);
    s->jb_armed = 1;
    if (setjmp(s->jb)) { s->jb_armed = 0; free(a); free(gg); free(out); return -1; }
    rk4_ctx c = {a, gg, out, (real_t)dz};
    o_parallel(s, s->NR, rk4_vjp_lines, &c);
    DISARM(s);
    put_c(gA, out, RT);
    free(a); free(gg); free(out);
    return 0;
}
