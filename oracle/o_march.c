/* o_march.c -- the z-step hot path of the oracle (test infrastructure only).
 *
 * Restates, with NONLINEARITY_ON:
 *   propagator.cu:125-266   propagator(): NL step, MAX reductions, accept/reject,
 *                           dz update and clamp, QDHT -> Lstep -> QDHT
 *   propagator.cu:376-425   calculate_NLstep() CPU branch, RK4 per radial line
 *   propagator.cu:570-625   calculate_Lstep() CPU branch (_UNIAXIAL, sqrtHO)
 *   propagator.cu:629-783   calculate_NLresponse() (PLASMA_DISPERSION branch)
 *   ionization.cpp:9-95     calculate_plasmadensity_losses_small() single level
 *   ionization.cpp:98-147   photoionization_function() table lookup
 *   ionization.h:77-87      avalanche / recombination
 *   fhatha.cpp:218-230      qdht_plan::run_many()
 *   extmath.h:234-252       sqrtHO (20-term Taylor series of sqrt(1+x)-1)
 *   solver_main.cpp:202-225 main loop with dumps; dump_types.cpp writers
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>
#include "o_internal.h"

/* A throw inside a worker of o_parallel returns to that worker's own frame
 * (o_tl_jb, with the message in o_tl_err); o_parallel re-throws on the
 * calling thread after every worker has been joined. */
static __thread jmp_buf* o_tl_jb = NULL;
static __thread char* o_tl_err = NULL;

void o_throw(orc_sim* s, const char* msg)
{
    if (o_tl_jb) {
        snprintf(o_tl_err, sizeof s->err, "%s", msg);
        longjmp(*o_tl_jb, 1);
    }
    if (msg != s->err) snprintf(s->err, sizeof s->err, "%s", msg);
    if (s->jb_armed) longjmp(s->jb, 1);
    fprintf(stderr, "oracle error outside API frame: %s\n", msg);
    abort();
}

void o_linebuf_alloc(o_linebuf* b, int nt)
{
    cplx_t** p[] = {&b->field_real, &b->b1, &b->b2, &b->b3, &b->b4, &b->out1, &b->out2, &b->out3, &b->out4, &b->in};
    for (unsigned i = 0; i < sizeof p / sizeof p[0]; ++i) *p[i] = (cplx_t*)calloc(nt, sizeof(cplx_t));
}

void o_linebuf_free(o_linebuf* b)
{
    cplx_t* p[] = {b->field_real, b->b1, b->b2, b->b3, b->b4, b->out1, b->out2, b->out3, b->out4, b->in};
    for (unsigned i = 0; i < sizeof p / sizeof p[0]; ++i) free(p[i]);
}

/* ---- tiny parallel-for over [0, n) ---- */
typedef struct {
    orc_sim* s;
    int tid, nth, n;
    void (*fn)(orc_sim*, int tid, int i0, int i1, void* ctx);
    void* ctx;
    int failed;
    char err[512];
} o_job;

static void* o_job_run(void* a)
{
    o_job* j = (o_job*)a;
    int chunk = (j->n + j->nth - 1) / j->nth;
    int i0 = j->tid * chunk, i1 = i0 + chunk;
    if (i1 > j->n) i1 = j->n;
    jmp_buf jb;
    jmp_buf* prev_jb = o_tl_jb;
    char* prev_err = o_tl_err;
    o_tl_jb = &jb;
    o_tl_err = j->err;
    if (setjmp(jb) == 0) {
        if (i0 < i1) j->fn(j->s, j->tid, i0, i1, j->ctx);
    } else {
        j->failed = 1;
    }
    o_tl_jb = prev_jb;
    o_tl_err = prev_err;
    return NULL;
}

void o_parallel(orc_sim* s, int n, void (*fn)(orc_sim*, int, int, int, void*), void* ctx)
{
    int nth = s->nthreads;
    if (nth <= 1 || n < 2) { fn(s, 0, 0, n, ctx); return; }
    pthread_t th[256];
    static __thread o_job jobs[256];
    if (nth > 256) nth = 256;
    for (int t = 0; t < nth; ++t) {
        jobs[t] = (o_job){s, t, nth, n, fn, ctx, 0, ""};
        if (t) pthread_create(&th[t], NULL, o_job_run, &jobs[t]);
    }
    o_job_run(&jobs[0]);
    for (int t = 1; t < nth; ++t) pthread_join(th[t], NULL);
    for (int t = 0; t < nth; ++t)
        if (jobs[t].failed) o_throw(s, jobs[t].err);
}

/* ---- ionization ---- */
static inline double abs2d(double re, double im) { return re * re + im * im; }

double o_photoionization(orc_sim* s, double reA, double imA, double ro)
{
    double Iv = abs2d(reA, imA);
    if (Iv < (O_ION_MIN_I / O_INTENSITY_DENOM)) return 0;
    if (ro > s->NEUTRAL_DENSITY) return 0;
    double lnI = log(Iv) + O_INTENSITY_DENOM_LN;
    if (isnan(lnI)) o_throw(s, "float_type photoionization_function(...): Intensity is nan!");
    int n = (int)floor((lnI - O_ION_MIN_LN) / O_ION_LN_STEP);
    if (n > O_ION_N - 2) o_throw(s, "Attention!!!! Intensity is too big! Keep away, blast is possible.");
    if (n < 0) return 0;
    double lnIn = O_ION_MIN_LN + n * O_ION_LN_STEP;
    double a = s->IONIZATION_RATE_LN[n], b = s->IONIZATION_RATE_LN[n + 1];
    double lnW = a + (lnI - lnIn) * (b - a) / O_ION_LN_STEP;
    return exp(lnW);
}

static inline real_t avalanche(orc_sim* s, real_t reA, real_t imA, real_t ro)
{
    real_t N = (real_t)s->NEUTRAL_DENSITY;
    if (ro < 0 || ro > N) return 0;
    real_t Iv = reA * reA + imA * imA;
    return (real_t)s->AVALANCHE_CROSSSECTION / ((real_t)s->IONIZATION_POTENTIAL + (real_t)s->PONDEROMOTIVE_COEFFICIENT * Iv) * Iv * ro * (1 - ro / N);
}

static inline real_t recombination(orc_sim* s, real_t ro)
{
    return ro > 0 ? ro / (real_t)s->RECOMBINATION_TAU : 0;
}

/* ionization.cpp:9-95, single-level branch.  pro[nt*stride] = rho(t_nt). */
void o_plasma_line(orc_sim* s, const cplx_t* field, real_t* pro, int stride, cplx_t* loss)
{
    int nt_ = s->NT;
    real_t tstep = (real_t)(s->TSTEP * O_IONRATE_DENOM);
    if (loss) loss[0] = 0;
    real_t ro1 = (real_t)(s->AMBIENT_CARRIER_DENSITY * 2.0 * (s->AMBIENT_CARRIER_DENSITY != 0 ? drand48() : 0.0));
    real_t k1, k2;
    if (pro) pro[0] = ro1;
    real_t reE0 = 0, imE0 = 0, I0 = 0;
    real_t N = (real_t)s->NEUTRAL_DENSITY;
    real_t U = (real_t)s->IONIZATION_POTENTIAL, pc = (real_t)s->PONDEROMOTIVE_COEFFICIENT;
    for (int nt = 1; nt < nt_; ++nt) {
        double ph = -s->OMEGA0 * (s->TMIN + nt * (s->TMAX - s->TMIN) / nt_);
        cplx_t pf = (cplx_t)cexp(I * ph);
        cplx_t e1 = field[nt] * pf;
        real_t reE1 = creal(e1), imE1 = cimag(e1);
        real_t I1 = reE1 * reE1 + imE1 * imE1;
        real_t W01 = (real_t)o_photoionization(s, reE0, imE0, ro1);
        k1 = tstep * (W01 + avalanche(s, reE0, imE0, ro1) - recombination(s, ro1));
        real_t W11 = (real_t)o_photoionization(s, reE1, imE1, ro1 + k1);
        k2 = tstep * (W11 + avalanche(s, reE1, imE1, ro1 + k1) - recombination(s, ro1 + k1));
        ro1 += (real_t)0.5 * (k1 + k2);
        if (ro1 > N) ro1 = N;
        if (pro) pro[(size_t)nt * stride] = ro1;
        if (loss) loss[nt - 1] = (I0 > (real_t)O_ION_MIN_LN) ? (-(real_t)0.5 * W01 * (U + pc * I0) / I0 * field[nt - 1]) : 0;
        reE0 = reE1;
        imE0 = imE1;
        I0 = I1;
    }
    if (loss) loss[nt_ - 1] = 0;
}

/* propagator.cu:629-783 */
void o_nl_response(orc_sim* s, o_linebuf* b, const cplx_t* in, cplx_t* out)
{
    int nt_ = s->NT;
    o_fft_exec(&s->fft, in, b->field_real, +1);
    for (int i = 0; i < nt_; ++i) b->field_real[i] /= (real_t)nt_;
    /* plasma density -> re(b3) (stride 2), loss -> b1 */
    o_plasma_line(s, b->field_real, (real_t*)b->b3, 2, b->b1);
    for (int i = 0; i < nt_; ++i) {
        real_t ro = creal(b->b3[i]);
        b->b3[i] = ro * b->field_real[i];
    }
    o_fft_exec(&s->fft, b->b1, b->b2, -1);
    o_fft_exec(&s->fft, b->b3, b->b1, -1);
    for (int w = 0; w < nt_; ++w) {
        b->b1[w] *= s->PLASMA_FUNC[w];
        b->b2[w] += b->b1[w];
    }
    if (s->NONLIN_REFRINDEX != 0) {
        for (int i = 0; i < nt_; ++i) {
            cplx_t e = b->field_real[i];
            b->b3[i] = creal(e) * creal(e) + cimag(e) * cimag(e);
        }
        if (s->RAMAN_FRACTION > 0.001) {
            o_fft_exec(&s->fft, b->b3, b->b1, -1);
            for (int w = 0; w < nt_; ++w) b->b1[w] *= s->RAMAN_FUNCTION[w];
            o_fft_exec(&s->fft, b->b1, b->b3, +1);
            for (int i = 0; i < nt_; ++i) b->b3[i] /= (real_t)nt_;
        }
        for (int i = 0; i < nt_; ++i) b->b3[i] *= b->field_real[i];
        o_fft_exec(&s->fft, b->b3, b->b1, -1);
        for (int w = 0; w < nt_; ++w) b->b2[w] += -I * s->KERR_PROFILE[w] * b->b1[w];
    }
    for (int w = 0; w < nt_; ++w) out[w] = b->b2[w];
}

/* ---- NL step (propagator.cu:376-425) ---- */
typedef struct { double maxI[256], maxNL[256]; } nl_ctx;

static void nl_lines(orc_sim* s, int tid, int r0, int r1, void* vctx)
{
    nl_ctx* c = (nl_ctx*)vctx;
    o_linebuf* b = &s->lb[tid];
    int nt_ = s->NT;
    real_t dz = (real_t)s->ZSTEP;
    real_t mI = 0, mNL = 0;
    for (int r = r0; r < r1; ++r) {
        cplx_t* F = s->FIELD + (size_t)r * nt_;
        cplx_t* B = s->BIG1 + (size_t)r * nt_;
        o_nl_response(s, b, F, b->out1);
        for (int t = 0; t < nt_; ++t) b->in[t] = F[t] + dz * (real_t)0.5 * b->out1[t];
        o_nl_response(s, b, b->in, b->out2);
        for (int t = 0; t < nt_; ++t) b->in[t] = F[t] + dz * (real_t)0.5 * b->out2[t];
        o_nl_response(s, b, b->in, b->out3);
        for (int t = 0; t < nt_; ++t) b->in[t] = F[t] + dz * b->out3[t];
        o_nl_response(s, b, b->in, b->out4);
        for (int t = 0; t < nt_; ++t) {
            B[t] = (b->out1[t] + (real_t)2.0 * b->out2[t] + (real_t)2.0 * b->out3[t] + b->out4[t]) / (real_t)6.0;
            real_t a = creal(B[t]) * creal(B[t]) + cimag(B[t]) * cimag(B[t]);
            real_t f = creal(F[t]) * creal(F[t]) + cimag(F[t]) * cimag(F[t]);
            if (a > mNL) mNL = a;
            if (f > mI) mI = f;
        }
    }
    c->maxI[tid] = mI;
    c->maxNL[tid] = mNL;
}

static void nl_step(orc_sim* s, double* maxA2, double* maxNL2)
{
    nl_ctx c;
    memset(&c, 0, sizeof c);
    o_parallel(s, s->NR, nl_lines, &c);
    double a = 0, b = 0;
    for (int t = 0; t < s->nthreads && t < 256; ++t) {
        if (c.maxI[t] > a) a = c.maxI[t];
        if (c.maxNL[t] > b) b = c.maxNL[t];
    }
    *maxA2 = a;
    *maxNL2 = b;
}

/* ---- QDHT run_many over the N_T columns (fhatha.cpp:218-230) ---- */
static void qdht_cols(orc_sim* s, int tid, int c0, int c1, void* vctx)
{
    (void)tid;
    cplx_t* data = (cplx_t*)vctx;
    int n = s->NR, nt_ = s->NT;
    cplx_t* buf = (cplx_t*)malloc(sizeof(cplx_t) * n);
    for (int c = c0; c < c1; ++c) {
        for (int i = 0; i < n; ++i) data[(size_t)i * nt_ + c] /= s->qm1_r[i];
        for (int i = 0; i < n; ++i) {
            cplx_t acc = 0;
            for (int j = 0; j < n; ++j) acc += s->qC_r[i + (size_t)n * j] * data[(size_t)j * nt_ + c];
            buf[i] = acc;
        }
        for (int i = 0; i < n; ++i) data[(size_t)i * nt_ + c] = buf[i] * s->qm1_r[i];
    }
    free(buf);
}

void o_qdht_run_many(orc_sim* s, cplx_t* data)
{
    o_parallel(s, s->NT, qdht_cols, data);
}

real_t o_sqrtho(real_t x)
{
    static const double k[20] = {0.5, -0.125, 0.0625, -0.0390625, 0.02734375, -0.0205078125, 0.01611328125,
                                 -0.013092041015625, 0.010910034179688, -0.009273529052734, 0.008008956909180,
                                 -0.007007837295532, 0.006199240684509, -0.005535036325455, 0.004981532692909,
                                 -0.004514514002949, 0.004116174532101, -0.003773159987759, 0.003475278936094,
                                 -0.003214633015887};
    real_t r = 0, xn = x;
    for (int j = 0; j < 20; ++j) { r += (real_t)k[j] * xn; xn *= x; }
    return r;
}

/* ---- linear step in (k_perp, omega), propagator.cu:570-625 ---- */
static void lstep_rows(orc_sim* s, int tid, int r0, int r1, void* vctx)
{
    (void)tid; (void)vctx;
    int nt_ = s->NT;
    real_t dz = (real_t)s->ZSTEP;
    for (int nx = r0; nx < r1; ++nx) {
        real_t kt = (real_t)(2 * M_PI * s->qx[nx] * s->qV / s->XMAX);
        real_t kt2 = kt * kt;
        for (int nw = 0; nw < nt_; ++nw) {
            size_t o = (size_t)nx * nt_ + nw;
            real_t w = s->OMEGA[nw];
            cplx_t k = s->WAVENUMBER[nw];
            if (w < (real_t)s->OMEGA_MIN || w > (real_t)s->OMEGA_MAX) { s->FIELD[o] = 0; continue; }
            if (kt2 > (real_t)O_MAX_KT2 * creal(k * k)) { s->FIELD[o] = 0; continue; }
            real_t k1 = creal(k), k2 = cimag(k);
            real_t sq = o_sqrtho(-kt2 / k1 / k1);
            s->BIG2[o] += dz * s->BIG1[o] / (1 + sq);
            s->FIELD[o] = s->BIG2[o] * exp(dz * (I * (-s->HO[nw] - sq * k1) + k2));
            s->BIG2[o] = s->FIELD[o];
        }
    }
}

static void lstep(orc_sim* s) { o_parallel(s, s->NR, lstep_rows, NULL); }

/* propagator.cu:125-266 without apertures: loops until accepted */
static void propagator(orc_sim* s)
{
    for (;;) {
        double maxA2, maxNL2;
        nl_step(s, &maxA2, &maxNL2);
        int repeat = 0;
        if (s->ZSTEP * s->ZSTEP * maxNL2 > s->MAX_TOLERANCE * s->MAX_TOLERANCE * maxA2) { repeat = 1; s->STEP_N++; s->n_reject++; }
        s->ZSTEP = 0.5 * s->MAX_TOLERANCE * sqrt(maxA2 / maxNL2);
        if (s->ZSTEP < O_MINSTEP_RATIO * (s->ZNET[s->n_Z] - s->ZNET[s->n_Z - 1]) || !(s->ZSTEP > 0))
            o_throw(s, "Minimum Z step size reached, collapse is possible!");
        double lim = s->ZNET[s->n_Z] - s->CURRENT_Z;
        if (lim < s->ZSTEP) s->ZSTEP = lim;
        if (!repeat) {
            o_qdht_run_many(s, s->BIG1);
            lstep(s);
            o_qdht_run_many(s, s->FIELD);
            return;
        }
    }
}

/* ---- dumps (dump_types.cpp) ---- */
typedef struct {
    int kind;   /* 0 full 1 maxI 2 flux 3 plasma_max 4 field_axis 5 average_spectrum 6 plasma_axis */
    FILE* f;
    int tN, xN, yN, zN, iscomplex, nz_denom;
    double* wfilter;
} o_dump;

static const char* dump_names[] = {"full", "maxI", "flux", "plasma_max", "field_axis", "average_spectrum", "plasma_axis"};

static void udf_header(o_dump* d)
{
    const char tag[] = "UDF1.0";
    int hdr[6] = {(int)sizeof(double), d->iscomplex, d->tN, d->xN, d->yN, d->zN};
    fwrite(tag, 1, sizeof tag, d->f);
    fwrite(hdr, sizeof(int), 6, d->f);
    fflush(d->f);
}

static void udf_write(o_dump* d, long ofs, const double* buf, size_t nelem)
{
    long long off = 31 + (long long)ofs * (d->iscomplex ? 16 : 8);
    fseeko(d->f, off, SEEK_SET);
    fwrite(buf, d->iscomplex ? 16 : 8, nelem, d->f);
}

static void field_time(orc_sim* s, int r, const double* wf, cplx_t* tmp, cplx_t* out)
{
    int nt_ = s->NT;
    for (int w = 0; w < nt_; ++w) tmp[w] = s->FIELD[(size_t)r * nt_ + w] * (real_t)(wf ? wf[w] : 1.0) / (real_t)nt_;
    o_fft_exec(&s->fft, tmp, out, +1);
}

static void do_dump(orc_sim* s, o_dump* d)
{
    int nt_ = s->NT, nr = s->NR;
    o_linebuf* b = &s->lb[0];
    switch (d->kind) {
    case 0: { /* full, dump_types.cpp:93-98 */
        if (s->n_Z % d->nz_denom) return;
        double* buf = (double*)malloc(sizeof(double) * 2 * nt_);
        for (int r = 0; r < nr; ++r) {
            o_fft_exec(&s->fft, s->FIELD + (size_t)r * nt_, b->b1, +1);
            for (int t = 0; t < nt_; ++t) {
                cplx_t v = b->b1[t] / (real_t)nt_;
                buf[2 * t] = creal(v);
                buf[2 * t + 1] = cimag(v);
            }
            udf_write(d, (long)d->tN * ((long)d->xN * d->yN * (s->n_Z / d->nz_denom) + r), buf, nt_);
        }
        free(buf);
        break;
    }
    case 1: case 2: { /* maxI / flux, dump_types.cpp:123-165 */
        double* buf = (double*)malloc(sizeof(double) * nr);
        for (int r = 0; r < nr; ++r) {
            field_time(s, r, d->wfilter, b->b2, b->b1);
            double acc = 0;
            for (int t = 0; t < nt_; ++t) {
                double a = creal(b->b1[t]) * creal(b->b1[t]) + cimag(b->b1[t]) * cimag(b->b1[t]);
                if (d->kind == 1) { if (a > acc) acc = a; }
                else acc += a;
            }
            buf[r] = d->kind == 1 ? acc : acc * s->TSTEP;
        }
        udf_write(d, (long)d->xN * (d->yN * s->n_Z), buf, nr);
        free(buf);
        break;
    }
    case 3: { /* plasma_max, dump_types.cpp:527-555 + propagator.cu:935-946 */
        double* buf = (double*)malloc(sizeof(double) * nr);
        real_t* ro = (real_t*)malloc(sizeof(real_t) * nt_);
        for (int r = 0; r < nr; ++r) {
            o_fft_exec(&s->fft, s->FIELD + (size_t)r * nt_, b->b1, +1);
            for (int t = 0; t < nt_; ++t) b->b1[t] /= (real_t)nt_;
            o_plasma_line(s, b->b1, ro, 1, NULL);
            double m = 0;
            for (int t = 0; t < nt_; ++t) if (ro[t] > m) m = ro[t];
            buf[r] = m;
        }
        udf_write(d, (long)d->xN * (d->yN * s->n_Z), buf, nr);
        free(ro);
        free(buf);
        break;
    }
    case 4: { /* field_axis, dump_types.cpp:405-419 */
        double* buf = (double*)malloc(sizeof(double) * 2 * nt_);
        o_fft_exec(&s->fft, s->FIELD, b->b1, +1);
        for (int t = 0; t < nt_; ++t) {
            cplx_t v = b->b1[t] / (real_t)nt_;
            buf[2 * t] = creal(v);
            buf[2 * t + 1] = cimag(v);
        }
        udf_write(d, (long)d->tN * s->n_Z, buf, nt_);
        free(buf);
        break;
    }
    case 5: { /* average_spectrum, dump_types.cpp:429-457 (_UNIAXIAL, one process) */
        double* buf = (double*)malloc(sizeof(double) * nt_);
        for (int w = 0; w < nt_; ++w) {
            double x1 = s->qx[0], x2 = s->qx[1];
            cplx_t v = s->FIELD[w];
            double f1 = creal(v) * creal(v) + cimag(v) * cimag(v);
            double acc = f1 * x1 * (x2 - x1);
            for (int cx = 1; cx < nr; ++cx) {
                x2 = s->qx[cx];
                v = s->FIELD[(size_t)cx * nt_ + w];
                f1 = creal(v) * creal(v) + cimag(v) * cimag(v);
                acc += f1 * x2 * (x2 - x1);
                x1 = x2;
            }
            buf[w] = acc * s->XMAX * s->XMAX / 2.0;
        }
        udf_write(d, (long)d->tN * s->n_Z, buf, nt_);
        free(buf);
        break;
    }
    case 6: { /* plasma_axis, dump_types.cpp:469-482 */
        double* buf = (double*)malloc(sizeof(double) * nt_);
        real_t* ro = (real_t*)malloc(sizeof(real_t) * nt_);
        o_fft_exec(&s->fft, s->FIELD, b->b1, +1);
        for (int t = 0; t < nt_; ++t) b->b1[t] /= (real_t)nt_;
        o_plasma_line(s, b->b1, ro, 1, NULL);
        for (int t = 0; t < nt_; ++t) buf[t] = ro[t];
        udf_write(d, (long)d->tN * s->n_Z, buf, nt_);
        free(ro);
        free(buf);
        break;
    }
    }
}

static void make_filter(orc_sim* s, const char* name, double* wf)
{
    int nt_ = s->NT;
    if (strncmp(name, "longpass", 8) == 0 || strncmp(name, "lp", 2) == 0 || strncmp(name, "shortpass", 9) == 0 || strncmp(name, "sp", 2) == 0) {
        int lp = (name[0] == 'l');
        const char* num = name + (strncmp(name, "longpass", 8) == 0 ? 8 : strncmp(name, "shortpass", 9) == 0 ? 9 : 2);
        double w = 2 * M_PI * O_LIGHT_VELOCITY / atof(num);
        for (int i = 0; i < nt_; ++i) wf[i] = lp ? (s->OMEGA_d[i] < w ? 1.0 : 0.0) : (s->OMEGA_d[i] > w ? 1.0 : 0.0);
    } else if (strncmp(name, "NO", 2) == 0) {
        for (int i = 0; i < nt_; ++i) wf[i] = 1.0;
    } else {
        o_throw(s, "create_wfilter: Unknown filter shape!");
    }
}

static int make_dumps(orc_sim* s, const char* outdir, o_dump** out)
{
    int n = o_load_int(s, "DUMPS_N", 1, 0);
    o_dump* d = (o_dump*)calloc(n > 0 ? n : 1, sizeof(o_dump));
    *out = d;
    for (int i = 0; i < n; ++i) {
        char type[100], fn1[512], mod[100], fn[1024], path[1600];
        o_load_numstring(s, "DUMPTYPE", i, type, 100, 0, "");
        int k = -1;
        for (int j = 0; j < 7; ++j) if (strcmp(type, dump_names[j]) == 0) k = j;
        if (k < 0) o_throw(s, "dump type not supported by this build");
        d[i].kind = k;
        d[i].nz_denom = 1;
        d[i].yN = 1;
        switch (k) {
        case 0: d[i].nz_denom = o_load_numint(s, "NZ_DENOM", i, 1, 1); d[i].xN = s->NR; d[i].zN = s->NZ / d[i].nz_denom; d[i].tN = s->NT; d[i].iscomplex = 1; break;
        case 1: case 2: case 3: d[i].xN = s->NR; d[i].zN = s->NZ; d[i].tN = 1; d[i].iscomplex = 0; break;
        case 4: d[i].tN = s->NT; d[i].xN = 1; d[i].zN = s->NZ; d[i].iscomplex = 1; break;
        case 5: case 6: d[i].tN = s->NT; d[i].xN = 1; d[i].zN = s->NZ; d[i].iscomplex = 0; break;
        }
        o_load_numstring(s, "DUMP_FILENAME", i, fn1, 200, 0, "");
        o_load_string(s, "DUMP_FILENAME_MOD", mod, 100, 1, "");
        snprintf(fn, sizeof fn, fn1, mod);
        snprintf(path, sizeof path, "%s/%s", outdir, fn);
        d[i].f = fopen(path, "w+b");
        if (!d[i].f) o_throw(s, "void dump_type_empty_class :: init_dumpfile : Unable to open file!");
        udf_header(&d[i]);
        if (k == 1 || k == 2) {
            char ft[100];
            o_load_numstring(s, "FILTER_TYPE", i, ft, 100, 1, "NO");
            d[i].wfilter = (double*)malloc(sizeof(double) * s->NT);
            make_filter(s, ft, d[i].wfilter);
        }
    }
    return n;
}

/* ========================================================================== */
/* API                                                                         */


static int read_tokens(const char* path, o_config* c)
{
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    int cap = 256;
    c->tok = (char**)malloc(sizeof(char*) * cap);
    c->ntok = 0;
    char buf[1024];
    while (fscanf(f, "%1023s", buf) == 1) {
        if (c->ntok == cap) { cap *= 2; c->tok = (char**)realloc(c->tok, sizeof(char*) * cap); }
        c->tok[c->ntok++] = strdup(buf);
    }
    fclose(f);
    return 0;
}

static void free_sim(orc_sim* s)
{
    if (!s) return;
    for (int i = 0; i < s->cfg.ntok; ++i) free(s->cfg.tok[i]);
    free(s->cfg.tok);
    for (int i = 0; i < s->cfg.nover; ++i) free(s->cfg.over[i]);
    free(s->cfg.over);
    if (s->lb) for (int t = 0; t < s->nthreads; ++t) o_linebuf_free(&s->lb[t]);
    free(s->lb);
    void* p[] = {s->ZNET, s->OMEGA_d, s->WAVENUMBER_d, s->HO_d, s->OMEGA, s->WAVENUMBER, s->HO, s->KERR_PROFILE,
                 s->PLASMA_FUNC, s->RAMAN_FUNCTION, s->qx, s->qm1, s->qC, s->qC_r, s->qm1_r, s->FIELD, s->BIG1, s->BIG2};
    for (unsigned i = 0; i < sizeof p / sizeof p[0]; ++i) free(p[i]);
    if (s->fft.n) o_fft_plan_free(&s->fft);
    free(s);
}

orc_sim* orc_create(const char* cfg_path, int nover, const char** over, const char* workdir, int nthreads,
                    char* err, int errlen)
{
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    s->nthreads = nthreads > 0 ? (nthreads > 256 ? 256 : nthreads) : 1;
    if (read_tokens(cfg_path, &s->cfg)) {
        if (err) snprintf(err, errlen, "Error opening file with starting condition.");
        free_sim(s);
        return NULL;
    }
    s->cfg.nover = nover;
    s->cfg.over = (char**)malloc(sizeof(char*) * (nover > 0 ? nover : 1));
    for (int i = 0; i < nover; ++i) s->cfg.over[i] = strdup(over[i]);
    s->jb_armed = 1;
    if (setjmp(s->jb)) {
        if (err) snprintf(err, errlen, "%s", s->err);
        free_sim(s);
        return NULL;
    }
    o_setup(s, workdir);
    s->jb_armed = 0;
    return s;
}

void orc_destroy(orc_sim* s) { free_sim(s); }
const char* orc_error(orc_sim* s) { return s->err; }

int orc_dims(orc_sim* s, int* R, int* T, int* NZ)
{
    *R = s->NR;
    *T = s->NT;
    *NZ = s->NZ;
    return 0;
}

double orc_scalar(orc_sim* s, const char* nm)
{
#define S(x) if (strcmp(nm, #x) == 0) return (double)s->x;
    S(TMIN) S(TMAX) S(TSTEP) S(XMAX) S(OMEGA0) S(OMEGA_MIN) S(OMEGA_MAX) S(GROUP_VELOCITY) S(MAX_TOLERANCE)
    S(NONLIN_REFRINDEX) S(RAMAN_FRACTION) S(TAU_RAMAN) S(OMEGA_RAMAN) S(NEUTRAL_DENSITY) S(RECOMBINATION_TAU)
    S(COLLISION_TAU) S(IONIZATION_POTENTIAL) S(AVALANCHE_CROSSSECTION) S(PONDEROMOTIVE_COEFFICIENT)
    S(AMBIENT_CARRIER_DENSITY) S(ZSTEP) S(CURRENT_Z) S(STEP_N) S(n_Z) S(qS) S(qV) S(n_reject) S(REDUCED_MASS)
#undef S
    if (strcmp(nm, "WAVENUMBER0_re") == 0) return creal(s->WAVENUMBER0);
    if (strcmp(nm, "WAVENUMBER0_im") == 0) return cimag(s->WAVENUMBER0);
    return NAN;
}

int orc_set_scalar(orc_sim* s, const char* nm, double v)
{
    if (strcmp(nm, "ZSTEP") == 0) { s->ZSTEP = v; return 0; }
    if (strcmp(nm, "CURRENT_Z") == 0) { s->CURRENT_Z = v; return 0; }
    if (strcmp(nm, "n_Z") == 0) { s->n_Z = (int)v; return 0; }
    if (strcmp(nm, "MAX_TOLERANCE") == 0) { s->MAX_TOLERANCE = v; return 0; }
    return -1;
}

static void put_c(double* out, const cplx_t* a, size_t n)
{
    for (size_t i = 0; i < n; ++i) { out[2 * i] = creal(a[i]); out[2 * i + 1] = cimag(a[i]); }
}
static void get_c(cplx_t* a, const double* in, size_t n)
{
    for (size_t i = 0; i < n; ++i) a[i] = (real_t)in[2 * i] + (real_t)in[2 * i + 1] * I;
}

int orc_get(orc_sim* s, const char* nm, double* out)
{
    int nt_ = s->NT, nr = s->NR;
    size_t n = (size_t)nr * nt_;
    if (!strcmp(nm, "omega")) { memcpy(out, s->OMEGA_d, sizeof(double) * nt_); return 0; }
    if (!strcmp(nm, "wavenum")) { memcpy(out, s->WAVENUMBER_d, sizeof(double) * 2 * nt_); return 0; }
    if (!strcmp(nm, "ho")) { memcpy(out, s->HO_d, sizeof(double) * nt_); return 0; }
    if (!strcmp(nm, "kerr")) { for (int i = 0; i < nt_; ++i) out[i] = s->KERR_PROFILE[i]; return 0; }
    if (!strcmp(nm, "plasma_func")) { put_c(out, s->PLASMA_FUNC, nt_); return 0; }
    if (!strcmp(nm, "raman_function")) { put_c(out, s->RAMAN_FUNCTION, nt_); return 0; }
    if (!strcmp(nm, "ionrate_ln")) { memcpy(out, s->IONIZATION_RATE_LN, sizeof(double) * O_ION_N); return 0; }
    if (!strcmp(nm, "qdht_x")) { memcpy(out, s->qx, sizeof(double) * nr); return 0; }
    if (!strcmp(nm, "qdht_m1")) { memcpy(out, s->qm1, sizeof(double) * nr); return 0; }
    if (!strcmp(nm, "qdht_c")) { memcpy(out, s->qC, sizeof(double) * n / nt_ * nr); return 0; }
    if (!strcmp(nm, "znet")) { memcpy(out, s->ZNET, sizeof(double) * s->NZ); return 0; }
    if (!strcmp(nm, "field")) { put_c(out, s->FIELD, n); return 0; }
    if (!strcmp(nm, "b1")) { put_c(out, s->BIG1, n); return 0; }
    if (!strcmp(nm, "b2")) { put_c(out, s->BIG2, n); return 0; }
    return -1;
}

int orc_put(orc_sim* s, const char* nm, const double* in)
{
    size_t n = (size_t)s->NR * s->NT;
    if (!strcmp(nm, "field")) { get_c(s->FIELD, in, n); return 0; }
    if (!strcmp(nm, "b1")) { get_c(s->BIG1, in, n); return 0; }
    if (!strcmp(nm, "b2")) { get_c(s->BIG2, in, n); return 0; }
    return -1;
}

int orc_nl_response(orc_sim* s, const double* in, double* out)
{
    int nt_ = s->NT;
    cplx_t* a = (cplx_t*)malloc(sizeof(cplx_t) * nt_);
    cplx_t* b = (cplx_t*)malloc(sizeof(cplx_t) * nt_);
    get_c(a, in, nt_);
    s->jb_armed = 1;
    if (setjmp(s->jb)) { s->jb_armed = 0; free(a); free(b); return -1; }
    o_nl_response(s, &s->lb[0], a, b);
    DISARM(s);
    put_c(out, b, nt_);
    free(a);
    free(b);
    return 0;
}

int orc_plasma(orc_sim* s, const double* E, double* rho, double* loss)
{
    int nt_ = s->NT;
    cplx_t* a = (cplx_t*)malloc(sizeof(cplx_t) * nt_);
    cplx_t* l = (cplx_t*)malloc(sizeof(cplx_t) * nt_);
    real_t* r = (real_t*)malloc(sizeof(real_t) * nt_);
    get_c(a, E, nt_);
    s->jb_armed = 1;
    if (setjmp(s->jb)) { s->jb_armed = 0; free(a); free(l); free(r); return -1; }
    o_plasma_line(s, a, r, 1, l);
    DISARM(s);
    for (int i = 0; i < nt_; ++i) rho[i] = r[i];
    if (loss) put_c(loss, l, nt_);
    free(a);
    free(l);
    free(r);
    return 0;
}

double orc_photoionization(orc_sim* s, double reA, double imA, double rho)
{
    ARM(s, NAN);
    double w = o_photoionization(s, reA, imA, rho);
    DISARM(s);
    return w;
}

double orc_keldysh_rate_ln(orc_sim* s, double lnI)
{
    /* the table is built while IONIZATION_POTENTIAL is still in joules */
    double save = s->IONIZATION_POTENTIAL;
    s->IONIZATION_POTENTIAL = save / (O_IONRATE_DENOM / O_INTENSITY_DENOM);
    double v = o_keldysh_rate_ln(s, lnI);
    s->IONIZATION_POTENTIAL = save;
    return v;
}

int orc_qdht_apply(orc_sim* s, double* data)
{
    size_t n = (size_t)s->NR * s->NT;
    cplx_t* a = (cplx_t*)malloc(sizeof(cplx_t) * n);
    get_c(a, data, n);
    o_qdht_run_many(s, a);
    put_c(data, a, n);
    free(a);
    return 0;
}

int orc_nl_step(orc_sim* s, double* maxA2, double* maxNL2)
{
    ARM(s, -1);
    nl_step(s, maxA2, maxNL2);
    DISARM(s);
    return 0;
}

int orc_lstep(orc_sim* s)
{
    lstep(s);
    return 0;
}

int orc_propagator(orc_sim* s)
{
    ARM(s, -1);
    if (s->n_Z == 0) s->n_Z = 1;
    propagator(s);
    s->STEP_N++;
    s->CURRENT_Z += s->ZSTEP;
    DISARM(s);
    return 0;
}

int orc_fixed_step(orc_sim* s, double dz, double* maxA2, double* maxNL2)
{
    ARM(s, -1);
    s->ZSTEP = dz;
    nl_step(s, maxA2, maxNL2);
    o_qdht_run_many(s, s->BIG1);
    lstep(s);
    o_qdht_run_many(s, s->FIELD);
    s->CURRENT_Z += dz;
    s->STEP_N++;
    DISARM(s);
    return 0;
}

/* one step with the NL RK4 at dz_rk and the linear step at dz_lin: the
 * accepted attempt of propagator() after a step-size change (the NL step runs
 * at the previous ZSTEP, the Lstep at the new one) */
void o_fixed_step2(orc_sim* s, double dz_rk, double dz_lin)
{
    double a, b;
    s->ZSTEP = dz_rk;
    nl_step(s, &a, &b);
    o_qdht_run_many(s, s->BIG1);
    s->ZSTEP = dz_lin;
    lstep(s);
    o_qdht_run_many(s, s->FIELD);
    s->CURRENT_Z += dz_lin;
    s->STEP_N++;
}

int orc_fixed_step2(orc_sim* s, double dz_rk, double dz_lin)
{
    ARM(s, -1);
    o_fixed_step2(s, dz_rk, dz_lin);
    DISARM(s);
    return 0;
}

int orc_run(orc_sim* s, const char* outdir, long max_steps)
{
    o_dump* volatile d = NULL;
    volatile int nd = 0;
    s->jb_armed = 1;
    if (setjmp(s->jb)) {
        s->jb_armed = 0;
        for (int i = 0; i < nd; ++i) { if (d[i].f) fclose(d[i].f); free(d[i].wfilter); }
        free(d);
        return -1;
    }
    {
        o_dump* dd = NULL;
        nd = make_dumps(s, outdir ? outdir : ".", &dd);
        d = dd;
    }
    long cycles = 0;
    int ret = 0;
    for (;;) {
        if (s->CURRENT_Z >= s->ZNET[s->n_Z]) {
            for (int i = 0; i < nd; ++i) do_dump(s, &d[i]);
            if (s->n_Z == s->NZ - 1) break;
            s->n_Z++;
        }
        if (max_steps > 0 && cycles >= max_steps) { ret = 1; break; }
        propagator(s);
        s->STEP_N++;
        s->CURRENT_Z += s->ZSTEP;
        cycles++;
    }
    s->jb_armed = 0;
    for (int i = 0; i < nd; ++i) { fclose(d[i].f); free(d[i].wfilter); }
    free(d);
    return ret;
}

double orc_sqrtho(double x) { return (double)o_sqrtho((real_t)x); }

double orc_raman_circular(int n, const double* Iv, double fr, double M, double zre, double zim, double* y)
{
    /* closed form of y = (1-fR) I + sum_m R[m] I[(n-m) mod T], R[m] = fR M Im(z^m):
     * s[n] = z s[n-1] + I[n];  y[n] = (1-fR) I[n] + fR M Im(s[n](1 - z^T) + z^{n+1} s[T-1]) */
    double _Complex z = zre + zim * I, sacc = 0;
    double _Complex* sv = (double _Complex*)malloc(sizeof(double _Complex) * n);
    for (int i = 0; i < n; ++i) { sacc = z * sacc + Iv[i]; sv[i] = sacc; }
    double _Complex zT = cpow(z, n), zp = z;
    for (int i = 0; i < n; ++i) {
        y[i] = (1 - fr) * Iv[i] + fr * M * cimag(sv[i] * (1.0 - zT) + zp * sv[n - 1]);
        zp *= z;
    }
    free(sv);
    return 0;
}
