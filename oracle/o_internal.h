/* o_internal.h -- private state of the oracle (test infrastructure only).
 * The struct mirrors the reference's process globals (solver.h:111-236) for
 * the single-process gsolver_solid_qdht configuration. */
#ifndef GSOLVER_ORACLE_INTERNAL_H
#define GSOLVER_ORACLE_INTERNAL_H

#include <complex.h>
#include <setjmp.h>
#include <stdio.h>
#include "oracle.h"

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
typedef ORACLE_REAL real_t;
typedef ORACLE_REAL _Complex cplx_t;

/* constants of solver.h:65-109 and SIconstants.h:4-38 */
#define O_LIGHT_VELOCITY      299792458.0
#define O_VACUUM_PERMEABILITY 1.2566370615e-6
#define O_ELECTRON_CHARGE     1.60217733e-19
#define O_ELECTRON_MASS       9.1093897e-31
#define O_HBAR                1.0545727e-34
#define O_MINSTEP_RATIO       1e-10
#define O_MAX_TOL_DEFAULT     1e-4
#define O_ION_LN_STEP         0.003304003148325
#define O_ION_MIN_LN          29.933606208922594
#define O_ION_MIN_I           1e13
#define O_ION_N               8192
#define O_INTENSITY_DENOM     1e16
#define O_FIELD_DENOM         1e8
#define O_INTENSITY_DENOM_LN  36.841361487904734
#define O_IONRATE_DENOM       1e20
#define O_IONRATE_DENOM_LN    46.0517018
#define O_ABS_LAYER_WIDTH     0.1
#define O_ABS_LAYER_BETA      10.0
#define O_MAX_KT2             0.25

/* ---- FFT plan (o_fft.c) ---- */
typedef struct {
    int n;
    int nfac;
    int fac[64];
    cplx_t* tw_fwd;   /* e^{-2 pi i j/n} */
    cplx_t* tw_bwd;   /* e^{+2 pi i j/n} */
} o_fftplan;
int  o_fft_plan_init(o_fftplan* p, int n);
void o_fft_plan_free(o_fftplan* p);
/* out-of-place, in != out; sign -1 forward, +1 backward; unnormalised */
void o_fft_exec(const o_fftplan* p, const cplx_t* in, cplx_t* out, int sign);

/* ---- special functions (o_special.c), always double ---- */
double o_j0(double x);
double o_j1(double x);
double o_j0_zero(int k);
double o_dawson(double x);
int    o_ellipke(double m, double* K, double* E);

/* ---- per-thread scratch for one radial line ---- */
typedef struct {
    cplx_t *field_real, *b1, *b2, *b3, *b4;   /* FIELD_REAL_SMALL, NL_SMALLBUFFER1..4 */
    cplx_t *out1, *out2, *out3, *out4, *in;   /* NL_OUTPUT1..4, NL_INPUT */
} o_linebuf;

typedef struct {
    int ntok;
    char** tok;
    int nover;
    char** over;
} o_config;

struct orc_sim {
    jmp_buf jb;
    int jb_armed;
    char err[512];

    o_config cfg;
    int nthreads;

    /* grids */
    int NT, NR, NZ, n_Z;
    double TMIN, TMAX, TSTEP;
    double XMAX;
    double* ZNET;
    double OMEGA0, OMEGA_MIN, OMEGA_MAX, GROUP_VELOCITY;
    double _Complex WAVENUMBER0;
    double* OMEGA_d;                /* double copies used by set-up */
    double _Complex* WAVENUMBER_d;
    double* HO_d;
    double MAX_TOLERANCE;
    double LAMBDA_V;

    /* nonlinear parameters (after the unit rescaling of solver_main.cpp:385-393) */
    double NONLIN_REFRINDEX, RAMAN_FRACTION, TAU_RAMAN, OMEGA_RAMAN;
    double NEUTRAL_DENSITY, RECOMBINATION_TAU, COLLISION_TAU, IONIZATION_POTENTIAL;
    double AVALANCHE_CROSSSECTION, PONDEROMOTIVE_COEFFICIENT, AMBIENT_CARRIER_DENSITY;
    double REDUCED_MASS;
    double IONIZATION_RATE_LN[O_ION_N];

    /* march-precision tables */
    real_t* OMEGA;
    cplx_t* WAVENUMBER;
    real_t* HO;
    real_t* KERR_PROFILE;
    cplx_t* PLASMA_FUNC;
    cplx_t* RAMAN_FUNCTION;

    /* QDHT plan (fhatha.cpp:177-198) */
    double qS, qV;
    double* qx;      /* j_i / S */
    double* qm1;     /* |J1(j_i)| */
    double* qC;      /* C[i + N*j] */
    real_t* qC_r;
    real_t* qm1_r;

    /* state */
    cplx_t *FIELD, *BIG1, *BIG2;
    double ZSTEP, CURRENT_Z;
    unsigned long STEP_N;
    long n_reject;

    o_fftplan fft;
    o_linebuf* lb;   /* nthreads entries */
};

/* error handling: the reference throws const char*; the oracle longjmps back
 * to the API entry that armed s->jb. */
void o_throw(orc_sim* s, const char* msg);
/* arms s->jb in an API entry: a throw returns failret from that entry */
#define ARM(s, failret)            \
    do {                           \
        (s)->jb_armed = 1;         \
        if (setjmp((s)->jb)) {     \
            (s)->jb_armed = 0;     \
            return failret;        \
        }                          \
    } while (0)
#define DISARM(s) ((s)->jb_armed = 0)

/* parallel for over [0, n) on s->nthreads threads; fn(s, tid, i0, i1, ctx) */
void o_parallel(orc_sim* s, int n, void (*fn)(orc_sim*, int, int, int, void*), void* ctx);

/* config access, process_input.cpp:297-482 semantics */
int    o_cfg_find(orc_sim* s, const char* name, const char** val);
double o_load_float(orc_sim* s, const char* name, int has_def, double def);
int    o_load_int(orc_sim* s, const char* name, int has_def, int def);
double o_load_numfloat(orc_sim* s, const char* name, int num, int has_def, double def);
int    o_load_numint(orc_sim* s, const char* name, int num, int has_def, int def);
void   o_load_string(orc_sim* s, const char* name, char* out, int n, int has_def, const char* def);
void   o_load_numstring(orc_sim* s, const char* name, int num, char* out, int n, int has_def, const char* def);

/* physics pieces (o_march.c) */
double o_photoionization(orc_sim* s, double reA, double imA, double ro);
void   o_plasma_line(orc_sim* s, const cplx_t* field, real_t* pro, int stride, cplx_t* loss);
void   o_nl_response(orc_sim* s, o_linebuf* b, const cplx_t* in, cplx_t* out);
void   o_qdht_run_many(orc_sim* s, cplx_t* data);
double o_keldysh_rate_ln(orc_sim* s, double lnI);
real_t o_sqrtho(real_t x);

void   o_fixed_step2(orc_sim* s, double dz_rk, double dz_lin);
void   o_setup(orc_sim* s, const char* workdir);   /* o_setup.c */
void   o_linebuf_alloc(o_linebuf* b, int nt);
void   o_linebuf_free(o_linebuf* b);

#endif
