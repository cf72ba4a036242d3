/*
 * oracle.h -- CPU restatement of the gsolver_solid_qdht UPPE z-march.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in gsolver_amd/ is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links or calls it.
 *
 * It restates, with NONLINEARITY_ON defined and the compile-time switches of
 * the gsolver_solid_qdht make target (KELDYSH_IONIZATION, PLASMA_DISPERSION,
 * _UNIAXIAL, QDHT, _ODE_RK4, TRANSVERSE_DIMENSIONS=1, one process), the
 * reference CPU algorithm in these files:
 *   propagator.cu:125-266 (driver), :376-425 (RK4 NL step), :570-625 (Lstep),
 *   :629-783 (NL response), ionization.cpp:9-147 (plasma Heun + W(I)),
 *   ionization.cpp:203-232,313-355 (Keldysh table), fhatha.cpp:177-230 (QDHT),
 *   solver_main.cpp:151-263,265-413 (main loop, tables, Raman),
 *   process_input.cpp:33-182,297-482,548-623,677-753,806-871,966-1021,
 *   pulse_gg.cpp:6-124, extmath.h:234-252, extmath.cpp:154-294,
 *   dump_types.cpp:25-165,391-457,527-555, dumps.cpp:15-52.
 *
 * Parity pins (see tests/test_oracle_*.py): bin/omega.bin, bin/wavenum.bin,
 * the 128-slab linear-regime dumps bin/fil_SiO2__{f,fa,mI,pm,s}.udf,
 * mpmath golden vectors (tests/golden/make_special.py) for Bessel J0/J1/zeros, Dawson and the
 * complete elliptic integrals, brute-force DFT for the FFT, closed forms
 * for the Raman convolution, the plasma recurrence and the RK4 combine.
 * The reverse mode of o_adjoint.c (no reference counterpart) is pinned by
 * central finite differences of this forward restatement.
 * Parity unpinned (no reference-held vector exists): the Keldysh rate table
 * values, and the nonlinear march end to end.
 *
 * Two builds come from the same sources: liboracle_fp64.so (ORACLE_REAL =
 * double, the reference's _EXTMATH_DOUBLE build) and liboracle_fp32.so
 * (ORACLE_REAL = float, its _EXTMATH_SINGLE counterpart for the march
 * arithmetic; set-up tables are always built in double and rounded).
 * Complex arrays are passed as interleaved (re, im) doubles in the API,
 * whatever the internal precision.
 */
#ifndef GSOLVER_ORACLE_H
#define GSOLVER_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* ---- standalone numerics ------------------------------------------------ */
/* FFTW sign convention: sign=-1 forward (e^{-2pi i kn/N}), +1 backward; unnormalised. */
int    orc_fft(int n, const double* in, double* out, int sign);
double orc_bessel_j0(double x);
double orc_bessel_j1(double x);
double orc_bessel_zero_j0(int k);            /* k-th positive zero, k >= 1 */
double orc_dawson(double x);
int    orc_ellipke(double m, double* K, double* E);   /* parameter m in [0,1] */
double orc_sqrtho(double x);                  /* extmath.h:245-252 */
double orc_raman_circular(int n, const double* intensity, double fr, double M,
                          double zre, double zim, double* y);  /* closed form check helper */

/* ---- simulation ---------------------------------------------------------- */
/* process_input(): config file + "NAME VALUE" overrides (argv[2..] of the
 * reference).  Writes omega.bin / wavenum.bin into workdir when non-NULL, as
 * process_input.cpp:118-119 does into the current directory.  Returns NULL and
 * fills err on failure (reference: throw const char*). */
orc_sim* orc_create(const char* cfg_path, int nover, const char** over,
                    const char* workdir, int nthreads, char* err, int errlen);
void     orc_destroy(orc_sim* s);
const char* orc_error(orc_sim* s);

int    orc_dims(orc_sim* s, int* R, int* T, int* NZ);
double orc_scalar(orc_sim* s, const char* name);
int    orc_set_scalar(orc_sim* s, const char* name, double v);
/* named arrays: "omega"(T) "wavenum"(T cplx) "ho"(T) "kerr"(T) "plasma_func"(T cplx)
 * "raman_function"(T cplx) "ionrate_ln"(8192) "qdht_x"(R) "qdht_m1"(R) "qdht_c"(R*R, C[i+R*j])
 * "znet"(NZ) "field"(R*T cplx) "b1"(R*T cplx) "b2"(R*T cplx) */
int    orc_get(orc_sim* s, const char* name, double* out);
int    orc_put(orc_sim* s, const char* name, const double* in);

/* single hot-path functions, on caller buffers (complex interleaved) */
int    orc_nl_response(orc_sim* s, const double* in, double* out);            /* a3 */
int    orc_plasma(orc_sim* s, const double* E, double* rho, double* loss);     /* a4 */
double orc_photoionization(orc_sim* s, double reA, double imA, double rho);   /* a5 */
double orc_keldysh_rate_ln(orc_sim* s, double lnI);                           /* a6 */
int    orc_qdht_apply(orc_sim* s, double* data);                              /* a10, [R][T] cplx in place */

/* stateful steps on FIELD / BIGBUFFER1 / BIGBUFFER2 */
int    orc_nl_step(orc_sim* s, double* maxA2, double* maxNL2);  /* a2: FIELD -> b1 */
int    orc_lstep(orc_sim* s);                                   /* a11 (uses ZSTEP), b1 must be in k-space */
int    orc_propagator(orc_sim* s);                              /* a1: one accepted step */
int    orc_fixed_step(orc_sim* s, double dz, double* maxA2, double* maxNL2);  /* one step at a given dz, no controller */
/* one step, NL RK4 at dz_rk and linear step at dz_lin (no controller) */
int    orc_fixed_step2(orc_sim* s, double dz_rk, double dz_lin);

/* ---- reverse mode (o_adjoint.c): exact transpose of the restated scheme ---- */
/* one line: gin = J^H lam of calculate_NLresponse at `in` */
int    orc_nl_response_vjp(orc_sim* s, const double* in, const double* lam, double* gin);
/* [R][T]: gA = J^H g of the RK4 NL step at FIELD = A and step dz */
int    orc_nl_step_vjp(orc_sim* s, const double* A, const double* g, double dz, double* gA);
/* n fixed steps from the current FIELD / BIG2 (state left at the end), then the
 * cotangents of the initial FIELD and BIG2 from those of the final ones
 * (gb2 may be NULL: zero).  Cotangent of complex z: dL/dRe z + i dL/dIm z. */
int    orc_adjoint_march(orc_sim* s, int n, const double* dz_rk, const double* dz_lin, const double* gfield,
                         const double* gb2, double* gfield0, double* gb2_0);

/* main loop (solver_main.cpp:202-225) with dumps written into outdir; stops
 * after max_steps cycles when max_steps > 0 (returns 1 then, 0 at Z_MAX). */
int    orc_run(orc_sim* s, const char* outdir, long max_steps);

#ifdef __cplusplus
}
#endif
#endif
