/*
 * gsolver_amd.h -- C-ABI of the MI355X (gfx950) implementation of the
 * gsolver_solid_qdht per-z-step UPPE right-hand side and RK march.
 *
 * The reference has no runtime plugin API: its seams are compile-time (the
 * ht_run dispatch of solver.h:323-346, the #ifdef _ENABLE_CUDA bodies of
 * calculate_NLstep / calculate_Lstep, propagator.cu:270-374,518-569) and
 * process globals (solver.h:111-236).  This header turns those seams into
 * entry points with plain pointers and sizes.  Each function says which
 * reference interface it replaces.
 *
 * Conventions
 *  - Every entry point returns 0 on success or a negative GS_E* code; the
 *    message of the last failure on the calling thread is gs_last_error().
 *    Messages reuse the reference's throw strings where one exists
 *    (e.g. "Minimum Z step size reached, collapse is possible!").
 *  - A plan owns its device memory; the caller owns host buffers.  Complex
 *    host arrays are interleaved (re, im) doubles, [N_R][N_T] row-major with
 *    omega (or t) contiguous, exactly the layout of FIELD / BIGBUFFER1/2.
 *  - One plan per host thread; calls on a plan are ordered on the plan's
 *    HIP stream and return after the stream has finished the call's work.
 */
#ifndef GSOLVER_AMD_H
#define GSOLVER_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gs_plan gs_plan;

/* arithmetic of the march */
enum gs_precision {
    GS_FP64 = 0,   /* all kernels in double: the FP64 GPU run results are checked against */
    GS_FP32 = 1,   /* plain FP32: FP32 MFMA QDHT, FP32 FFT, scans and propagator */
    GS_MIXED = 2   /* split-BF16 3-pass MFMA QDHT with FP32 accumulation; FP32 FFT,
                      Kerr/Raman and W(I); FP64 plasma-density scan, propagator
                      phase reduction, step-control norms and dump integrals */
};

enum gs_status {
    GS_OK = 0,
    GS_E_ARG = -1,        /* bad argument / unsupported configuration */
    GS_E_CONFIG = -2,     /* missing or invalid config key, unreadable input file */
    GS_E_HIP = -3,        /* HIP runtime failure */
    GS_E_PHYSICS = -4,    /* the reference's run-time throws: blast, NaN, step collapse */
    GS_E_NODEVICE = -5    /* no GPU visible: the library has no CPU fallback */
};

/* buffer ids for gs_get / gs_put / gs_qdht */
enum gs_buffer {
    GS_FIELD = 0,    /* FIELD:      the field in (r, omega) */
    GS_B1 = 1,       /* BIGBUFFER1: NL increment (r, omega); k-space after the first QDHT */
    GS_B2 = 2,       /* BIGBUFFER2: the field carried across steps in (k_perp, omega) */
    GS_FK = 3        /* FIELD in (k_perp, omega) after the linear step, before the QDHT back;
                        equal to BIGBUFFER2 except where the linear step zeroes FIELD (there
                        BIGBUFFER2 keeps its value), so putting it also sets BIGBUFFER2 there */
};

/* Set-up from a reference config file: replaces process_input()
 * (process_input.cpp:33-95) = load_info + load_nonlindata +
 * initialize_variables (solver_main.cpp:265-398) + load_kerr_th_profiles +
 * create_gg (pulse_gg.cpp:6-124) + FIELD/FIELD_DENOM + init_zstep_kerr +
 * BIGBUFFER2 = QDHT(FIELD).  `over` holds argv[2..] of the reference command
 * line ("NAME VALUE" pairs searched before the file, process_input.cpp:300).
 * Relative file names (DISPERSION_FILE, dump names) resolve against the
 * current working directory, as in the reference.  When workdir is non-NULL,
 * omega.bin and wavenum.bin are written there (process_input.cpp:118-119). */
int gs_plan_create(const char* cfg_path, int nover, const char* const* over,
                   int precision, const char* workdir, gs_plan** out);
void gs_plan_destroy(gs_plan* p);
const char* gs_last_error(void);

/* The host-side set-up of gs_plan_create without a device: parses the config
 * and builds the tables (process_input.cpp:98-182,211-295, solver_main.cpp:
 * 340-413, ionization.cpp:313-331, fhatha.cpp:177-198), then copies table
 * `name` (see gs_get_table) into out.  Lets a host check its inputs, and the
 * set-up be tested, on machines without a GPU. */
int gs_setup_table(const char* cfg_path, int nover, const char* const* over, const char* name,
                   double* out, size_t n);

/* sizes: N_R, N_T, N_Z */
int gs_dims(gs_plan* p, int* nr, int* nt, int* nz);
/* named scalars of the plan (ZSTEP, CURRENT_Z, OMEGA0, TSTEP, MAX_TOLERANCE,
 * STEP_N, n_Z, n_reject, GROUP_VELOCITY, OMEGA_MIN, OMEGA_MAX, ...).  Settable:
 * ZSTEP, CURRENT_Z, STEP_N, n_Z, MAX_TOLERANCE, ADJOINT_CHECKPOINT_BYTES and
 * the GEMM switches QDHT_SKIP_COLUMNS, QDHT_K_GROUPS, QDHT_TILE_ORDER,
 * QDHT_F32_NARROW_FULL (these change no result). */
int gs_get_scalar(gs_plan* p, const char* name, double* v);
int gs_set_scalar(gs_plan* p, const char* name, double v);
/* named host-side tables built at set-up, FP64: "omega" (T), "wavenum" (T cplx),
 * "ho" (T), "kerr" (T), "plasma_func" (T cplx), "raman_function" (T cplx),
 * "ionrate_ln" (8192), "qdht_x" (R), "qdht_m1" (R), "qdht_c" (R*R, row i of C),
 * "znet" (N_Z) */
int gs_get_table(gs_plan* p, const char* name, double* out, size_t n);

/* host <-> device copies of a [N_R][N_T] complex buffer (replaces direct access
 * to the FIELD / BIGBUFFER1 / BIGBUFFER2 globals) */
int gs_get(gs_plan* p, int buffer, double* host);
int gs_put(gs_plan* p, int buffer, const double* host);

/* calculate_NLstep (propagator.cu:268-426): BIGBUFFER1 = RK4 average of the
 * nonlinear RHS over each radial line of FIELD at step dz; returns
 * max|FIELD|^2 and max|BIGBUFFER1|^2 over (r, omega) samples. */
int gs_nl_step(gs_plan* p, double dz, double* maxA2, double* maxNL2);
/* calculate_NLresponse (propagator.cu:629-783) for n_lines lines given in
 * host memory (n_lines * N_T complex) -> out (same shape). */
int gs_nl_response(gs_plan* p, const double* in, double* out, int n_lines);
/* gs_nl_response without any device allocation: the same kernels, with the
 * FP64 staging copy and the field-precision input and output in a fixed
 * device scratch array of the library (12 MiB per GPU, held until the process
 * ends) instead of buffers allocated per call.  No counterpart in the
 * reference: for callers that cannot allocate device memory per call (tight
 * serving loops, harnesses that forbid hipMalloc).  n_lines * N_T must not
 * exceed 262144 samples (GS_E_ARG otherwise).  All calls on one GPU share the
 * scratch: calls that overlap on the same GPU (plans on other threads or
 * streams) race and must be serialised by the caller. */
int gs_nl_response_noalloc(gs_plan* p, const double* in, double* out, int n_lines);
/* qdht_plan::run_many (fhatha.cpp:218-230) via ht_run (solver.h:337-346),
 * in place on a device buffer. */
int gs_qdht(gs_plan* p, int buffer);
/* calculate_Lstep (propagator.cu:516-626): with BIGBUFFER1 in k-space,
 * BIGBUFFER2 += dz B1/(1+sqrtHO); FIELD_k = BIGBUFFER2 * exp(dz(...)). */
int gs_linear_step(gs_plan* p, double dz);
/* propagator() (propagator.cu:125-266): one accepted step with the
 * reference's controller, then CURRENT_Z += ZSTEP (solver_main.cpp:217-220). */
int gs_step(gs_plan* p);
/* one step at a fixed dz (no controller): NL step, fused QDHT+Lstep, QDHT. */
int gs_fixed_step(gs_plan* p, double dz, double* maxA2, double* maxNL2);
/* the main loop of solver_main.cpp:202-225 with the .udf dumps of the config
 * (dump_types.cpp; types flux, maxI, field_axis, plasma_max, average_spectrum,
 * full, plasma_axis) written into outdir.  Stops after max_steps accepted
 * steps when max_steps > 0 (returns 1 then; 0 at the last plane).  A call
 * after such a stop continues the march; dump files already in outdir keep
 * the planes written so far (they are reopened, not truncated, once the plan
 * has written its first plane), so a march run in chunks leaves the same
 * files as one uninterrupted call. */
int gs_march(gs_plan* p, const char* outdir, long max_steps);

/* gs_march for n independent runs at once (energy / width scans, uneven
 * grids): each plan marches to its last plane with its own controller and
 * dumps (outdirs[i], NULL = "."), the runs' step attempts overlapping on the
 * device.  rcs (may be NULL) receives 0 per finished run.  No counterpart in
 * the reference, which runs one configuration per process. */
int gs_march_many(gs_plan* const* plans, int n, const char* const* outdirs, int* rcs);

/* Dump-side reductions of the current FIELD, per radial line (N_R each):
 * flux = TSTEP * sum_t |E|^2, maxI = max_t |E|^2, rhomax = max_t rho(t),
 * with E = IFFT(FIELD)/N_T (dump_types.cpp:123-165, propagator.cu:935-946).
 * Any output pointer may be NULL. */
int gs_line_diagnostics(gs_plan* p, double* flux, double* maxI, double* rhomax);

/* Stand-alone batched t-axis transform of n_lines host lines of N_T complex
 * values with the plan's device FFT: sign -1 forward, +1 backward,
 * unnormalised (the FFTW plans of solver_main.cpp:339-341). */
int gs_fft_lines(gs_plan* p, const double* in, double* out, int n_lines, int sign);

/* n steps of gs_fixed_step enqueued back to back on the plan's stream with one
 * synchronisation at the end (the fixed-dz throughput mode of the benchmark;
 * the reference has no counterpart).  ms4 (may be NULL) receives the mean
 * device time per step of the four kernel groups, as gs_last_timings; with
 * ms4 NULL no timing events are recorded between the kernels. */
int gs_fixed_steps(gs_plan* p, double dz, int n, double* ms4);

/* Reverse mode (no counterpart in the reference, which has no adjoint).
 *
 * gs_adjoint_forward: n fixed steps from the plan's current state, step k with
 * the NL RK4 at dz_rk[k] and the linear step at dz_lin[k] (the accepted attempt
 * of propagator(), propagator.cu:125-266, runs the NL step at the previous
 * ZSTEP and the linear step at the new one; scalar "LAST_DZ_NL" after gs_step
 * gives the former, "ZSTEP" the latter).  The state is left at the end of the
 * march.  FIELD is kept on the device before every step when n of them fit in
 * the scalar "ADJOINT_CHECKPOINT_BYTES" (gs_set_scalar; default, or negative:
 * 60 % of the device memory free at the call); otherwise FIELD and FIELD_k
 * (the live BIGBUFFER2) are kept every K = ceil(sqrt(2n)) steps and
 * gs_adjoint_backward recomputes one segment of K steps at a time (memory
 * (2 ceil(n/K) + K) fields, one extra forward).
 *
 * gs_adjoint_backward: from the cotangents of the final FIELD (gfield) and
 * BIGBUFFER2 (gb2, NULL = 0), the cotangents of the initial FIELD (gfield0) and
 * BIGBUFFER2 (gb2_0, may be NULL), the exact transpose of the recorded steps
 * with the step sizes held fixed.  tie_b2 != 0 adds the BIGBUFFER2 part to
 * gfield0 through the QDHT transpose, for runs whose initial BIGBUFFER2 is
 * QDHT(FIELD) (as gs_plan_create sets it up).  Cotangent of a complex value z
 * of a real objective L: dL/dRe z + i dL/dIm z.  Host arrays [N_R][N_T] complex.
 * May be called again with other cotangents for the same recorded steps. */
int gs_adjoint_forward(gs_plan* p, int n, const double* dz_rk, const double* dz_lin);
int gs_adjoint_backward(gs_plan* p, const double* gfield, const double* gb2, int tie_b2, double* gfield0,
                        double* gb2_0);

/* Device time of the kernels of the last gs_fixed_step / gs_step, in ms:
 * [0] NL RK4 kernel, [1] QDHT + linear step, [2] QDHT back, [3] reductions. */
int gs_last_timings(gs_plan* p, double* ms4);

#ifdef __cplusplus
}
#endif
#endif
