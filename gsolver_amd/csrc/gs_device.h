// gs_device.h -- device-side data of a plan and the launchers of the gfx950
// kernels (line kernels in gs_line.hip, QDHT GEMMs and element-wise kernels in
// gs_qdht.hip).  Host code (gs_plan.cpp) fills these structs and calls the
// launchers; nothing here is visible through the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gs {

enum Prec { kFP64 = 0, kFP32 = 1, kMIXED = 2 };

// status bits written by kernels (atomicOr into DevStatus::flags)
enum : int {
    kFlagBlast = 1,      // ionization table index beyond 8190 (ionization.cpp:114-118)
    kFlagNaN = 2,        // NaN intensity (ionization.cpp:104-105)
    kFlagScanSlow = 4    // plasma scan needed the chunk-sequential fallback
};

struct DevStatus {
    double maxA2, maxNL2;
    int flags;
    int fallbacks;
};

// Step-controller state of the device-driven march (propagator.cu:125-266
// and the main loop of solver_main.cpp:202-225 without host round trips).
// One z-step attempt ("slot") is: NL RK4 at dz -> controller -> linear half at
// dz_lin if accepted -> finish.  Kernels of a slot exit at once when state != 0
// (and the linear half when the attempt was rejected).
struct DevCtl {
    double dz;          // ZSTEP: dz of the next NL attempt
    double z;           // CURRENT_Z
    double z_plane;     // ZNET[n_Z], the plane the march is heading for
    double dz_floor;    // 1e-10 (ZNET[n_Z] - ZNET[n_Z - 1]): below it the run has collapsed
    double tol;         // MAX_TOLERANCE
    double dz_lin;      // dz of this slot's linear half (the new ZSTEP)
    long long steps;    // STEP_N (accepted and rejected attempts)
    long long rejects;  // n_reject
    long long accepted; // accepted steps since the counter was last cleared by the host
    int accept;         // this slot: 1 accepted, 0 rejected
    int state;          // 0 running, 1 reached z_plane, 2 error
    int err;            // 1 step collapse, 2 NaN intensity, 3 blast
    int dead_dirty;     // FIELD still holds values in the columns the linear step zeroes:
                        // the next accepted QDHT back clears them (and this flag)
};

// Column selection of a QDHT GEMM that skips the omega columns the linear step
// zeroes (omega outside [OMEGA_MIN, OMEGA_MAX], propagator.cu:601-603): only the
// 32-column blocks listed in ctile are computed.  The QDHT back can also
// clear FIELD on the other blocks (ztile), once: when zero_dead is set, or,
// in the device-driven march, when DevCtl::dead_dirty is.
struct QCols {
    const int* ctile;   // first column of each computed 32-column block
    int nct;
    const int* ztile;   // first column of each skipped 32-column block
    int nzt;
    int zero_dead;      // host-driven launches: clear the skipped blocks of the output
    int split3;         // mixed: operands in three bf16 parts, six products (FP32-level accuracy)
    int ks;             // mixed / FP32: K groups per workgroup, 1 or 2 (same bits; 0: the default)
    int order;          // tiles per XCD: 0 all row tiles of a column tile, 1 all column tiles of a row tile
    int wide;           // FP32 all-column transform: 64-column tiles (else the 32-column ones; same bits)
};

// Scalars and table pointers the NL kernels read.  Table pointers refer to
// device arrays in the plan's field precision unless noted.
struct NLArgs {
    int nr;                     // number of real radial lines (<= padded rows)
    double inv_nt;              // 1/N_T
    // plasma (solver units)
    double dt;                  // TSTEP * 1e20
    double neutral, inv_neutral, inv_recomb, U, pond, sigma, rho_init;
    const double* ionrate;      // 8192 FP64 ln W table
    // Raman / Kerr
    int use_kerr, use_raman;
    double one_m_fr, frM;
    double onemzT_re, onemzT_im;   // 1 - z^T
    const void* zpow;           // z^k, k = 0..T (C2<R>)
    const void* pf;             // PLASMA_FUNC (C2<R>)
    const void* kerr;           // KERR_PROFILE (R)
    const void* tw;             // e^{-2 pi i k/T}, k < T (C2<R>)
    const void* dft16;          // mixed: the radix-16 DFTs (sign -1, +1) as split-BF16 MFMA A
                                // fragments [sign][k slice][hi, lo][lane][8] (bf16); null otherwise
    const void* mfma_tw;        // mixed: lane-order twiddles of the MFMA passes after the first,
                                // [pass][group j][r] = e^{-2 pi i r (j mod Ns) / (16 Ns)} (C2<float>)
    int mfma_fwd;               // mixed: forward line FFTs on MFMA (0 none, 1 the Kerr/Raman term, 2 all)
    int mfma_inv;               // mixed: inverse line FFTs of non-ionizing lines on MFMA
    DevStatus* status;
    int nl_e;                   // samples per thread of the RK4 kernel (0: default)
    const DevCtl* ctl;          // device-driven march: dz from ctl->dz, skip when ctl->state != 0
};

// Linear step (propagator.cu:570-625) from per-point set-up tables:
//   rate[i][w] = -HO[w] - s k1[w],  g[i][w] = 1/(1+s),  s = sqrtHO(-kt2_i / k1_w^2)
// with g < 0 marking points where FIELD is zeroed and B2 left alone (omega
// outside [OMEGA_MIN, OMEGA_MAX], kt2 > MAX_KT2 Re(k^2), padding rows).
// After every linear step B2 equals FIELD_k on the other points, and B2 never
// changes on the masked ones, so the device keeps one array: fk holds FIELD_k
// (zero where masked) and is the B2 the next step updates; the plan's b2
// buffer keeps only the masked points' values, merged back for gs_get.
struct LstepArgs {
    double dz;
    const double* k2;           // Im k(omega)
    const void* rate;           // [rp][T], P precision of the policy
    const void* gfac;           // [rp][T], field precision
    void* fk;                   // FIELD in k-space = B2 on unmasked points (C2<R>), updated in place
    const DevCtl* ctl;          // device-driven march: dz = ctl->dz_lin, seed = ctl->steps; skip unless accepted
    uint32_t seed;              // the step's STEP_N (attempts before this linear step): salt of the
                                // mixed policy's unbiased-rounding draws
};

// sqrt(1+x) - 1 as the 20-term Taylor sum of extmath.h:245-252, Horner form
constexpr double kSqrtHO[20] = {0.5, -0.125, 0.0625, -0.0390625, 0.02734375, -0.0205078125, 0.01611328125,
                                -0.013092041015625, 0.010910034179688, -0.009273529052734, 0.008008956909180,
                                -0.007007837295532, 0.006199240684509, -0.005535036325455, 0.004981532692909,
                                -0.004514514002949, 0.004116174532101, -0.003773159987759, 0.003475278936094,
                                -0.003214633015887};
__host__ __device__ inline double sqrtho(double x)
{
    double r = kSqrtHO[19];
    for (int j = 18; j >= 0; --j) r = r * x + kSqrtHO[j];
    return r * x;
}

// Fixed device scratch of the allocation-free entry point gs_nl_response_noalloc:
// at most kNoAllocMaxSamples line samples per call (n_lines * N_T), staged as the
// FP64 host copy plus the field-precision input and output (16 + 2 x 16 bytes
// per sample at most, each region 256-byte aligned).
constexpr size_t kNoAllocMaxSamples = 262144;
constexpr size_t kNoAllocScratchBytes = 48 * kNoAllocMaxSamples + 1024;
// Device address of that scratch on the current device: hipGetSymbolAddress
// runs once per device (with it current) and the address is kept per device.
// Returns 0 or a HIP error code.
int noalloc_scratch(void** ptr);

// ---- batched march (gs_march_many): per-run descriptors in device memory
struct NLRun {                  // the NL RK4 step of one run
    const void* field;
    void* b1;
    double* linemax;
    NLArgs a;                   // a.ctl: the run's controller
};
struct GemmRun {                // one QDHT of one run's step (X -> Y, the linear step fused when la.fk)
    const void* A;
    const void* X;
    void* Y;
    int K, NC;                  // padded rows, columns
    int b0, nb;                 // the run's first block in the launch (a multiple of 8), its block count
    LstepArgs la;               // la.ctl: the run's controller
    QCols qc;
};
struct CtlRun {
    int nr;
    const double* linemax;
    DevStatus* status;
    DevCtl* ctl;
};

// ---- launchers (all asynchronous on `st`) ----
// RK4 NL step over nr lines of field -> b1; per-line maxima into linemax[2*r]
int launch_nl_rk4(Prec p, int nt, const void* field, void* b1, double dz, const NLArgs& a, double* linemax, hipStream_t st);
// one RHS evaluation per line (calculate_NLresponse)
int launch_nl_response(Prec p, int nt, const void* in, void* out, int nlines, const NLArgs& a, hipStream_t st);
// batched line FFT, optional scale and spectral filter (double per omega, may be null)
int launch_fft_lines(Prec p, int nt, const void* in, void* out, int nlines, int sign, double scale,
                     const double* wfilter, const void* tw, hipStream_t st);
// per-line diagnostics of a spectral buffer: flux, maxI (with filter), rhomax (unfiltered);
// optional E(t) rows (et, scaled 1/T, unfiltered) and rho(t) rows (rho, double)
int launch_line_diag(Prec p, int nt, const void* field, int nlines, const double* wfilter, double tstep,
                     const NLArgs& a, double* flux, double* maxI, double* rhomax, void* et, double* rho, hipStream_t st);
// QDHT GEMM: out = A * in, A = diag(m1) C diag(1/m1) padded to rp x rp.
// amat: FP64/FP32 transposed matrix, or for MIXED the bf16 hi/lo split in
// MFMA fragment order (see the plan set-up).
int launch_qdht(Prec p, int rp, int nt, const void* amat, const void* in, void* out,
                const LstepArgs* lstep, hipStream_t st, const DevCtl* gate = nullptr, const QCols* cols = nullptr);
// BIGBUFFER2 <-> FIELD_k: out = masked ? b2 : fk (the full B2), or fk = masked ? 0 : b2
int launch_b2_merge(Prec p, const void* fk, const void* b2, void* out, const void* gfac, size_t n_complex, hipStream_t st);
int launch_b2_split(Prec p, const void* b2, void* fk, const void* gfac, size_t n_complex, hipStream_t st);
// batched phases of gs_march_many: the NL RK4 step of items[b] = (run, line) for runs
// of one precision, line length and line shape (nl_e); one QDHT of every run of a
// precision (brun[b]: the run of block b); the controllers (one block per run) and
// the bookkeeping after the linear half of n runs
int launch_nl_rk4_batch(Prec p, int nt, int nl_e, bool mfma_fft, const NLRun* runs, const int2* items, int nitems,
                        hipStream_t st);
int launch_qdht_batch(Prec p, bool lstep, const GemmRun* runs, const int* brun, int nblocks, hipStream_t st, int ks = 0);
int launch_controller_batch(const CtlRun* runs, int n, hipStream_t st);
int launch_finish_batch(DevCtl* ctls, int n, hipStream_t st);
// blocks of one run's QDHT in a batched launch (the single-run grid of launch_qdht)
int qdht_blocks(Prec p, int rp, int nt, const QCols* cols, bool back, bool gated);
// element-wise linear step from a k-space increment (calculate_Lstep)
int launch_lstep(Prec p, int rp, int nt, const void* b1k, const LstepArgs& la, hipStream_t st);
// max of the per-line maxima -> status
int launch_reduce_max(int nr, const double* linemax, DevStatus* st_out, hipStream_t st);
// device-driven march: maxima -> accept / reject, next dz, clamp to the plane (controller),
// and the bookkeeping after the linear half (finish)
int launch_controller(int nr, const double* linemax, DevStatus* status, DevCtl* ctl, hipStream_t st);
int launch_finish(DevCtl* ctl, hipStream_t st);
// conversions between the host FP64 layout and the device field precision
int launch_convert(Prec p, const double* src64, void* dst, size_t n_complex, int to_device, hipStream_t st);
// reverse mode (gs_adjoint.hip).  Transposed RK4 NL step over a.nr lines: gfield =
// J^H gb1 at FIELD = field and step dz; stage holds three planes of `plane`
// complex elements (the recomputed stage inputs)
int launch_nl_rk4_vjp(Prec p, int nt, const void* field, const void* gb1, void* gfield, void* stage, size_t plane,
                      double dz, const NLArgs& a, hipStream_t st);
// transposed linear step at la.dz: gb2 <- conj(P) (gb2 + gfk), gb1k <- dz g gb2
// (masked points: gb2 passes through, gb1k = 0)
int launch_lstep_adj(Prec p, int rp, int nt, const void* gfk, void* gb2, void* gb1k, const LstepArgs& la, hipStream_t st);
// y += x over n complex values in the field precision
int launch_axpy(Prec p, void* y, const void* x, size_t n_complex, hipStream_t st);
// average spectrum (dump_types.cpp:429-457): out[w] = XMAX^2/2 sum_r |F_rw|^2 x_r (x_r - x_{r-1})
int launch_avg_spectrum(Prec p, int nr, int nt, const void* field, const double* wts, double* out, hipStream_t st);

}  // namespace gs
