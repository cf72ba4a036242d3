// gs_line.hip -- per-radial-line kernels of the z-step: the fused RK4 of the
// nonlinear RHS (propagator.cu:376-425 around calculate_NLresponse,
// propagator.cu:629-783), a single RHS evaluation, batched line FFTs and the
// dump-side line diagnostics.  The line layout, FFTs, plasma scan and RHS
// are in gs_line.h.
#include "gs_line.h"

#include <mutex>

// The allocation-free entry point's staging area: one copy per GPU, held until
// the process ends.  Host code never passes this name to a kernel or a copy
// (in host code it is a stand-in, not a device address); it uses the address
// noalloc_scratch() returns for the current device.
static __device__ __align__(256) unsigned char scratch[gs::kNoAllocScratchBytes];

namespace gs {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch(void** ptr)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return (int)hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return 0;
}

// FP32 field lines keep the RK4 input row A and the k1 + 2k2 + 2k3 + k4
// accumulator in LDS (2 x 32 KiB at N_T = 4096, after LineSmem); FP64 lines
// do not fit and use the global rows.
template <class Pol, int T, int E>
struct RK4Rows {
    static constexpr size_t row_bytes = 2 * (size_t)T * sizeof(C2<typename Pol::R>);
    static constexpr bool in_lds = sizeof(typename Pol::R) == 4 && sizeof(LineSmem<Pol, T, E>) + row_bytes <= 160 * 1024;
    static constexpr size_t bytes = in_lds ? row_bytes : 0;
};

// RK4 NL step of one radial line (the body of k_nl_rk4 and k_nl_rk4_batch)
template <class Pol, int T, int E>
__device__ __forceinline__ void nl_rk4_line(const C2<typename Pol::R>* __restrict__ field, C2<typename Pol::R>* __restrict__ b1,
                                            double dz, const NLArgs& a, double* __restrict__ linemax, int line,
                                            unsigned char* smem_raw)
{
    using R = typename Pol::R;
    constexpr int NT = Shape<Pol, T, E>::NT;
    constexpr bool ROWS_LDS = RK4Rows<Pol, T, E>::in_lds;
    LineSmem<Pol, T, E>& sm = *reinterpret_cast<LineSmem<Pol, T, E>*>(smem_raw);
    const int tid = threadIdx.x;
    const C2<R>* Ag = field + (size_t)line * T;
    C2<R>* b1row = b1 + (size_t)line * T;
    // A and the accumulator: LDS rows (FP32) or the global rows (FP64, the output row accumulates)
    C2<R>* Arow = ROWS_LDS ? reinterpret_cast<C2<R>*>(smem_raw + sizeof(LineSmem<Pol, T, E>)) : const_cast<C2<R>*>(Ag);
    C2<R>* acc = ROWS_LDS ? Arow + T : b1row;
    int flags = 0;
    double mA = 0, mN = 0;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const C2<R> av = Ag[tid + r * NT];
        if constexpr (ROWS_LDS) Arow[tid + r * NT] = av;
        mA = fmax(mA, (double)av.x * av.x + (double)av.y * av.y);
    }
    C2<R> k[E];
#pragma unroll
    for (int r = 0; r < E; ++r) k[r] = C2<R>{0, 0};
    const R rdz = (R)dz;
    // classical RK4 without linear propagation between stages: stage inputs
    // A, A + dz/2 k1, A + dz/2 k2, A + dz k3; B1 = (k1 + 2 k2 + 2 k3 + k4)/6.
    // Each thread reads and writes only its own entries of A / acc.
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
        const R c = s == 0 ? (R)0 : (s == 3 ? rdz : rdz * (R)0.5);
        const R wk = (s == 1 || s == 2) ? (R)2 : (R)1;
        const R wold = s == 0 ? (R)0 : (R)1;
        C2<R> xin[E];
#pragma unroll
        for (int r = 0; r < E; ++r) xin[r] = cadd(Arow[tid + r * NT], cscale(k[r], c));
        nl_rhs<Pol, T, E>(xin, k, sm, a, &flags);
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const int o = tid + r * NT;
            const C2<R> old = s == 0 ? C2<R>{0, 0} : acc[o];
            acc[o] = cadd(cscale(old, wold), cscale(k[r], wk));
        }
    }
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int o = tid + r * NT;
        const C2<R> t = acc[o];
        const C2<R> v = {t.x / (R)6, t.y / (R)6};
        b1row[o] = v;
        mN = fmax(mN, (double)v.x * v.x + (double)v.y * v.y);
    }
    block_max2<NT>(mA, mN, sm.red);
    if (tid == 0) {
        linemax[2 * line] = mA;
        linemax[2 * line + 1] = mN;
    }
    report_flags(a.status, flags);
}

template <class Pol, int T, int E>
__global__ __launch_bounds__((Shape<Pol, T, E>::NT), GS_NL_WAVES) void k_nl_rk4(const C2<typename Pol::R>* __restrict__ field,
                                                                    C2<typename Pol::R>* __restrict__ b1, double dz,
                                                                    NLArgs a, double* __restrict__ linemax)
{
    if (a.ctl) {
        if (a.ctl->state != 0) return;
        dz = a.ctl->dz;
    }
    extern __shared__ __align__(16) unsigned char smem_raw[];
    nl_rk4_line<Pol, T, E>(field, b1, dz, a, linemax, blockIdx.x, smem_raw);
}

// the NL step of several runs in one launch: block b is line items[b].y of run items[b].x
template <class Pol, int T, int E>
__global__ __launch_bounds__((Shape<Pol, T, E>::NT), GS_NL_WAVES) void k_nl_rk4_batch(const NLRun* __restrict__ runs,
                                                                                 const int2* __restrict__ items)
{
    using R = typename Pol::R;
    const int2 it = items[blockIdx.x];
    const NLRun& r = runs[it.x];
    if (r.a.ctl->state != 0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    nl_rk4_line<Pol, T, E>((const C2<R>*)r.field, (C2<R>*)r.b1, r.a.ctl->dz, r.a, r.linemax, it.y, smem_raw);
}

template <class Pol, int T, int E>
__global__ __launch_bounds__((Shape<Pol, T, E>::NT)) void k_nl_response(const C2<typename Pol::R>* __restrict__ in,
                                                                         C2<typename Pol::R>* __restrict__ out, NLArgs a)
{
    using R = typename Pol::R;
    constexpr int NT = Shape<Pol, T, E>::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    LineSmem<Pol, T, E>& sm = *reinterpret_cast<LineSmem<Pol, T, E>*>(smem_raw);
    const int line = blockIdx.x, tid = threadIdx.x;
    int flags = 0;
    C2<R> x[E], y[E];
#pragma unroll
    for (int r = 0; r < E; ++r) x[r] = in[(size_t)line * T + tid + r * NT];
    nl_rhs<Pol, T, E>(x, y, sm, a, &flags);
#pragma unroll
    for (int r = 0; r < E; ++r) out[(size_t)line * T + tid + r * NT] = y[r];
    report_flags(a.status, flags);
}

// ------------------------------------------------------------------ line FFT / diagnostics
// E consecutive complex values (one chunk) from / to global memory, as 16-byte
// accesses (the chunk starts on a 16-byte boundary)
template <class R, int E>
__device__ __forceinline__ void load_chunk(const C2<R>* __restrict__ p, C2<R> (&x)[E])
{
    if constexpr (sizeof(R) == 4) {
        const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
        for (int i = 0; i < E / 2; ++i) {
            const float4 v = q[i];
            x[2 * i] = C2<R>{v.x, v.y};
            x[2 * i + 1] = C2<R>{v.z, v.w};
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) x[e] = p[e];
    }
}
template <class R, int E>
__device__ __forceinline__ void store_chunk(C2<R>* __restrict__ p, const C2<R> (&x)[E])
{
    if constexpr (sizeof(R) == 4) {
        float4* q = reinterpret_cast<float4*>(p);
#pragma unroll
        for (int i = 0; i < E / 2; ++i) q[i] = float4{x[2 * i].x, x[2 * i].y, x[2 * i + 1].x, x[2 * i + 1].y};
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) p[e] = x[e];
    }
}

template <class R, int T>
struct FftShape {
    static constexpr int E = T >= 1024 ? 8 : (T >= 512 ? 8 : 4);
    static constexpr int NT = T / E;
};

// batched FFT of nlines lines (natural order in and out), optional scale and filter
template <class R, int T>
__global__ __launch_bounds__((FftShape<R, T>::NT)) void k_fft_lines(const C2<R>* __restrict__ in, C2<R>* __restrict__ out,
                                                                  int sign, double scale, const double* __restrict__ filt,
                                                                  const C2<R>* __restrict__ tw)
{
    constexpr int E = FftShape<R, T>::E, NT = FftShape<R, T>::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<R>* b0 = reinterpret_cast<C2<R>*>(smem_raw);
    C2<R>* b1 = b0 + (T + T / E);
    const size_t base = (size_t)blockIdx.x * T;
    const int tid = threadIdx.x;
    C2<R> x[E], y[E];
    if (sign < 0) {
        // chunk in -> strided out
        R f[E];
        load_chunk<R, E>(in + base + (size_t)tid * E, x);
#pragma unroll
        for (int e = 0; e < E; ++e) f[e] = (R)(scale * (filt ? filt[tid * E + e] : 1.0));
#pragma unroll
        for (int r = 0; r < E; ++r) y[r] = C2<R>{0, 0};
        fft_fwd_acc<0, R, T, NT, E>(x, f, y, b0, b1, tw, nullptr);
#pragma unroll
        for (int r = 0; r < E; ++r) out[base + tid + r * NT] = y[r];
    } else {
        // strided in -> chunk out
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const int n = tid + r * NT;
            y[r] = cscale(in[base + n], (R)(scale * (filt ? filt[n] : 1.0)));
        }
        fft_inv<R, T, NT, E>(y, x, b0, b1, tw);
        store_chunk<R, E>(out + base + (size_t)tid * E, x);
    }
}

template <class Pol, int T, int E>
__global__ __launch_bounds__((Shape<Pol, T, E>::NT)) void k_line_diag(const C2<typename Pol::R>* __restrict__ field,
                                                                 const double* __restrict__ filt, double tstep, NLArgs a,
                                                                 double* flux, double* maxI, double* rhomax,
                                                                 C2<typename Pol::R>* et, double* rho_out)
{
    using R = typename Pol::R;
    using S = typename Pol::S;
    using W = typename Pol::W;
    constexpr int NT = Shape<Pol, T, E>::NT, NW = Shape<Pol, T, E>::NW;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    LineSmem<Pol, T, E>& sm = *reinterpret_cast<LineSmem<Pol, T, E>*>(smem_raw);
    const C2<R>* tw = (const C2<R>*)a.tw;
    const int line = blockIdx.x, tid = threadIdx.x, n0 = tid * E;
    const size_t base = (size_t)line * T;
    int flags = 0;
    C2<R> y[E], Ev[E];
#pragma unroll
    for (int r = 0; r < E; ++r) y[r] = field[base + tid + r * NT];
    fft_inv<R, T, NT, E>(y, Ev, sm.b0, sm.b1, tw);
    S Wf[E], cf[E];
    S rho[E];
    double fl = 0, mi = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        Ev[e] = cscale(Ev[e], (R)a.inv_nt);
        const W I = (W)Ev[e].x * (W)Ev[e].x + (W)Ev[e].y * (W)Ev[e].y;
        Wf[e] = (S)photoionization<W>(I, a.ionrate, &flags);
        cf[e] = (S)((W)a.sigma / ((W)a.U + (W)a.pond * I) * I);
        fl += (double)I;
        mi = fmax(mi, (double)I);
        if (et) et[base + n0 + e] = Ev[e];
    }
    plasma_scan2<S, E, NT>(Wf, cf, rho, a, scratch_of(sm), &flags);
    double rm = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        rm = fmax(rm, (double)rho[e]);
        if (rho_out) rho_out[base + n0 + e] = (double)rho[e];
    }
    if (filt) {
        // flux / maxI of the filtered field (dump_types.cpp:123-165)
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const int n = tid + r * NT;
            y[r] = cscale(field[base + n], (R)(filt[n] * a.inv_nt));
        }
        fft_inv<R, T, NT, E>(y, Ev, sm.b0, sm.b1, tw);
        fl = 0;
        mi = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const double I = (double)Ev[e].x * Ev[e].x + (double)Ev[e].y * Ev[e].y;
            fl += I;
            mi = fmax(mi, I);
        }
    }
    // flux: block sum; maxI, rhomax: block max
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) fl += __shfl_xor(fl, d);
    if ((tid & 63) == 0) sm.pW[tid >> 6] = fl;
    block_max2<NT>(mi, rm, sm.red);
    if (tid == 0) {
        double s = 0;
        for (int w = 0; w < NW; ++w) s += sm.pW[w];
        if (flux) flux[line] = s * tstep;
        if (maxI) maxI[line] = mi;
        if (rhomax) rhomax[line] = rm;
    }
    report_flags(a.status, flags);
}

// ------------------------------------------------------------------ launchers
template <class Pol, int T, int E>
static int nl_rk4_TE(const void* field, void* b1, double dz, const NLArgs& a, double* linemax, hipStream_t st)
{
    using R = typename Pol::R;
    const size_t sm = sizeof(LineSmem<Pol, T, E>) + RK4Rows<Pol, T, E>::bytes;
    auto kf = k_nl_rk4<Pol, T, E>;
    HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(kf, dim3(a.nr), dim3(Shape<Pol, T, E>::NT), sm, st, (const C2<R>*)field, (C2<R>*)b1, dz, a, linemax);
    HIPCHK(hipGetLastError());
    return 0;
}

// samples per thread: the default, or 4/8/16 when a.nl_e asks for it (FP32/mixed, N_T >= 1024)
template <class Pol, int T>
static int nl_rk4_T(const void* field, void* b1, double dz, const NLArgs& a, double* linemax, hipStream_t st)
{
    if constexpr (sizeof(typename Pol::R) == 4 && T >= 1024) {
        if (a.nl_e == 4) return nl_rk4_TE<Pol, T, 4>(field, b1, dz, a, linemax, st);
        if (a.nl_e == 8) return nl_rk4_TE<Pol, T, 8>(field, b1, dz, a, linemax, st);
        if (a.nl_e == 16) return nl_rk4_TE<Pol, T, 16>(field, b1, dz, a, linemax, st);
    }
    return nl_rk4_TE<Pol, T, DefaultE<Pol, T>::v>(field, b1, dz, a, linemax, st);
}

template <class Pol, int T>
static int nl_resp_T(const void* in, void* out, int nlines, const NLArgs& a, hipStream_t st)
{
    using R = typename Pol::R;
    constexpr int E = DefaultE<Pol, T>::v;
    const size_t sm = sizeof(LineSmem<Pol, T, E>);
    auto kf = k_nl_response<Pol, T, E>;
    HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(kf, dim3(nlines), dim3(Shape<Pol, T, E>::NT), sm, st, (const C2<R>*)in, (C2<R>*)out, a);
    HIPCHK(hipGetLastError());
    return 0;
}

template <class Pol, int T>
static int diag_T(const void* field, int nlines, const double* filt, double tstep, const NLArgs& a, double* flux,
                  double* maxI, double* rhomax, void* et, double* rho, hipStream_t st)
{
    using R = typename Pol::R;
    constexpr int E = DefaultE<Pol, T>::v;
    const size_t sm = sizeof(LineSmem<Pol, T, E>);
    auto kf = k_line_diag<Pol, T, E>;
    HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(kf, dim3(nlines), dim3(Shape<Pol, T, E>::NT), sm, st, (const C2<R>*)field, filt, tstep, a, flux,
                       maxI, rhomax, (C2<R>*)et, rho);
    HIPCHK(hipGetLastError());
    return 0;
}

template <class R, int T>
static int fft_T(const void* in, void* out, int nlines, int sign, double scale, const double* filt, const void* tw,
                 hipStream_t st)
{
    constexpr int E = FftShape<R, T>::E;
    const size_t sm = 2 * sizeof(C2<R>) * (T + T / E);
    auto kf = k_fft_lines<R, T>;
    HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(kf, dim3(nlines), dim3(FftShape<R, T>::NT), sm, st, (const C2<R>*)in, (C2<R>*)out, sign, scale,
                       filt, (const C2<R>*)tw);
    HIPCHK(hipGetLastError());
    return 0;
}


int launch_nl_rk4(Prec p, int nt, const void* field, void* b1, double dz, const NLArgs& a, double* linemax, hipStream_t st)
{
    if (p == kFP64) GS_T_SWITCH(nt, (nl_rk4_T<PolF64, T_>(field, b1, dz, a, linemax, st)));
    if (p == kFP32) GS_T_SWITCH(nt, (nl_rk4_T<PolF32, T_>(field, b1, dz, a, linemax, st)));
    if (a.mfma_fwd || a.mfma_inv)
        GS_T_SWITCH(nt, (nl_rk4_TE<PolMixMfma, T_, DefaultE<PolMixMfma, T_>::v>(field, b1, dz, a, linemax, st)));
    GS_T_SWITCH(nt, (nl_rk4_T<PolMix, T_>(field, b1, dz, a, linemax, st)));
}

template <class Pol, int T, int E>
static int nl_rk4_batch_TE(const NLRun* runs, const int2* items, int nitems, hipStream_t st)
{
    const size_t sm = sizeof(LineSmem<Pol, T, E>) + RK4Rows<Pol, T, E>::bytes;
    auto kf = k_nl_rk4_batch<Pol, T, E>;
    HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(kf, dim3(nitems), dim3(Shape<Pol, T, E>::NT), sm, st, runs, items);
    HIPCHK(hipGetLastError());
    return 0;
}

// the line shape k_nl_rk4 uses for this (policy, N_T, nl_e): as nl_rk4_T
template <class Pol, int T>
static int nl_rk4_batch_T(int nl_e, const NLRun* runs, const int2* items, int nitems, hipStream_t st)
{
    if constexpr (sizeof(typename Pol::R) == 4 && T >= 1024) {
        if (nl_e == 4) return nl_rk4_batch_TE<Pol, T, 4>(runs, items, nitems, st);
        if (nl_e == 8) return nl_rk4_batch_TE<Pol, T, 8>(runs, items, nitems, st);
        if (nl_e == 16) return nl_rk4_batch_TE<Pol, T, 16>(runs, items, nitems, st);
    }
    return nl_rk4_batch_TE<Pol, T, DefaultE<Pol, T>::v>(runs, items, nitems, st);
}

int launch_nl_rk4_batch(Prec p, int nt, int nl_e, bool mfma_fft, const NLRun* runs, const int2* items, int nitems,
                        hipStream_t st)
{
    if (p == kFP64) GS_T_SWITCH(nt, (nl_rk4_batch_T<PolF64, T_>(nl_e, runs, items, nitems, st)));
    if (p == kFP32) GS_T_SWITCH(nt, (nl_rk4_batch_T<PolF32, T_>(nl_e, runs, items, nitems, st)));
    if (mfma_fft)
        GS_T_SWITCH(nt, (nl_rk4_batch_TE<PolMixMfma, T_, DefaultE<PolMixMfma, T_>::v>(runs, items, nitems, st)));
    GS_T_SWITCH(nt, (nl_rk4_batch_T<PolMix, T_>(nl_e, runs, items, nitems, st)));
}

int launch_nl_response(Prec p, int nt, const void* in, void* out, int nlines, const NLArgs& a, hipStream_t st)
{
    if (p == kFP64) GS_T_SWITCH(nt, (nl_resp_T<PolF64, T_>(in, out, nlines, a, st)));
    if (p == kFP32) GS_T_SWITCH(nt, (nl_resp_T<PolF32, T_>(in, out, nlines, a, st)));
    if (a.mfma_fwd || a.mfma_inv) GS_T_SWITCH(nt, (nl_resp_T<PolMixMfma, T_>(in, out, nlines, a, st)));
    GS_T_SWITCH(nt, (nl_resp_T<PolMix, T_>(in, out, nlines, a, st)));
}

int launch_fft_lines(Prec p, int nt, const void* in, void* out, int nlines, int sign, double scale, const double* wfilter,
                     const void* tw, hipStream_t st)
{
    if (p == kFP64) GS_T_SWITCH(nt, (fft_T<double, T_>(in, out, nlines, sign, scale, wfilter, tw, st)));
    GS_T_SWITCH(nt, (fft_T<float, T_>(in, out, nlines, sign, scale, wfilter, tw, st)));
}

int launch_line_diag(Prec p, int nt, const void* field, int nlines, const double* wfilter, double tstep, const NLArgs& a,
                     double* flux, double* maxI, double* rhomax, void* et, double* rho, hipStream_t st)
{
    if (p == kFP64) GS_T_SWITCH(nt, (diag_T<PolF64, T_>(field, nlines, wfilter, tstep, a, flux, maxI, rhomax, et, rho, st)));
    if (p == kFP32) GS_T_SWITCH(nt, (diag_T<PolF32, T_>(field, nlines, wfilter, tstep, a, flux, maxI, rhomax, et, rho, st)));
    GS_T_SWITCH(nt, (diag_T<PolMix, T_>(field, nlines, wfilter, tstep, a, flux, maxI, rhomax, et, rho, st)));
}

}  // namespace gs
