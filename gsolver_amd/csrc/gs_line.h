// gs_line.h -- building blocks of the per-radial-line kernels (gs_line.hip,
// gs_adjoint.hip): line shapes, LDS Stockham FFTs, the plasma scan and the
// nonlinear RHS of one line (calculate_NLresponse, propagator.cu:629-783).
//
// One workgroup owns one radial line of N_T samples, E samples per thread.
// Time-domain data live in "chunk" layout (thread j holds t = jE..jE+E-1,
// which the plasma and Raman scans need), spectral data in the "strided"
// layout (thread j holds omega = j + NT*r), which is exactly where a Stockham
// FFT leaves its output after a final radix-E pass and where an inverse
// Stockham FFT wants its input for a first radix-E pass.  So spectral data
// never round-trip through LDS: the forward FFT ends in registers and the
// inverse starts from them.  Passes ping-pong between two padded LDS buffers
// (index n + n/E, which keeps the chunk accesses free of bank conflicts), one
// barrier per pass.
//
// The plasma density (ionization.cpp:9-95) is a nonlinear recurrence in t;
// it is solved by Newton iterations on the whole line, each an affine
// block scan of the linearised Heun maps, until the update is at rounding
// level, with a chunk-sequential exact fallback.  Lines without any
// ionization skip it (rho stays at its initial value 0).  The Raman
// convolution (propagator.cu:696-703, circular) is evaluated in closed form
// from one complex affine scan: y = (1-fR) I + fR M Im(s (1 - z^T) + z^{n+1} s_{T-1}).
#pragma once

#include "gs_dev.h"

#include <type_traits>

#ifndef GS_NL_E32
#define GS_NL_E32 8    // samples per thread of the FP32/mixed line kernels for N_T >= 1024
#endif
#ifndef GS_SCAN_TOL_F32OUT
#define GS_SCAN_TOL_F32OUT 4e-15   // Newton stop of the FP64 plasma scan when rho is stored as FP32
#endif
#ifndef GS_NL_WAVES
#define GS_NL_WAVES 1  // minimum waves per SIMD the RK4 kernel is register-allocated for
#endif

// launcher dispatch over the supported line lengths (T_ the compile-time N_T)
#define GS_T_SWITCH(nt, CALL)                                   \
    switch (nt) {                                               \
    case 256: { constexpr int T_ = 256; return CALL; }          \
    case 512: { constexpr int T_ = 512; return CALL; }          \
    case 1024: { constexpr int T_ = 1024; return CALL; }        \
    case 2048: { constexpr int T_ = 2048; return CALL; }        \
    case 4096: { constexpr int T_ = 4096; return CALL; }        \
    default: return -1000;                                      \
    }

namespace gs {

template <class Pol, int T>
struct DefaultE {
    using R = typename Pol::R;
    static constexpr int v = T >= 1024 ? (sizeof(R) == 8 ? 8 : GS_NL_E32) : (T >= 512 ? 8 : 4);
};

template <class Pol, int T, int E_ = DefaultE<Pol, T>::v>
struct Shape {
    using R = typename Pol::R;
    static constexpr int E = E_;
    static constexpr int NT = T / E;
    static constexpr int NW = NT / 64;
    static constexpr int PADN = T + T / E;     // padded LDS buffer length
    static_assert(NT >= 64 && NT <= 1024, "unsupported N_T");
};

__device__ __forceinline__ int pidx(int n, int E) { return n + n / E; }

// largest radix <= E (so every thread has at least one butterfly) that divides M
template <int M, int E> struct PickRadix {
    static constexpr int v = (E >= 16 && M % 16 == 0) ? 16 : (E >= 8 && M % 8 == 0) ? 8 : (M % 4 == 0) ? 4 : 2;
};

// Twiddles of a pass of span L = Ns RADIX: v[r] *= w_L^{r k}, w_L = e^{-2 pi i/L}
// (conjugated for SIGN > 0), from the one table entry w_L^k = tw[k T/L] by the
// product tree of twiddle_pow.  (A precomputed per-pass table in lane order,
// one 64-byte load per butterfly instead of six complex products, measured
// slower: the line kernels run two waves per SIMD and do not hide the loads.)
template <class R, int T, int RADIX, int SIGN>
__device__ __forceinline__ void twiddle_apply(C2<R>* v, const C2<R>* __restrict__ tw, int L, int k)
{
    C2<R> w = tw[k * (T / L)];
    if (SIGN > 0) w.y = -w.y;
    twiddle_pow<RADIX>(v, w);
}

// One Stockham pass src -> dst (both padded LDS buffers), no barrier.
template <class R, int T, int NT, int E, int RADIX, int SIGN>
__device__ __forceinline__ void spass(const C2<R>* __restrict__ src, C2<R>* __restrict__ dst,
                                      const C2<R>* __restrict__ tw, int Ns)
{
    constexpr int NB = T / RADIX;
    constexpr int PER = (NB + NT - 1) / NT;
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = threadIdx.x + b * NT;
        if (NB % NT == 0 || j < NB) {
            C2<R> v[RADIX];
#pragma unroll
            for (int r = 0; r < RADIX; ++r) v[r] = src[pidx(j + r * NB, E)];
            const int k = j & (Ns - 1);
            if (Ns > 1) twiddle_apply<R, T, RADIX, SIGN>(v, tw, Ns * RADIX, k);
            dft_small<RADIX, SIGN>(v);
            const int d = (j - k) * RADIX + k;
#pragma unroll
            for (int r = 0; r < RADIX; ++r) dst[pidx(d + r * Ns, E)] = v[r];
        }
    }
}

// The passes that cover M more points, ping-ponging b0 <-> b1 starting from
// buffer P; PassChain<...>::out is the buffer holding the result.
template <class R, int T, int NT, int E, int M, int SIGN, int P>
struct PassChain {
    static constexpr int RAD = PickRadix<M, E>::v;
    using Next = PassChain<R, T, NT, E, M / RAD, SIGN, 1 - P>;
    static constexpr int out = Next::out;
    static __device__ __forceinline__ void run(C2<R>* b0, C2<R>* b1, const C2<R>* tw, int Ns)
    {
        spass<R, T, NT, E, RAD, SIGN>(P ? b1 : b0, P ? b0 : b1, tw, Ns);
        __syncthreads();
        Next::run(b0, b1, tw, Ns * RAD);
    }
};
template <class R, int T, int NT, int E, int SIGN, int P>
struct PassChain<R, T, NT, E, 1, SIGN, P> {
    static constexpr int out = P;
    static __device__ __forceinline__ void run(C2<R>*, C2<R>*, const C2<R>*, int) {}
};

// forward (SIGN -1) FFT of the chunk-layout signal fac[e]*x[e]; the spectrum
// (strided layout) times a per-omega factor G is added to out, with G by MODE:
//   0: G = 1     1: G = M[omega] (complex table)     2: G = -i K[omega] (real table)
template <int MODE, class R, int T, int NT, int E, class F>
__device__ __forceinline__ void fft_fwd_acc(const C2<R> (&x)[E], const F (&fac)[E], C2<R> (&out)[E], C2<R>* b0,
                                            C2<R>* b1, const C2<R>* tw, const void* table)
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < E; ++e) b0[pidx(tid * E + e, E)] = cscale(x[e], (R)fac[e]);
    __syncthreads();
    using Chain = PassChain<R, T, NT, E, T / E, -1, 0>;
    Chain::run(b0, b1, tw, 1);
    const C2<R>* src = Chain::out ? b1 : b0;
    // last pass: radix E at Ns = T/E = NT, butterfly j = tid, output omega = tid + NT r
    C2<R> y[E];
#pragma unroll
    for (int r = 0; r < E; ++r) y[r] = src[pidx(tid + r * NT, E)];
    twiddle_apply<R, T, E, -1>(y, tw, T, tid);
    dft_small<E, -1>(y);
#pragma unroll
    for (int r = 0; r < E; ++r) {
        C2<R> g;
        if constexpr (MODE == 0) {
            g = y[r];
        } else if constexpr (MODE == 1) {
            g = cmul(((const C2<R>*)table)[tid + r * NT], y[r]);
        } else {
            const R k = ((const R*)table)[tid + r * NT];
            g = C2<R>{k * y[r].y, -k * y[r].x};
        }
        out[r] = cadd(out[r], g);
    }
    // the next transform writes b0 first and passes a barrier before it writes
    // b1, so only a result left in b0 needs a barrier here
    if constexpr (Chain::out == 0) __syncthreads();
}

// inverse (SIGN +1) FFT, unnormalised: spectrum in strided layout -> time in chunk layout
template <class R, int T, int NT, int E>
__device__ __forceinline__ void fft_inv(const C2<R> (&y)[E], C2<R> (&x)[E], C2<R>* b0, C2<R>* b1, const C2<R>* tw)
{
    const int tid = threadIdx.x;
    // first pass: radix E at Ns = 1 over j = tid, inputs tid + NT r (registers), outputs tid*E + r
    C2<R> v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = y[r];
    dft_small<E, +1>(v);
#pragma unroll
    for (int r = 0; r < E; ++r) b0[pidx(tid * E + r, E)] = v[r];
    __syncthreads();
    using Chain = PassChain<R, T, NT, E, T / E, +1, 0>;
    Chain::run(b0, b1, tw, E);
    const C2<R>* src = Chain::out ? b1 : b0;
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = src[pidx(tid * E + e, E)];
    if constexpr (Chain::out == 0) __syncthreads();   // as in fft_fwd_acc
}

// ------------------------------------------------------------------ MFMA FFT
// Mixed policy: the line FFTs run their radix-16 passes on the matrix cores.
// One pass maps the 16 inputs of each butterfly group j (src[j + r T/16],
// twiddled by w^{r k}, k = j mod Ns) to its 16 outputs Y_q = sum_r x_r
// e^{SIGN 2 pi i q r/16} as one real GEMM: [Re Y; Im Y] (32 rows) = [[C, -S],
// [S, C]] (32 x 32, the k columns holding Re x then Im x) times 32 groups,
// v_mfma_f32_32x32x16_bf16 with both operands split into bf16 hi + lo and the
// three products lo*hi, hi*lo, hi*hi accumulated in FP32 (the operator
// fragments come from the plan, dft16).  Each wave owns 32 groups: lane l
// holds group 32 w + l%32 and inputs r = 8 (l/32) .. +7 (the B operand
// layout); its results are Y_q for q in {4h..4h+3, 8+4h..8+4h+3}, h = l/32.
// A transform is good to ~5e-6 relative (FP32 VALU: ~1e-7).  The forward
// transforms of the RHS terms (loss, rho E, Kerr/Raman product) enter N
// linearly.  The inverse transform's E(t) also feeds W(I) ~ I^6, so it runs
// on the matrix cores only for lines whose intensity stays below the
// ionization threshold; other lines transform again in FP32.
// twp: this pass's twiddles w^{r k} in lane order, [group j][r] (null at Ns = 1)
template <int T, int E, int SIGN>
__device__ __forceinline__ void mfma_pass16(const C2<float>* __restrict__ src, C2<float>* __restrict__ dst,
                                            const C2<float>* __restrict__ twp, int Ns, const v8bf (&fa)[2][2])
{
    constexpr int NG = T / 16, NW = T / E / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, h = lane >> 5;
    for (int g0 = wv * 32; g0 < NG; g0 += NW * 32) {     // wave-uniform: MFMA needs every lane
        const int j = g0 + (lane & 31);
        const int k = j & (Ns - 1);
        float4 tq[4];
        if (Ns > 1) {
            const float4* tp = reinterpret_cast<const float4*>(twp + (size_t)j * 16 + 8 * h);
#pragma unroll
            for (int i = 0; i < 4; ++i) tq[i] = tp[i];
        }
        float re[8], im[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            C2<float> v = src[pidx(j + NG * (8 * h + i), E)];
            if (Ns > 1) {
                const float4 t = tq[i >> 1];
                C2<float> w = (i & 1) ? C2<float>{t.z, t.w} : C2<float>{t.x, t.y};
                if (SIGN > 0) w.y = -w.y;
                v = cmul(v, w);
            }
            re[i] = v.x;
            im[i] = v.y;
        }
        uint32_t rh[4], rl[4], ih[4], il[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            split2(re[2 * i], re[2 * i + 1], &rh[i], &rl[i]);
            split2(im[2 * i], im[2 * i + 1], &ih[i], &il[i]);
        }
        typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
        const v8bf brh = __builtin_bit_cast(v8bf, u32x4_t{rh[0], rh[1], rh[2], rh[3]});
        const v8bf brl = __builtin_bit_cast(v8bf, u32x4_t{rl[0], rl[1], rl[2], rl[3]});
        const v8bf bih = __builtin_bit_cast(v8bf, u32x4_t{ih[0], ih[1], ih[2], ih[3]});
        const v8bf bil = __builtin_bit_cast(v8bf, u32x4_t{il[0], il[1], il[2], il[3]});
        // two independent chains (the Re x and Im x halves of K), summed at the end
        v16f a0, a1;
#pragma unroll
        for (int e = 0; e < 16; ++e) a0[e] = a1[e] = 0.f;
        a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0][1], brh, a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[1][1], bih, a1, 0, 0, 0);
        a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0][0], brl, a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[1][0], bil, a1, 0, 0, 0);
        a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0][0], brh, a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[1][0], bih, a1, 0, 0, 0);
        const int d = (j - k) * 16 + k;
        // result row 8 (e/4) + 4 h + e%4 of the 32: Re Y_q for rows < 16, Im Y_q at row + 16 (e + 8)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int q = (e < 4 ? 0 : 8) + 4 * h + (e & 3);
            dst[pidx(d + q * Ns, E)] = C2<float>{a0[e] + a1[e], a0[e + 8] + a1[e + 8]};
        }
    }
}

// The passes over M more points from buffer P: radix 16 on MFMA while 16
// divides M, then one VALU pass of the remaining radix (2, 4 or 8).
template <int T, int NT, int E, int M, int SIGN, int P>
struct MfmaChain {
    static constexpr bool MF = M >= 16 && M % 16 == 0;
    static constexpr int RAD = MF ? 16 : M;
    using Next = MfmaChain<T, NT, E, M / RAD, SIGN, 1 - P>;
    static constexpr int out = Next::out;
    // twp: the lane-order twiddle tables of the MFMA passes after the first
    // ([pass][group][r], T entries each); tw: the plain table for a VALU pass
    static __device__ __forceinline__ void run(C2<float>* b0, C2<float>* b1, const C2<float>* tw,
                                               const C2<float>* twp, const v8bf (&fa)[2][2], int Ns)
    {
        if constexpr (MF) {
            mfma_pass16<T, E, SIGN>(P ? b1 : b0, P ? b0 : b1, Ns > 1 ? twp : nullptr, Ns, fa);
            __syncthreads();
            Next::run(b0, b1, tw, Ns > 1 ? twp + T : twp, fa, Ns * RAD);
        } else {
            spass<float, T, NT, E, RAD, SIGN>(P ? b1 : b0, P ? b0 : b1, tw, Ns);
            __syncthreads();
            Next::run(b0, b1, tw, twp, fa, Ns * RAD);
        }
    }
};
template <int T, int NT, int E, int SIGN, int P>
struct MfmaChain<T, NT, E, 1, SIGN, P> {
    static constexpr int out = P;
    static __device__ __forceinline__ void run(C2<float>*, C2<float>*, const C2<float>*, const C2<float>*,
                                               const v8bf (&)[2][2], int) {}
};

// the operator fragments of sign SIGN (dft16 holds [sign: -1, +1][k slice][hi, lo][lane][8])
template <int SIGN>
__device__ __forceinline__ void load_dft16(const uint16_t* dft16, v8bf (&fa)[2][2])
{
    const int lane = threadIdx.x & 63;
    const uint16_t* f = dft16 + (SIGN > 0 ? 2 * 2 * 64 * 8 : 0);
#pragma unroll
    for (int sl = 0; sl < 2; ++sl)
#pragma unroll
        for (int hl = 0; hl < 2; ++hl) fa[sl][hl] = *(const v8bf*)(f + ((sl * 2 + hl) * 64 + lane) * 8);
}

// fft_inv with the passes of MfmaChain (FP32 data, T >= 512): strided-layout
// spectrum in, chunk-layout time signal out, unnormalised
template <int T, int NT, int E>
__device__ __forceinline__ void fft_inv_mfma(const C2<float> (&y)[E], C2<float> (&x)[E], C2<float>* b0, C2<float>* b1,
                                             const C2<float>* tw, const uint16_t* dft16, const C2<float>* twp)
{
    const int tid = threadIdx.x;
    v8bf fa[2][2];
    load_dft16<+1>(dft16, fa);
#pragma unroll
    for (int r = 0; r < E; ++r) b0[pidx(tid + r * NT, E)] = y[r];
    __syncthreads();
    using Chain = MfmaChain<T, NT, E, T, +1, 0>;
    Chain::run(b0, b1, tw, twp, fa, 1);
    const C2<float>* src = Chain::out ? b1 : b0;
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = src[pidx(tid * E + e, E)];
    if constexpr (Chain::out == 0) __syncthreads();   // as in fft_fwd_acc
}

// fft_fwd_acc with the passes of MfmaChain (FP32 data, T >= 512): same
// arguments and result layout
template <int MODE, int T, int NT, int E, class F>
__device__ __forceinline__ void fft_fwd_acc_mfma(const C2<float> (&x)[E], const F (&fac)[E], C2<float> (&out)[E],
                                                 C2<float>* b0, C2<float>* b1, const C2<float>* tw, const void* table,
                                                 const uint16_t* dft16, const C2<float>* twp)
{
    const int tid = threadIdx.x;
    v8bf fa[2][2];
    load_dft16<-1>(dft16, fa);
#pragma unroll
    for (int e = 0; e < E; ++e) b0[pidx(tid * E + e, E)] = cscale(x[e], (float)fac[e]);
    __syncthreads();
    using Chain = MfmaChain<T, NT, E, T, -1, 0>;
    Chain::run(b0, b1, tw, twp, fa, 1);
    const C2<float>* src = Chain::out ? b1 : b0;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int w = tid + r * NT;
        const C2<float> y = src[pidx(w, E)];
        C2<float> g;
        if constexpr (MODE == 0) {
            g = y;
        } else if constexpr (MODE == 1) {
            g = cmul(((const C2<float>*)table)[w], y);
        } else {
            const float kk = ((const float*)table)[w];
            g = C2<float>{kk * y.y, -kk * y.x};
        }
        out[r] = cadd(out[r], g);
    }
    // the next transform writes b0 first and passes a barrier before it writes b1
    if constexpr (Chain::out == 0) __syncthreads();
}

// the forward transform of one RHS term: on MFMA for the mixed policy when
// asked (mfma_fwd 1: the Kerr/Raman term; 2: every term), VALU otherwise
template <int MODE, class Pol, int T, int NT, int E, class F>
__device__ __forceinline__ void fft_fwd_term(const C2<typename Pol::R> (&x)[E], const F (&fac)[E],
                                             C2<typename Pol::R> (&out)[E], C2<typename Pol::R>* b0,
                                             C2<typename Pol::R>* b1, const C2<typename Pol::R>* tw, const void* table,
                                             const NLArgs& a)
{
    if constexpr (Pol::mfma_fft && T >= 512) {
        if (a.mfma_fwd >= (MODE == 2 ? 1 : 2)) {
            fft_fwd_acc_mfma<MODE, T, NT, E>(x, fac, out, b0, b1, tw, table, (const uint16_t*)a.dft16,
                                             (const C2<float>*)a.mfma_tw);
            return;
        }
    }
    fft_fwd_acc<MODE, typename Pol::R, T, NT, E>(x, fac, out, b0, b1, tw, table);
}

// ------------------------------------------------------------------ plasma
struct PlasmaScratch {
    // laid out in a separate LDS area: [NT] W, [NT] c, [NT] rho (doubles), scan + reduction slots
    double* W;
    double* c;
    double* rho;
    void* scan;      // (NW+1) affine maps
    double* red;     // 2*NW doubles
};

// rho(t) of one line from per-sample rates Wv (W(I)) and cv (avalanche
// coefficient), chunk layout.  Newton on the chunk boundaries: thread j
// runs its E Heun steps exactly from a guessed start value g_j (the density
// at sample jE-1), carrying the derivative d out/d g_j; the affine block scan
// of these linearised chunk maps gives the next guesses.  Once the guesses
// stop moving (the start each thread used is within the tolerance of the
// one the scan returns) the values of that pass are rho.
template <class S, int E, int NT, class Wt, class O>
__device__ bool plasma_scan2(const Wt (&Wv)[E], const Wt (&cv)[E], O (&rho)[E], const NLArgs& a,
                             const PlasmaScratch& ps, int* flags)
{
    const int tid = threadIdx.x;
    bool any = false;
#pragma unroll
    for (int e = 0; e < E; ++e) any |= Wv[e] > (Wt)0;
    const bool line_any = __syncthreads_or(any);
    if (!line_any && a.rho_init == 0.0) {
#pragma unroll
        for (int e = 0; e < E; ++e) rho[e] = 0;
        return false;
    }
    const S dt = (S)a.dt, N = (S)a.neutral, invN = (S)a.inv_neutral, invtau = (S)a.inv_recomb;
    ps.W[tid] = (double)Wv[E - 1];
    ps.c[tid] = (double)cv[E - 1];
    __syncthreads();
    // rates entering the step into sample jE come from sample jE-1; the step
    // into n = 1 starts from E0 = 0 (ionization.cpp:25)
    const S Wm1 = tid > 0 ? (S)ps.W[tid - 1] : (S)0, cm1 = tid > 0 ? (S)ps.c[tid - 1] : (S)0;
    const S rho0 = (S)a.rho_init;
    AffR<S>* sc = (AffR<S>*)ps.scan;
    const S tol = sizeof(S) == 8 ? (S)(sizeof(O) == 4 ? GS_SCAN_TOL_F32OUT : 4e-15) : (S)2e-7;
    S g = 0;
    bool done = false;
    for (int it = 0; it < 24 && !done; ++it) {
        S x = g, D = 1;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (tid == 0 && e == 0) {
                x = rho0;
                D = 0;
            } else {
                const S Wp = e == 0 ? Wm1 : ((tid == 0 && e == 1) ? (S)0 : (S)Wv[e - 1]);
                const S cp = e == 0 ? cm1 : ((tid == 0 && e == 1) ? (S)0 : (S)cv[e - 1]);
                S d;
                x = heun_step(x, Wp, cp, (S)Wv[e], (S)cv[e], dt, N, invN, invtau, &d);
                D *= d;
            }
            rho[e] = (O)x;
        }
        AffR<S> tot;
        const AffR<S> pre = block_scan_excl<NT>(AffR<S>{D, x - D * g}, sc, &tot);
        const S gn = tid == 0 ? (S)0 : pre.b;
        double dm = fabs((double)(gn - g)), sm = fmax(fabs((double)gn), fabs((double)x));
        g = gn;
        block_max2<NT>(dm, sm, ps.red);
        done = (dm <= (double)tol * sm) || sm == 0.0;
    }
    if (!done) {
        // chunk-sequential exact pass: thread j starts from thread j-1's end value
        if (tid == 0) *flags |= kFlagScanSlow;
        for (int j = 0; j < NT; ++j) {
            if (tid == j) {
                S x = j > 0 ? (S)ps.rho[j - 1] : (S)0;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if (j == 0 && e == 0) {
                        x = rho0;
                    } else {
                        const S Wp = e == 0 ? Wm1 : ((j == 0 && e == 1) ? (S)0 : (S)Wv[e - 1]);
                        const S cp = e == 0 ? cm1 : ((j == 0 && e == 1) ? (S)0 : (S)cv[e - 1]);
                        S d;
                        x = heun_step(x, Wp, cp, (S)Wv[e], (S)cv[e], dt, N, invN, invtau, &d);
                    }
                }
                ps.rho[j] = (double)x;
            }
            __syncthreads();
        }
        // exact pass from the chunk starts of the sequential sweep
        S x = tid > 0 ? (S)ps.rho[tid - 1] : (S)0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (tid == 0 && e == 0) {
                x = rho0;
            } else {
                const S Wp = e == 0 ? Wm1 : ((tid == 0 && e == 1) ? (S)0 : (S)Wv[e - 1]);
                const S cp = e == 0 ? cm1 : ((tid == 0 && e == 1) ? (S)0 : (S)cv[e - 1]);
                S d;
                x = heun_step(x, Wp, cp, (S)Wv[e], (S)cv[e], dt, N, invN, invtau, &d);
            }
            rho[e] = (O)x;
        }
    }
    return true;
}

// ------------------------------------------------------------------ NL RHS
template <class Pol, int T, int E>
struct LineSmem {
    using R = typename Pol::R;
    using SH = Shape<Pol, T, E>;
    C2<R> b0[SH::PADN];
    C2<R> b1[SH::PADN];
    double pW[SH::NT], pc[SH::NT], prho[SH::NT];
    double red[2 * SH::NW + 2];
    AffC<double> scan[SH::NW + 1];     // large enough for every affine map type
};

template <class Pol, int T, int E>
__device__ __forceinline__ PlasmaScratch scratch_of(LineSmem<Pol, T, E>& sm)
{
    return PlasmaScratch{sm.pW, sm.pc, sm.prho, (void*)sm.scan, sm.red};
}

// calculate_NLresponse of one line: x spectral (strided layout) -> out spectral (strided).
// out = FFT(loss) + PLASMA_FUNC FFT(rho E) - i KERR FFT(((1-fR) I + R*I) E)
// (propagator.cu:629-783), accumulated in that order.  A term whose time signal
// is zero on the whole line (no sample above the loss threshold, rho = 0 when
// nothing ionizes and there are no ambient carriers) has a zero spectrum and
// its FFT is skipped.
template <class Pol, int T, int E>
__device__ __forceinline__ void nl_rhs(const C2<typename Pol::R> (&x)[E], C2<typename Pol::R> (&out)[E], LineSmem<Pol, T, E>& sm,
                       const NLArgs& a, int* flags)
{
    using R = typename Pol::R;
    using S = typename Pol::S;
    using W = typename Pol::W;
    constexpr int NT = Shape<Pol, T, E>::NT;
    const int tid = threadIdx.x, n0 = tid * E;
    const C2<R>* tw = (const C2<R>*)a.tw;

    // E(t) = IFFT(A)/N_T, chunk layout.  Mixed: on the matrix cores, kept when the
    // whole line stays below the ionization threshold I = 1e-3 (with a 1e-4
    // margin, far above the transform's error), else repeated in FP32
    C2<R> Ev[E];
    bool have_e = false;
    if constexpr (Pol::mfma_fft && T >= 512) {
        if (a.mfma_inv && a.rho_init == 0.0) {
            fft_inv_mfma<T, NT, E>(x, Ev, sm.b0, sm.b1, tw, (const uint16_t*)a.dft16, (const C2<float>*)a.mfma_tw);
            const R s2 = (R)(a.inv_nt * a.inv_nt);
            bool hot = false;
#pragma unroll
            for (int e = 0; e < E; ++e) hot |= (Ev[e].x * Ev[e].x + Ev[e].y * Ev[e].y) * s2 >= (R)0.9999e-3;
            have_e = !__syncthreads_or(hot);
        }
    }
    if (!have_e) fft_inv<R, T, NT, E>(x, Ev, sm.b0, sm.b1, tw);
    R lossf[E];
    R rho[E];
    bool anyloss = false, has_rho = false, ionizing;
    {
        W Ws[E];
        bool anyw = false;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            Ev[e] = cscale(Ev[e], (R)a.inv_nt);
            const W I = (W)Ev[e].x * (W)Ev[e].x + (W)Ev[e].y * (W)Ev[e].y;
            Ws[e] = photoionization<W>(I, a.ionrate, flags);
            anyw |= Ws[e] > (W)0;
        }
        // avalanche and loss need W > 0 somewhere (or ambient carriers)
        ionizing = __syncthreads_or(anyw) || a.rho_init != 0.0;
        if (ionizing) {
            W cs[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const W I = (W)Ev[e].x * (W)Ev[e].x + (W)Ev[e].y * (W)Ev[e].y;
                cs[e] = (W)a.sigma / ((W)a.U + (W)a.pond * I) * I;
                const int n = n0 + e;
                // loss[n] = -W (U + p I)/I E for 1 <= n <= T-2 above the ln-threshold (quirk, ionization.cpp:43)
                lossf[e] = (n >= 1 && n <= T - 2 && I > (W)29.933606208922594)
                               ? (R)(-(W)0.5 * Ws[e] * ((W)a.U + (W)a.pond * I) / I) : (R)0;
                anyloss |= lossf[e] != (R)0;
            }
            has_rho = plasma_scan2<S, E, NT>(Ws, cs, rho, a, scratch_of(sm), flags);
        }
    }

#pragma unroll
    for (int r = 0; r < E; ++r) out[r] = C2<R>{0, 0};
    if (ionizing && __syncthreads_or(anyloss)) fft_fwd_term<0, Pol, T, NT, E>(Ev, lossf, out, sm.b0, sm.b1, tw, nullptr, a);
    if (has_rho) fft_fwd_term<1, Pol, T, NT, E>(Ev, rho, out, sm.b0, sm.b1, tw, a.pf, a);
    if (!a.use_kerr) return;

    // Raman: s_n = z s_{n-1} + I_n; y = (1-fR) I + fR M Im(s (1 - z^T) + z^{n+1} s_{T-1})
    R yv[E];
    if (a.use_raman) {
        const C2<R>* zp = (const C2<R>*)a.zpow;
        const C2<R> z = zp[1];
        C2<R> loc = {0, 0};
#pragma unroll
        for (int e = 0; e < E; ++e) {
            loc = cmul(z, loc);
            loc.x += Ev[e].x * Ev[e].x + Ev[e].y * Ev[e].y;
        }
        const C2<R> zE = zp[E];
        AffC<R> tot;
        const AffC<R> pre = block_scan_excl<NT>(AffC<R>{zE.x, zE.y, loc.x, loc.y}, (AffC<R>*)sm.scan, &tot);
        const C2<R> slast = {tot.br, tot.bi};
        const C2<R> omz = {(R)a.onemzT_re, (R)a.onemzT_im};
        C2<R> sv = {pre.br, pre.bi};     // s at sample n0 - 1
        C2<R> zq = zp[n0 + 1];           // z^{n+1} at n = n0
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const R I = Ev[e].x * Ev[e].x + Ev[e].y * Ev[e].y;
            sv = cmul(z, sv);
            sv.x += I;
            const R qy = sv.x * omz.y + sv.y * omz.x + zq.x * slast.y + zq.y * slast.x;
            yv[e] = (R)a.one_m_fr * I + (R)a.frM * qy;
            zq = cmul(zq, z);
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) yv[e] = Ev[e].x * Ev[e].x + Ev[e].y * Ev[e].y;
    }
    fft_fwd_term<2, Pol, T, NT, E>(Ev, yv, out, sm.b0, sm.b1, tw, a.kerr, a);
}

}  // namespace gs
