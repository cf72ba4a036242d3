// gs_device.hip -- gfx950 kernels of the gsolver_solid_qdht z-step.
//
//  * k_nl_rk4        one workgroup per radial line: the whole classical RK4 of
//                    the nonlinear RHS (propagator.cu:376-425) with the line in
//                    LDS; per stage an inverse FFT, W(I) table lookups, the
//                    plasma recurrence (ionization.cpp:9-95) as Newton
//                    iterations over a block-wide affine scan, the Raman
//                    convolution (propagator.cu:696-703) as a complex affine
//                    scan in closed form, and three forward FFTs.
//  * k_qdht_*        the QDHT (fhatha.cpp:218-230) as a real GEMM on MFMA:
//                    Y[R][2T] = A[R][R] X[R][2T], A = diag(m1) C diag(1/m1);
//                    FP32 (v_mfma_f32_32x32x2_f32), split-BF16 3-pass
//                    (v_mfma_f32_32x32x16_bf16, hi*hi + hi*lo + lo*hi, FP32
//                    accumulate) and FP64 (v_mfma_f64_16x16x4_f64).  Complex
//                    columns are split into re/im planes in LDS so one lane
//                    owns both parts of an output, which lets the linear step
//                    (propagator.cu:570-625) run as the GEMM epilogue.
//  * helpers         batched line FFTs, dump-side line diagnostics, the
//                    stand-alone linear step, reductions, conversions.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gs_device.h"

namespace gs {

// ------------------------------------------------------------------ types
template <class R>
struct alignas(2 * sizeof(R)) C2 {
    R x, y;
};

template <class R> __device__ __forceinline__ C2<R> cadd(C2<R> a, C2<R> b) { return {a.x + b.x, a.y + b.y}; }
template <class R> __device__ __forceinline__ C2<R> csub(C2<R> a, C2<R> b) { return {a.x - b.x, a.y - b.y}; }
template <class R> __device__ __forceinline__ C2<R> cmul(C2<R> a, C2<R> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <class R> __device__ __forceinline__ C2<R> cscale(C2<R> a, R s) { return {a.x * s, a.y * s}; }
// multiply by sign*i
template <int SIGN, class R> __device__ __forceinline__ C2<R> cmuli(C2<R> a) { return SIGN > 0 ? C2<R>{-a.y, a.x} : C2<R>{a.y, -a.x}; }

struct PolF64 { using R = double; using S = double; using W = double; using P = double; };
struct PolF32 { using R = float;  using S = float;  using W = float;  using P = float;  };
struct PolMix { using R = float;  using S = double; using W = double; using P = double; };

template <int T> struct LineShape {
    static constexpr int E = T >= 512 ? 8 : 4;          // elements per thread
    static constexpr int NT = T / E;                     // threads per line
    static constexpr int NW = NT / 64;                   // waves per line
    static_assert(NT >= 64 && NT <= 1024, "unsupported N_T");
};

#define HIPCHK(x)                                  \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return (int)e_;      \
    } while (0)

// ------------------------------------------------------------------ block FFT
// Stockham autosort, radix 8 passes then one radix 4 or 2 pass, in LDS.
// X[q] = sum_r x[r] e^{SIGN 2 pi i r q / n}.  tw[k] = e^{-2 pi i k / T}.
template <int RADIX, int SIGN, class R>
__device__ __forceinline__ void dft_small(C2<R>* v)
{
    if constexpr (RADIX == 2) {
        C2<R> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (RADIX == 4) {
        C2<R> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        C2<R> s13 = cadd(v[1], v[3]), d13 = cmuli<SIGN>(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[1] = cadd(d02, d13);
        v[2] = csub(s02, s13);
        v[3] = csub(d02, d13);
    } else {
        const R h = (R)0.70710678118654752440084436210485;
        C2<R> a[4], b[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            a[r] = cadd(v[r], v[r + 4]);
            b[r] = csub(v[r], v[r + 4]);
        }
        // b[r] *= w8^r, w8 = e^{SIGN i pi/4}
        b[1] = C2<R>{h * (b[1].x - (R)SIGN * b[1].y), h * (b[1].y + (R)SIGN * b[1].x)};
        b[2] = cmuli<SIGN>(b[2]);
        b[3] = C2<R>{h * (-b[3].x - (R)SIGN * b[3].y), h * (-b[3].y + (R)SIGN * b[3].x)};
        dft_small<4, SIGN>(a);
        dft_small<4, SIGN>(b);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[2 * q] = a[q];
            v[2 * q + 1] = b[q];
        }
    }
}

template <class R, int T, int NT, int RADIX, int SIGN>
__device__ __forceinline__ void fft_pass(C2<R>* buf, const C2<R>* __restrict__ tw, int Ns)
{
    constexpr int NB = T / RADIX;
    constexpr int PER = (NB + NT - 1) / NT;
    C2<R> v[PER][RADIX];
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = threadIdx.x + b * NT;
        if (NB % NT == 0 || j < NB) {
#pragma unroll
            for (int r = 0; r < RADIX; ++r) v[b][r] = buf[j + r * NB];
            if (Ns > 1) {
                const int k = j & (Ns - 1);
                const int step = k * (T / (Ns * RADIX));
#pragma unroll
                for (int r = 1; r < RADIX; ++r) {
                    C2<R> w = tw[r * step];
                    if (SIGN > 0) w.y = -w.y;
                    v[b][r] = cmul(v[b][r], w);
                }
            }
            dft_small<RADIX, SIGN>(v[b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = threadIdx.x + b * NT;
        if (NB % NT == 0 || j < NB) {
            const int k = j & (Ns - 1);
            const int d = (j - k) * RADIX + k;
#pragma unroll
            for (int r = 0; r < RADIX; ++r) buf[d + r * Ns] = v[b][r];
        }
    }
    __syncthreads();
}

template <class R, int T, int NT, int SIGN>
__device__ void block_fft(C2<R>* buf, const C2<R>* __restrict__ tw)
{
    constexpr int L = __builtin_ctz(T);
    int Ns = 1;
#pragma unroll
    for (int p = 0; p < L / 3; ++p) {
        fft_pass<R, T, NT, 8, SIGN>(buf, tw, Ns);
        Ns *= 8;
    }
    if constexpr (L % 3 == 2) fft_pass<R, T, NT, 4, SIGN>(buf, tw, Ns);
    if constexpr (L % 3 == 1) fft_pass<R, T, NT, 2, SIGN>(buf, tw, Ns);
}

// ------------------------------------------------------------------ block scans
template <class V> struct AffR { V a, b; };               // x -> a x + b
template <class V> struct AffC { V ar, ai, br, bi; };     // complex x -> a x + b

template <class V> __device__ __forceinline__ AffR<V> compose(AffR<V> e, AffR<V> l)
{   // apply e then l
    return {l.a * e.a, l.a * e.b + l.b};
}
template <class V> __device__ __forceinline__ AffC<V> compose(AffC<V> e, AffC<V> l)
{
    return {l.ar * e.ar - l.ai * e.ai, l.ar * e.ai + l.ai * e.ar,
            l.ar * e.br - l.ai * e.bi + l.br, l.ar * e.bi + l.ai * e.br + l.bi};
}
template <class V> __device__ __forceinline__ AffR<V> shfl_up(AffR<V> v, int d) { return {__shfl_up(v.a, d), __shfl_up(v.b, d)}; }
template <class V> __device__ __forceinline__ AffC<V> shfl_up(AffC<V> v, int d)
{
    return {__shfl_up(v.ar, d), __shfl_up(v.ai, d), __shfl_up(v.br, d), __shfl_up(v.bi, d)};
}
template <class V> __device__ __forceinline__ AffR<V> aff_id(AffR<V>*) { return {(V)1, (V)0}; }
template <class V> __device__ __forceinline__ AffC<V> aff_id(AffC<V>*) { return {(V)1, (V)0, (V)0, (V)0}; }

// Exclusive block scan of affine maps (composition, earlier threads first).
// lds must hold NW+1 entries.  Returns the prefix; *total = all threads composed.
template <int NT, class A>
__device__ A block_scan_excl(A v, A* lds, A* total)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const A id = aff_id((A*)nullptr);
    A inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        A o = shfl_up(inc, d);
        if (lane >= d) inc = compose(o, inc);
    }
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        A run = id;
        for (int w = 0; w < NW; ++w) {
            A t = lds[w];
            lds[w] = run;
            run = compose(run, t);
        }
        lds[NW] = run;
    }
    __syncthreads();
    A exc = shfl_up(inc, 1);
    if (lane == 0) exc = id;
    A res = compose(lds[wid], exc);
    *total = lds[NW];
    __syncthreads();
    return res;
}

template <int NT>
__device__ double block_max(double v, double* lds)
{
    constexpr int NW = NT / 64;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = lds[0];
    for (int w = 1; w < NW; ++w) r = fmax(r, lds[w]);
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------ plasma
// W(I), ionization.cpp:98-147 (solid, table branch).  Sets status bits on
// the reference's throw conditions and returns 0 there.
template <class W>
__device__ __forceinline__ W photoionization(W I, const double* __restrict__ tab, int* flags)
{
    if (I < (W)1e-3) return 0;   // IONIZATION_MIN_I / INTENSITY_DENOM
    const W lnI = log(I) + (W)36.841361487904734;
    if (lnI != lnI) { *flags |= kFlagNaN; return 0; }
    const W u = (lnI - (W)29.933606208922594) / (W)0.003304003148325;
    const int n = (int)floor(u);
    if (n > 8190) { *flags |= kFlagBlast; return 0; }
    if (n < 0) return 0;
    const W lnIn = (W)29.933606208922594 + (W)n * (W)0.003304003148325;
    const W a = (W)tab[n], b = (W)tab[n + 1];
    return exp(a + (lnI - lnIn) * (b - a) / (W)0.003304003148325);
}

// One Heun step of ionization.cpp:36-41 from rho = x with the rates of the
// previous sample (Wp, cp) and of this sample (Wn, cn); returns the new rho
// and the derivative d(new)/dx (0 where the clamp at N is active).
template <class S>
__device__ __forceinline__ S heun_step(S x, S Wp, S cp, S Wn, S cn, S dt, S N, S invN, S invtau, S* deriv)
{
    const S W01 = x > N ? (S)0 : Wp;
    const S av0 = (x < 0 || x > N) ? (S)0 : cp * x * ((S)1 - x * invN);
    const S re0 = x > 0 ? x * invtau : (S)0;
    const S k1 = dt * (W01 + av0 - re0);
    const S xm = x + k1;
    const S W11 = xm > N ? (S)0 : Wn;
    const S av1 = (xm < 0 || xm > N) ? (S)0 : cn * xm * ((S)1 - xm * invN);
    const S re1 = xm > 0 ? xm * invtau : (S)0;
    const S k2 = dt * (W11 + av1 - re1);
    S y = x + (S)0.5 * (k1 + k2);
    // derivative of the active branch (x >= 0 side at the kink of rec)
    const S dav0 = (x < 0 || x > N) ? (S)0 : cp * ((S)1 - (S)2 * x * invN);
    const S dre0 = x >= 0 ? invtau : (S)0;
    const S dk1 = dt * (dav0 - dre0);
    const S dav1 = (xm < 0 || xm > N) ? (S)0 : cn * ((S)1 - (S)2 * xm * invN);
    const S dre1 = xm >= 0 ? invtau : (S)0;
    const S dk2 = dt * (dav1 - dre1) * ((S)1 + dk1);
    S d = (S)1 + (S)0.5 * (dk1 + dk2);
    if (y > N) { y = N; d = 0; }
    *deriv = d;
    return y;
}

// Plasma density of one line from the per-sample rates, thread-chunk layout
// (thread j owns samples j*E .. j*E+E-1).  Solves rho_n = f_n(rho_{n-1})
// with Newton iterations, each a block affine scan of the linearised maps;
// falls back to a chunk-sequential pass if Newton has not converged.
template <class S, int T>
__device__ void plasma_scan(const S (&Wv)[LineShape<T>::E], const S (&cv)[LineShape<T>::E], S (&rho)[LineShape<T>::E],
                            const NLArgs& a, void* scratch, int* flags)
{
    constexpr int E = LineShape<T>::E, NT = LineShape<T>::NT;
    const int tid = threadIdx.x;
    S* xW = (S*)scratch;              // [NT] last W of each thread
    S* xc = xW + NT;                  // [NT] last c
    S* xr = xc + NT;                  // [NT] last rho (current iterate)
    AffR<S>* sc = (AffR<S>*)(xr + NT);   // [NW+1]
    double* red = (double*)(sc + LineShape<T>::NW + 1);

    const S dt = (S)a.dt, N = (S)a.neutral, invN = (S)a.inv_neutral, invtau = (S)a.inv_recomb;
    xW[tid] = Wv[E - 1];
    xc[tid] = cv[E - 1];
    __syncthreads();
    // rates entering step n: (W, c) of sample n-1; sample 0 is the E0 = 0 start
    S Wp[E], cp[E];
    {
        const S w0 = tid > 0 ? xW[tid - 1] : (S)0, c0 = tid > 0 ? xc[tid - 1] : (S)0;
        Wp[0] = w0;
        cp[0] = c0;
#pragma unroll
        for (int e = 1; e < E; ++e) { Wp[e] = Wv[e - 1]; cp[e] = cv[e - 1]; }
        if (tid == 0 && E > 1) { Wp[1] = 0; cp[1] = 0; }   // n = 1: E0 = 0 (ionization.cpp:25)
    }
    const S tol = sizeof(S) == 8 ? (S)4e-15 : (S)2e-7;
#pragma unroll
    for (int e = 0; e < E; ++e) rho[e] = 0;
    bool done = false;
    for (int it = 0; it < 24 && !done; ++it) {
        xr[tid] = rho[E - 1];
        __syncthreads();
        S prev = tid > 0 ? xr[tid - 1] : (S)0;
        AffR<S> m[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const S x = e == 0 ? prev : rho[e - 1];
            if (tid == 0 && e == 0) {
                m[e] = {(S)0, (S)a.rho_init};
            } else {
                S d;
                const S f = heun_step(x, Wp[e], cp[e], Wv[e], cv[e], dt, N, invN, invtau, &d);
                m[e] = {d, f - d * x};
            }
        }
        AffR<S> agg = m[0];
#pragma unroll
        for (int e = 1; e < E; ++e) agg = compose(agg, m[e]);
        AffR<S> tot;
        AffR<S> pre = block_scan_excl<NT>(agg, sc, &tot);
        S x = pre.b;   // applied to rho_{-1} = 0
        double dmax = 0, smax = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            x = m[e].a * x + m[e].b;
            dmax = fmax(dmax, (double)fabs(x - rho[e]));
            smax = fmax(smax, (double)fabs(x));
            rho[e] = x;
        }
        const double D = block_max<NT>(dmax, red);
        const double M = block_max<NT>(smax, red + 32);
        done = (D <= (double)tol * M) || M == 0.0;
    }
    if (!done) {
        // chunk-sequential exact pass: thread j starts from thread j-1's end value
        if (tid == 0) *flags |= kFlagScanSlow;
        for (int j = 0; j < NT; ++j) {
            if (tid == j) {
                S x = j > 0 ? xr[j - 1] : (S)0;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    S d;
                    x = (j == 0 && e == 0) ? (S)a.rho_init : heun_step(x, Wp[e], cp[e], Wv[e], cv[e], dt, N, invN, invtau, &d);
                    rho[e] = x;
                }
                xr[j] = x;
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------ NL RHS
template <class Pol, int T>
struct NLSmem {
    using R = typename Pol::R;
    C2<R> A[T];      // the line input (spectral)
    C2<R> buf[T];    // FFT work area; doubles as scan scratch
};

// calculate_NLresponse (propagator.cu:629-783) of one line; x holds the
// spectral input in thread-chunk layout, out receives the spectral RHS.
template <class Pol, int T>
__device__ void nl_rhs(const C2<typename Pol::R> (&x)[LineShape<T>::E], C2<typename Pol::R> (&out)[LineShape<T>::E],
                       C2<typename Pol::R>* buf, const NLArgs& a, int* flags)
{
    using R = typename Pol::R;
    using S = typename Pol::S;
    using W = typename Pol::W;
    constexpr int E = LineShape<T>::E, NT = LineShape<T>::NT;
    const int tid = threadIdx.x, n0 = tid * E;
    const C2<R>* tw = (const C2<R>*)a.tw;

    // E(t) = IFFT(A) / N_T
#pragma unroll
    for (int e = 0; e < E; ++e) buf[n0 + e] = x[e];
    __syncthreads();
    block_fft<R, T, NT, +1>(buf, tw);
    C2<R> Ev[E];
    R Iv[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        Ev[e] = cscale(buf[n0 + e], (R)a.inv_nt);
        Iv[e] = Ev[e].x * Ev[e].x + Ev[e].y * Ev[e].y;
    }
    __syncthreads();   // buf becomes scan scratch

    // ionization rates and avalanche coefficients per sample
    S Ws[E], cs[E], rho[E];
    R lossf[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const W I = (W)Iv[e];
        const W Wr = photoionization<W>(I, a.ionrate, flags);
        const W c = (W)a.sigma / ((W)a.U + (W)a.pond * I) * I;
        Ws[e] = (S)Wr;
        cs[e] = (S)c;
        const int n = n0 + e;
        // loss[n] = -W (U + p I)/I E for 1 <= n <= T-2 above the ln-threshold (quirk)
        lossf[e] = (n >= 1 && n <= T - 2 && I > (W)29.933606208922594)
                       ? (R)(-(W)0.5 * Wr * ((W)a.U + (W)a.pond * I) / I) : (R)0;
    }
    plasma_scan<S, T>(Ws, cs, rho, a, (void*)buf, flags);

    // Raman: y = (1-fR) I + fR M Im(s (1 - z^T) + z^{n+1} s_{T-1}), s_n = z s_{n-1} + I_n
    R yv[E];
    if (a.use_kerr) {
        if (a.use_raman) {
            const C2<R>* zp = (const C2<R>*)a.zpow;
            const C2<R> z = zp[1];
            C2<R> sl[E];
            C2<R> acc = {0, 0};
#pragma unroll
            for (int e = 0; e < E; ++e) {
                acc = cmul(z, acc);
                acc.x += Iv[e];
                sl[e] = acc;
            }
            const C2<R> zE = zp[E];
            AffC<R> agg = {zE.x, zE.y, acc.x, acc.y};
            AffC<R> tot;
            AffC<R> pre = block_scan_excl<NT>(agg, (AffC<R>*)buf, &tot);
            const C2<R> sstart = {pre.br, pre.bi};
            const C2<R> slast = {tot.br, tot.bi};
            const C2<R> omz = {(R)a.onemzT_re, (R)a.onemzT_im};
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const C2<R> s = cadd(sl[e], cmul(zp[e + 1], sstart));
                const C2<R> q = cadd(cmul(s, omz), cmul(zp[n0 + e + 1], slast));
                yv[e] = (R)a.one_m_fr * Iv[e] + (R)a.frM * q.y;
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) yv[e] = Iv[e];
        }
    }

    // three forward FFTs: loss, rho E, y E
    const C2<R>* pf = (const C2<R>*)a.pf;
    const R* kp = (const R*)a.kerr;
#pragma unroll
    for (int e = 0; e < E; ++e) buf[n0 + e] = cscale(Ev[e], lossf[e]);
    __syncthreads();
    block_fft<R, T, NT, -1>(buf, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) out[e] = buf[n0 + e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) buf[n0 + e] = cscale(Ev[e], (R)rho[e]);
    __syncthreads();
    block_fft<R, T, NT, -1>(buf, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) out[e] = cadd(out[e], cmul(pf[n0 + e], buf[n0 + e]));
    __syncthreads();
    if (a.use_kerr) {
#pragma unroll
        for (int e = 0; e < E; ++e) buf[n0 + e] = cscale(Ev[e], yv[e]);
        __syncthreads();
        block_fft<R, T, NT, -1>(buf, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const C2<R> f = buf[n0 + e];
            const R k = kp[n0 + e];
            out[e] = cadd(out[e], C2<R>{k * f.y, -k * f.x});   // += -i K F
        }
        __syncthreads();
    }
}

template <class Pol, int T>
__global__ __launch_bounds__(LineShape<T>::NT) void k_nl_rk4(const C2<typename Pol::R>* __restrict__ field,
                                                              C2<typename Pol::R>* __restrict__ b1, double dz, NLArgs a,
                                                              double* __restrict__ linemax)
{
    using R = typename Pol::R;
    constexpr int E = LineShape<T>::E, NT = LineShape<T>::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    NLSmem<Pol, T>& sm = *reinterpret_cast<NLSmem<Pol, T>*>(smem_raw);
    const int line = blockIdx.x, tid = threadIdx.x, n0 = tid * E;
    int flags = 0;
    const C2<R>* src = field + (size_t)line * T;
    double mA = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const C2<R> v = src[tid + e * NT];
        sm.A[tid + e * NT] = v;
        mA = fmax(mA, (double)v.x * v.x + (double)v.y * v.y);
    }
    __syncthreads();
    C2<R> k[E], acc[E], xin[E];
    const R rdz = (R)dz;
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
        const R c = s == 0 ? (R)0 : (s == 3 ? rdz : rdz * (R)0.5);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const C2<R> av = sm.A[n0 + e];
            xin[e] = s == 0 ? av : cadd(av, cscale(k[e], c));
        }
        nl_rhs<Pol, T>(xin, k, sm.buf, a, &flags);
        const R wgt = (s == 1 || s == 2) ? (R)2 : (R)1;
#pragma unroll
        for (int e = 0; e < E; ++e) acc[e] = s == 0 ? k[e] : cadd(acc[e], cscale(k[e], wgt));
    }
    double mN = 0;
    C2<R>* dst = b1 + (size_t)line * T;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const C2<R> v = {acc[e].x / (R)6, acc[e].y / (R)6};
        sm.buf[n0 + e] = v;
        mN = fmax(mN, (double)v.x * v.x + (double)v.y * v.y);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) dst[tid + e * NT] = sm.buf[tid + e * NT];
    double* red = (double*)sm.A;
    const double MA = block_max<NT>(mA, red);
    const double MN = block_max<NT>(mN, red + 32);
    if (tid == 0) {
        linemax[2 * line] = MA;
        linemax[2 * line + 1] = MN;
        if (flags) atomicOr(&a.status->flags, flags);
    }
    if (tid != 0 && flags) atomicOr(&a.status->flags, flags);
}

template <class Pol, int T>
__global__ __launch_bounds__(LineShape<T>::NT) void k_nl_response(const C2<typename Pol::R>* __restrict__ in,
                                                                   C2<typename Pol::R>* __restrict__ out, NLArgs a)
{
    using R = typename Pol::R;
    constexpr int E = LineShape<T>::E, NT = LineShape<T>::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    NLSmem<Pol, T>& sm = *reinterpret_cast<NLSmem<Pol, T>*>(smem_raw);
    const int line = blockIdx.x, tid = threadIdx.x, n0 = tid * E;
    int flags = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) sm.A[tid + e * NT] = in[(size_t)line * T + tid + e * NT];
    __syncthreads();
    C2<R> x[E], y[E];
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = sm.A[n0 + e];
    nl_rhs<Pol, T>(x, y, sm.buf, a, &flags);
#pragma unroll
    for (int e = 0; e < E; ++e) sm.buf[n0 + e] = y[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) out[(size_t)line * T + tid + e * NT] = sm.buf[tid + e * NT];
    if (flags) atomicOr(&a.status->flags, flags);
}

// ------------------------------------------------------------------ line FFT / diagnostics
template <class R, int T>
__global__ __launch_bounds__(LineShape<T>::NT) void k_fft_lines(const C2<R>* __restrict__ in, C2<R>* __restrict__ out,
                                                                int sign, double scale, const double* __restrict__ filt,
                                                                const C2<R>* __restrict__ tw)
{
    constexpr int E = LineShape<T>::E, NT = LineShape<T>::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<R>* buf = reinterpret_cast<C2<R>*>(smem_raw);
    const size_t base = (size_t)blockIdx.x * T;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = threadIdx.x + e * NT;
        C2<R> v = in[base + n];
        const R f = (R)(scale * (filt ? filt[n] : 1.0));
        buf[n] = cscale(v, f);
    }
    __syncthreads();
    if (sign > 0) block_fft<R, T, NT, +1>(buf, tw);
    else block_fft<R, T, NT, -1>(buf, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = threadIdx.x + e * NT;
        out[base + n] = buf[n];
    }
}

template <class Pol, int T>
__global__ __launch_bounds__(LineShape<T>::NT) void k_line_diag(const C2<typename Pol::R>* __restrict__ field,
                                                                const double* __restrict__ filt, double tstep, NLArgs a,
                                                                double* flux, double* maxI, double* rhomax,
                                                                C2<typename Pol::R>* et, double* rho_out)
{
    using R = typename Pol::R;
    using S = typename Pol::S;
    using W = typename Pol::W;
    constexpr int E = LineShape<T>::E, NT = LineShape<T>::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    NLSmem<Pol, T>& sm = *reinterpret_cast<NLSmem<Pol, T>*>(smem_raw);
    const C2<R>* tw = (const C2<R>*)a.tw;
    const int line = blockIdx.x, tid = threadIdx.x, n0 = tid * E;
    const size_t base = (size_t)line * T;
    int flags = 0;
    // unfiltered E(t) for the plasma density
#pragma unroll
    for (int e = 0; e < E; ++e) sm.buf[tid + e * NT] = field[base + tid + e * NT];
    __syncthreads();
    block_fft<R, T, NT, +1>(sm.buf, tw);
    C2<R> Ev[E];
    S Ws[E], cs[E], rho[E];
    double fl = 0, mi = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        Ev[e] = cscale(sm.buf[n0 + e], (R)a.inv_nt);
        const W I = (W)Ev[e].x * (W)Ev[e].x + (W)Ev[e].y * (W)Ev[e].y;
        Ws[e] = (S)photoionization<W>(I, a.ionrate, &flags);
        cs[e] = (S)((W)a.sigma / ((W)a.U + (W)a.pond * I) * I);
        fl += (double)I;
        mi = fmax(mi, (double)I);
    }
    if (et) {
#pragma unroll
        for (int e = 0; e < E; ++e) sm.A[n0 + e] = Ev[e];
    }
    __syncthreads();
    plasma_scan<S, T>(Ws, cs, rho, a, (void*)sm.buf, &flags);
    double rm = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        rm = fmax(rm, (double)rho[e]);
        if (rho_out) rho_out[base + n0 + e] = (double)rho[e];
    }
    if (et) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) et[base + tid + e * NT] = sm.A[tid + e * NT];
    }
    if (filt) {
        // flux / maxI of the filtered field (dump_types.cpp:123-165)
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int n = tid + e * NT;
            sm.buf[n] = cscale(field[base + n], (R)(filt[n] * a.inv_nt));
        }
        __syncthreads();
        block_fft<R, T, NT, +1>(sm.buf, tw);
        fl = 0;
        mi = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const C2<R> v = sm.buf[n0 + e];
            const double I = (double)v.x * v.x + (double)v.y * v.y;
            fl += I;
            mi = fmax(mi, I);
        }
    }
    __syncthreads();
    double* red = (double*)sm.A;
    // sum via max-style reduction tree on a separate slot
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) fl += __shfl_xor(fl, d);
    if ((tid & 63) == 0) red[64 + (tid >> 6)] = fl;
    const double MI = block_max<NT>(mi, red);
    const double RM = block_max<NT>(rm, red + 32);
    if (tid == 0) {
        double s = 0;
        for (int w = 0; w < LineShape<T>::NW; ++w) s += red[64 + w];
        if (flux) flux[line] = s * tstep;
        if (maxI) maxI[line] = MI;
        if (rhomax) rhomax[line] = RM;
    }
    if (flags) atomicOr(&a.status->flags, flags);
}

// ------------------------------------------------------------------ linear step
__device__ __forceinline__ double sqrtho_d(double x)
{
    // sqrt(1+x)-1 as the 20-term Taylor sum of extmath.h:245-252, Horner form
    const double c[20] = {0.5, -0.125, 0.0625, -0.0390625, 0.02734375, -0.0205078125, 0.01611328125,
                          -0.013092041015625, 0.010910034179688, -0.009273529052734, 0.008008956909180,
                          -0.007007837295532, 0.006199240684509, -0.005535036325455, 0.004981532692909,
                          -0.004514514002949, 0.004116174532101, -0.003773159987759, 0.003475278936094,
                          -0.003214633015887};
    double r = c[19];
#pragma unroll
    for (int j = 18; j >= 0; --j) r = r * x + c[j];
    return r * x;
}
__device__ __forceinline__ float sqrtho_f(float x)
{
    const float c[20] = {0.5f, -0.125f, 0.0625f, -0.0390625f, 0.02734375f, -0.0205078125f, 0.01611328125f,
                         -0.013092041015625f, 0.010910034179688f, -0.009273529052734f, 0.008008956909180f,
                         -0.007007837295532f, 0.006199240684509f, -0.005535036325455f, 0.004981532692909f,
                         -0.004514514002949f, 0.004116174532101f, -0.003773159987759f, 0.003475278936094f,
                         -0.003214633015887f};
    float r = c[19];
#pragma unroll
    for (int j = 18; j >= 0; --j) r = r * x + c[j];
    return r * x;
}

template <class Pol>
__device__ __forceinline__ void lstep_point(int i, int w, size_t o, typename Pol::R yr, typename Pol::R yi, const LstepArgs& la)
{
    using R = typename Pol::R;
    using P = typename Pol::P;
    C2<R>* fk = (C2<R>*)la.fk;
    C2<R>* b2 = (C2<R>*)la.b2;
    if (i >= la.nr || !la.wok[w]) { fk[o] = C2<R>{0, 0}; return; }
    const P kt2 = (P)la.kt2[i];
    if (kt2 > (P)la.lim[w]) { fk[o] = C2<R>{0, 0}; return; }
    const P k1 = (P)la.k1[w];
    const P s = sizeof(P) == 8 ? (P)sqrtho_d((double)(-kt2 / k1 / k1)) : (P)sqrtho_f((float)(-kt2 / k1 / k1));
    const P dz = (P)la.dz;
    const C2<R> b = b2[o];
    const P br = (P)b.x + dz * (P)yr / ((P)1 + s);
    const P bi = (P)b.y + dz * (P)yi / ((P)1 + s);
    const P ph = dz * (-(P)la.ho[w] - s * k1);
    const P amp = exp(dz * (P)la.k2[w]);
    P sn, cs;
    if constexpr (sizeof(P) == 8) sincos(ph, &sn, &cs);
    else sincosf(ph, &sn, &cs);
    const C2<R> f = {(R)(amp * (br * cs - bi * sn)), (R)(amp * (br * sn + bi * cs))};
    fk[o] = f;
    b2[o] = f;
}

template <class Pol>
__global__ void k_lstep(const C2<typename Pol::R>* __restrict__ b1k, int nt, size_t n, LstepArgs la)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const int i = (int)(o / nt), w = (int)(o % nt);
    const C2<typename Pol::R> y = b1k[o];
    lstep_point<Pol>(i, w, o, y.x, y.y, la);
}

// ------------------------------------------------------------------ QDHT GEMMs
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// blockIdx -> tile, grouping all row tiles of a column tile on one XCD so the
// B columns are reused from that XCD's L2 (dispatch is round-robin over 8 XCDs)
__device__ __forceinline__ void tile_of(int mt, int ntl, int* mi, int* ni)
{
    const int tiles = mt * ntl, b = blockIdx.x;
    int t = b;
    if ((tiles & 7) == 0) t = (b & 7) * (tiles >> 3) + (b >> 3);
    *mi = t % mt;
    *ni = t / mt;
}

// FP32: tile 128 rows x 64 complex columns, K step 16, 4 waves (2x2)
template <bool LSTEP>
__global__ __launch_bounds__(256) void k_qdht_f32(const float* __restrict__ At, const C2<float>* __restrict__ X,
                                                  C2<float>* __restrict__ Y, int K, int NC, LstepArgs la)
{
    constexpr int BM = 128, BNC = 64, BK = 16, PA = 32, PB = 32;
    __shared__ float As[BK][BM + PA];
    __shared__ float Br[BK][BNC + PB];
    __shared__ float Bi[BK][BNC + PB];
    const int M = K;
    int mi, ni;
    tile_of(M / BM, NC / BNC, &mi, &ni);
    const int m0 = mi * BM, n0 = ni * BNC;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 64, wn = (w & 1) * 32;
    const int lr = tid >> 4, lc = (tid & 15);

    v16f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][c][r] = 0.f;

    float4 ra0, ra1, rb0, rb1;
    auto gload = [&](int k0) {
        const float* pa = At + (size_t)(k0 + lr) * M + m0 + lc * 8;
        ra0 = *(const float4*)pa;
        ra1 = *(const float4*)(pa + 4);
        const float* pb = (const float*)(X + (size_t)(k0 + lr) * NC + n0 + lc * 4);
        rb0 = *(const float4*)pb;
        rb1 = *(const float4*)(pb + 4);
    };
    auto sstore = [&]() {
        *(float4*)&As[lr][lc * 8] = ra0;
        *(float4*)&As[lr][lc * 8 + 4] = ra1;
        Br[lr][lc * 4 + 0] = rb0.x; Bi[lr][lc * 4 + 0] = rb0.y;
        Br[lr][lc * 4 + 1] = rb0.z; Bi[lr][lc * 4 + 1] = rb0.w;
        Br[lr][lc * 4 + 2] = rb1.x; Bi[lr][lc * 4 + 2] = rb1.y;
        Br[lr][lc * 4 + 3] = rb1.z; Bi[lr][lc * 4 + 3] = rb1.w;
    };
    gload(0);
    sstore();
    __syncthreads();
    for (int k0 = 0; k0 < K; k0 += BK) {
        const bool more = k0 + BK < K;
        if (more) gload(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 2) {
            const int kr = kk + (lane >> 5), cl = lane & 31;
            const float a0 = As[kr][wm + cl], a1 = As[kr][wm + 32 + cl];
            const float br = Br[kr][wn + cl], bi = Bi[kr][wn + cl];
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, br, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, bi, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, br, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bi, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
        if (more) {
            sstore();
            __syncthreads();
        }
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = m0 + wm + rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int col = n0 + wn + (lane & 31);
            const size_t o = (size_t)row * NC + col;
            if (LSTEP) lstep_point<PolF32>(row, col, o, acc[rb][0][r], acc[rb][1][r], la);
            else Y[o] = C2<float>{acc[rb][0][r], acc[rb][1][r]};
        }
}

// BF16 3-pass: same tile, K step 32, operands k-contiguous in LDS
__device__ __forceinline__ uint32_t bf16_rn(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

template <bool LSTEP>
__global__ __launch_bounds__(256) void k_qdht_bf16x3(const uint16_t* __restrict__ Ah, const uint16_t* __restrict__ Al,
                                                     const C2<float>* __restrict__ X, C2<float>* __restrict__ Y, int K,
                                                     int NC, LstepArgs la)
{
    constexpr int BM = 128, BNC = 64, BK = 32, LD = BK + 8;
    __shared__ __align__(16) uint16_t sAh[BM][LD];
    __shared__ __align__(16) uint16_t sAl[BM][LD];
    __shared__ __align__(16) uint16_t sBh[2][BNC][LD];
    __shared__ __align__(16) uint16_t sBl[2][BNC][LD];
    const int M = K;
    int mi, ni;
    tile_of(M / BM, NC / BNC, &mi, &ni);
    const int m0 = mi * BM, n0 = ni * BNC;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 64, wn = (w & 1) * 32;
    // A loader: row tid/2, k half (tid&1)*16; B loader: k pair tid/16, 4 columns (tid&15)*4
    const int arow = tid >> 1, ak = (tid & 1) * 16;
    const int bkp = tid >> 4, bc = (tid & 15) * 4;

    v16f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][c][r] = 0.f;

    uint4 rah0, rah1, ral0, ral1;
    float4 x0a, x0b, x1a, x1b;
    auto gload = [&](int k0) {
        const size_t ao = (size_t)(m0 + arow) * K + k0 + ak;
        rah0 = *(const uint4*)(Ah + ao);
        rah1 = *(const uint4*)(Ah + ao + 8);
        ral0 = *(const uint4*)(Al + ao);
        ral1 = *(const uint4*)(Al + ao + 8);
        const float* p0 = (const float*)(X + (size_t)(k0 + 2 * bkp) * NC + n0 + bc);
        const float* p1 = (const float*)(X + (size_t)(k0 + 2 * bkp + 1) * NC + n0 + bc);
        x0a = *(const float4*)p0;
        x0b = *(const float4*)(p0 + 4);
        x1a = *(const float4*)p1;
        x1b = *(const float4*)(p1 + 4);
    };
    auto split_store = [&](int c, int part, float v0, float v1) {
        const uint32_t h0 = bf16_rn(v0), h1 = bf16_rn(v1);
        const uint32_t l0 = bf16_rn(v0 - __uint_as_float(h0 << 16)), l1 = bf16_rn(v1 - __uint_as_float(h1 << 16));
        *(uint32_t*)&sBh[part][c][2 * bkp] = h0 | (h1 << 16);
        *(uint32_t*)&sBl[part][c][2 * bkp] = l0 | (l1 << 16);
    };
    auto sstore = [&]() {
        *(uint4*)&sAh[arow][ak] = rah0;
        *(uint4*)&sAh[arow][ak + 8] = rah1;
        *(uint4*)&sAl[arow][ak] = ral0;
        *(uint4*)&sAl[arow][ak + 8] = ral1;
        split_store(bc + 0, 0, x0a.x, x1a.x); split_store(bc + 0, 1, x0a.y, x1a.y);
        split_store(bc + 1, 0, x0a.z, x1a.z); split_store(bc + 1, 1, x0a.w, x1a.w);
        split_store(bc + 2, 0, x0b.x, x1b.x); split_store(bc + 2, 1, x0b.y, x1b.y);
        split_store(bc + 3, 0, x0b.z, x1b.z); split_store(bc + 3, 1, x0b.w, x1b.w);
    };
    gload(0);
    sstore();
    __syncthreads();
    for (int k0 = 0; k0 < K; k0 += BK) {
        const bool more = k0 + BK < K;
        if (more) gload(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const int ko = kk + 8 * (lane >> 5), cl = lane & 31;
            v8bf ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) {
                ah[rb] = *(const v8bf*)&sAh[wm + rb * 32 + cl][ko];
                al[rb] = *(const v8bf*)&sAl[wm + rb * 32 + cl][ko];
            }
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                bh[c] = *(const v8bf*)&sBh[c][wn + cl][ko];
                bl[c] = *(const v8bf*)&sBl[c][wn + cl][ko];
            }
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    acc[rb][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[rb], bh[c], acc[rb][c], 0, 0, 0);
                    acc[rb][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[rb], bl[c], acc[rb][c], 0, 0, 0);
                    acc[rb][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[rb], bh[c], acc[rb][c], 0, 0, 0);
                }
        }
        __syncthreads();
        if (more) {
            sstore();
            __syncthreads();
        }
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = m0 + wm + rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int col = n0 + wn + (lane & 31);
            const size_t o = (size_t)row * NC + col;
            if (LSTEP) lstep_point<PolMix>(row, col, o, acc[rb][0][r], acc[rb][1][r], la);
            else Y[o] = C2<float>{acc[rb][0][r], acc[rb][1][r]};
        }
}

// FP64: tile 64 rows x 32 complex columns, K step 16, 4 waves (2x2)
template <bool LSTEP>
__global__ __launch_bounds__(256) void k_qdht_f64(const double* __restrict__ At, const C2<double>* __restrict__ X,
                                                  C2<double>* __restrict__ Y, int K, int NC, LstepArgs la)
{
    constexpr int BM = 64, BNC = 32, BK = 16, PA = 8, PB = 8;
    __shared__ double As[BK][BM + PA];
    __shared__ double Br[BK][BNC + PB];
    __shared__ double Bi[BK][BNC + PB];
    const int M = K;
    int mi, ni;
    tile_of(M / BM, NC / BNC, &mi, &ni);
    const int m0 = mi * BM, n0 = ni * BNC;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 16;
    const int lr = tid >> 4, lc = tid & 15;
    v4d acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[a][c][r] = 0.0;
    double2 ra0, ra1, rb0, rb1;
    auto gload = [&](int k0) {
        const double* pa = At + (size_t)(k0 + lr) * M + m0 + lc * 4;
        ra0 = *(const double2*)pa;
        ra1 = *(const double2*)(pa + 2);
        const double* pb = (const double*)(X + (size_t)(k0 + lr) * NC + n0 + lc * 2);
        rb0 = *(const double2*)pb;
        rb1 = *(const double2*)(pb + 2);
    };
    auto sstore = [&]() {
        As[lr][lc * 4 + 0] = ra0.x; As[lr][lc * 4 + 1] = ra0.y;
        As[lr][lc * 4 + 2] = ra1.x; As[lr][lc * 4 + 3] = ra1.y;
        Br[lr][lc * 2 + 0] = rb0.x; Bi[lr][lc * 2 + 0] = rb0.y;
        Br[lr][lc * 2 + 1] = rb1.x; Bi[lr][lc * 2 + 1] = rb1.y;
    };
    gload(0);
    sstore();
    __syncthreads();
    for (int k0 = 0; k0 < K; k0 += BK) {
        const bool more = k0 + BK < K;
        if (more) gload(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4), cl = lane & 15;
            const double a0 = As[kr][wm + cl], a1 = As[kr][wm + 16 + cl];
            const double br = Br[kr][wn + cl], bi = Bi[kr][wn + cl];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, br, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bi, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, br, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bi, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
        if (more) {
            sstore();
            __syncthreads();
        }
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // v_mfma_f64_16x16x4_f64 result layout: row = lane/16 + 4 r, column = lane%16
            const int row = m0 + wm + rb * 16 + (lane >> 4) + 4 * r;
            const int col = n0 + wn + (lane & 15);
            const size_t o = (size_t)row * NC + col;
            if (LSTEP) lstep_point<PolF64>(row, col, o, acc[rb][0][r], acc[rb][1][r], la);
            else Y[o] = C2<double>{acc[rb][0][r], acc[rb][1][r]};
        }
}

// ------------------------------------------------------------------ small kernels
__global__ void k_reduce_max(int nr, const double* __restrict__ lm, DevStatus* st)
{
    __shared__ double red[64];
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < nr; i += blockDim.x) {
        a = fmax(a, lm[2 * i]);
        b = fmax(b, lm[2 * i + 1]);
    }
    const double A = block_max<1024>(a, red);
    const double B = block_max<1024>(b, red + 32);
    if (threadIdx.x == 0) {
        st->maxA2 = A;
        st->maxNL2 = B;
    }
}

template <class R>
__global__ void k_convert(const double* __restrict__ src, C2<R>* __restrict__ dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = C2<R>{(R)src[2 * i], (R)src[2 * i + 1]};
}
template <class R>
__global__ void k_convert_back(const C2<R>* __restrict__ src, double* __restrict__ dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        dst[2 * i] = (double)src[i].x;
        dst[2 * i + 1] = (double)src[i].y;
    }
}

template <class R>
__global__ void k_avg_spectrum(int nr, int nt, const C2<R>* __restrict__ f, const double* __restrict__ wts, double* out)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nt) return;
    double acc = 0;
    for (int r = 0; r < nr; ++r) {
        const C2<R> v = f[(size_t)r * nt + w];
        acc += ((double)v.x * v.x + (double)v.y * v.y) * wts[r];
    }
    out[w] = acc;
}

// ------------------------------------------------------------------ launchers
template <class Pol, int T>
static int nl_rk4_T(const void* field, void* b1, double dz, const NLArgs& a, double* linemax, hipStream_t st)
{
    using R = typename Pol::R;
    const size_t sm = sizeof(NLSmem<Pol, T>);
    auto kf = k_nl_rk4<Pol, T>;
    HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(kf, dim3(a.nr), dim3(LineShape<T>::NT), sm, st, (const C2<R>*)field, (C2<R>*)b1, dz, a, linemax);
    HIPCHK(hipGetLastError());
    return 0;
}

template <class Pol, int T>
static int nl_resp_T(const void* in, void* out, int nlines, const NLArgs& a, hipStream_t st)
{
    using R = typename Pol::R;
    const size_t sm = sizeof(NLSmem<Pol, T>);
    auto kf = k_nl_response<Pol, T>;
    HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(kf, dim3(nlines), dim3(LineShape<T>::NT), sm, st, (const C2<R>*)in, (C2<R>*)out, a);
    HIPCHK(hipGetLastError());
    return 0;
}

template <class Pol, int T>
static int diag_T(const void* field, int nlines, const double* filt, double tstep, const NLArgs& a, double* flux,
                  double* maxI, double* rhomax, void* et, double* rho, hipStream_t st)
{
    using R = typename Pol::R;
    const size_t sm = sizeof(NLSmem<Pol, T>);
    auto kf = k_line_diag<Pol, T>;
    HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(kf, dim3(nlines), dim3(LineShape<T>::NT), sm, st, (const C2<R>*)field, filt, tstep, a, flux, maxI,
                       rhomax, (C2<R>*)et, rho);
    HIPCHK(hipGetLastError());
    return 0;
}

template <class R, int T>
static int fft_T(const void* in, void* out, int nlines, int sign, double scale, const double* filt, const void* tw,
                 hipStream_t st)
{
    const size_t sm = sizeof(C2<R>) * T;
    auto kf = k_fft_lines<R, T>;
    HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(kf, dim3(nlines), dim3(LineShape<T>::NT), sm, st, (const C2<R>*)in, (C2<R>*)out, sign, scale,
                       filt, (const C2<R>*)tw);
    HIPCHK(hipGetLastError());
    return 0;
}

#define GS_T_SWITCH(nt, CALL)                                   \
    switch (nt) {                                               \
    case 256: { constexpr int T_ = 256; return CALL; }          \
    case 512: { constexpr int T_ = 512; return CALL; }          \
    case 1024: { constexpr int T_ = 1024; return CALL; }        \
    case 2048: { constexpr int T_ = 2048; return CALL; }        \
    case 4096: { constexpr int T_ = 4096; return CALL; }        \
    default: return -1000;                                      \
    }

int launch_nl_rk4(Prec p, int nt, const void* field, void* b1, double dz, const NLArgs& a, double* linemax, hipStream_t st)
{
    if (p == kFP64) GS_T_SWITCH(nt, (nl_rk4_T<PolF64, T_>(field, b1, dz, a, linemax, st)));
    if (p == kFP32) GS_T_SWITCH(nt, (nl_rk4_T<PolF32, T_>(field, b1, dz, a, linemax, st)));
    GS_T_SWITCH(nt, (nl_rk4_T<PolMix, T_>(field, b1, dz, a, linemax, st)));
}

int launch_nl_response(Prec p, int nt, const void* in, void* out, int nlines, const NLArgs& a, hipStream_t st)
{
    if (p == kFP64) GS_T_SWITCH(nt, (nl_resp_T<PolF64, T_>(in, out, nlines, a, st)));
    if (p == kFP32) GS_T_SWITCH(nt, (nl_resp_T<PolF32, T_>(in, out, nlines, a, st)));
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

int launch_qdht(Prec p, int rp, int nt, const void* amat, const void* amat_lo, const void* in, void* out,
                const LstepArgs* ls, hipStream_t st)
{
    LstepArgs la = ls ? *ls : LstepArgs{};
    if (p == kFP64) {
        if (rp % 64 || nt % 32) return -1001;
        const dim3 grid((rp / 64) * (nt / 32));
        if (ls) hipLaunchKernelGGL(k_qdht_f64<true>, grid, dim3(256), 0, st, (const double*)amat, (const C2<double>*)in, (C2<double>*)out, rp, nt, la);
        else hipLaunchKernelGGL(k_qdht_f64<false>, grid, dim3(256), 0, st, (const double*)amat, (const C2<double>*)in, (C2<double>*)out, rp, nt, la);
    } else if (p == kFP32) {
        if (rp % 128 || nt % 64) return -1001;
        const dim3 grid((rp / 128) * (nt / 64));
        if (ls) hipLaunchKernelGGL(k_qdht_f32<true>, grid, dim3(256), 0, st, (const float*)amat, (const C2<float>*)in, (C2<float>*)out, rp, nt, la);
        else hipLaunchKernelGGL(k_qdht_f32<false>, grid, dim3(256), 0, st, (const float*)amat, (const C2<float>*)in, (C2<float>*)out, rp, nt, la);
    } else {
        if (rp % 128 || nt % 64) return -1001;
        const dim3 grid((rp / 128) * (nt / 64));
        if (ls) hipLaunchKernelGGL(k_qdht_bf16x3<true>, grid, dim3(256), 0, st, (const uint16_t*)amat, (const uint16_t*)amat_lo, (const C2<float>*)in, (C2<float>*)out, rp, nt, la);
        else hipLaunchKernelGGL(k_qdht_bf16x3<false>, grid, dim3(256), 0, st, (const uint16_t*)amat, (const uint16_t*)amat_lo, (const C2<float>*)in, (C2<float>*)out, rp, nt, la);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int launch_lstep(Prec p, int rp, int nt, const void* b1k, const LstepArgs& la, hipStream_t st)
{
    const size_t n = (size_t)rp * nt;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (p == kFP64) hipLaunchKernelGGL(k_lstep<PolF64>, grid, dim3(256), 0, st, (const C2<double>*)b1k, nt, n, la);
    else if (p == kFP32) hipLaunchKernelGGL(k_lstep<PolF32>, grid, dim3(256), 0, st, (const C2<float>*)b1k, nt, n, la);
    else hipLaunchKernelGGL(k_lstep<PolMix>, grid, dim3(256), 0, st, (const C2<float>*)b1k, nt, n, la);
    HIPCHK(hipGetLastError());
    return 0;
}

int launch_reduce_max(int nr, const double* linemax, DevStatus* s, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, st, nr, linemax, s);
    HIPCHK(hipGetLastError());
    return 0;
}

int launch_convert(Prec p, const double* src64, void* dst, size_t n, int to_device, hipStream_t st)
{
    const dim3 grid((unsigned)((n + 255) / 256));
    if (p == kFP64) {
        if (to_device) hipLaunchKernelGGL(k_convert<double>, grid, dim3(256), 0, st, src64, (C2<double>*)dst, n);
        else hipLaunchKernelGGL(k_convert_back<double>, grid, dim3(256), 0, st, (const C2<double>*)dst, (double*)src64, n);
    } else {
        if (to_device) hipLaunchKernelGGL(k_convert<float>, grid, dim3(256), 0, st, src64, (C2<float>*)dst, n);
        else hipLaunchKernelGGL(k_convert_back<float>, grid, dim3(256), 0, st, (const C2<float>*)dst, (double*)src64, n);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int launch_avg_spectrum(Prec p, int nr, int nt, const void* field, const double* wts, double* out, hipStream_t st)
{
    const dim3 grid((nt + 255) / 256);
    if (p == kFP64) hipLaunchKernelGGL(k_avg_spectrum<double>, grid, dim3(256), 0, st, nr, nt, (const C2<double>*)field, wts, out);
    else hipLaunchKernelGGL(k_avg_spectrum<float>, grid, dim3(256), 0, st, nr, nt, (const C2<float>*)field, wts, out);
    HIPCHK(hipGetLastError());
    return 0;
}

}  // namespace gs
