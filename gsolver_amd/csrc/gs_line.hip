// gs_line.hip -- per-radial-line kernels of the z-step: the fused RK4 of the
// nonlinear RHS (propagator.cu:376-425 around calculate_NLresponse,
// propagator.cu:629-783), a single RHS evaluation, batched line FFTs and the
// dump-side line diagnostics.
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
#include "gs_dev.h"

#ifndef GS_NL_E32
#define GS_NL_E32 8    // samples per thread of the FP32/mixed line kernels for N_T >= 1024
#endif
#ifndef GS_NL_WAVES
#define GS_NL_WAVES 1  // minimum waves per SIMD the RK4 kernel is register-allocated for
#endif

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
            if (Ns > 1) {
                C2<R> w = tw[k * (T / (Ns * RADIX))];
                if (SIGN > 0) w.y = -w.y;
                twiddle_pow<RADIX>(v, w);
            }
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
    twiddle_pow<E>(y, tw[tid]);
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
    const S tol = sizeof(S) == 8 ? (S)4e-15 : (S)2e-7;
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

    // E(t) = IFFT(A)/N_T, chunk layout
    C2<R> Ev[E];
    fft_inv<R, T, NT, E>(x, Ev, sm.b0, sm.b1, tw);
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
    if (ionizing && __syncthreads_or(anyloss)) fft_fwd_acc<0, R, T, NT, E>(Ev, lossf, out, sm.b0, sm.b1, tw, nullptr);
    if (has_rho) fft_fwd_acc<1, R, T, NT, E>(Ev, rho, out, sm.b0, sm.b1, tw, a.pf);
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
    fft_fwd_acc<2, R, T, NT, E>(Ev, yv, out, sm.b0, sm.b1, tw, a.kerr);
}

// FP32 field lines keep the RK4 input row A and the k1 + 2k2 + 2k3 + k4
// accumulator in LDS (2 x 32 KiB at N_T = 4096, after LineSmem); FP64 lines
// do not fit and use the global rows.
template <class Pol, int T>
struct RK4Rows {
    static constexpr bool in_lds = sizeof(typename Pol::R) == 4;
    static constexpr size_t bytes = in_lds ? 2 * (size_t)T * sizeof(C2<typename Pol::R>) : 0;
};

template <class Pol, int T, int E>
__global__ __launch_bounds__((Shape<Pol, T, E>::NT), GS_NL_WAVES) void k_nl_rk4(const C2<typename Pol::R>* __restrict__ field,
                                                                    C2<typename Pol::R>* __restrict__ b1, double dz,
                                                                    NLArgs a, double* __restrict__ linemax)
{
    using R = typename Pol::R;
    constexpr int NT = Shape<Pol, T, E>::NT;
    constexpr bool ROWS_LDS = RK4Rows<Pol, T>::in_lds;
    if (a.ctl) {
        if (a.ctl->state != 0) return;
        dz = a.ctl->dz;
    }
    extern __shared__ __align__(16) unsigned char smem_raw[];
    LineSmem<Pol, T, E>& sm = *reinterpret_cast<LineSmem<Pol, T, E>*>(smem_raw);
    const int line = blockIdx.x, tid = threadIdx.x;
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
    if (flags) atomicOr(&a.status->flags, flags);
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
    if (flags) atomicOr(&a.status->flags, flags);
}

// ------------------------------------------------------------------ line FFT / diagnostics
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
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int n = tid * E + e;
            x[e] = in[base + n];
            f[e] = (R)(scale * (filt ? filt[n] : 1.0));
        }
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
#pragma unroll
        for (int e = 0; e < E; ++e) out[base + tid * E + e] = x[e];
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
    if (flags) atomicOr(&a.status->flags, flags);
}

// ------------------------------------------------------------------ launchers
template <class Pol, int T, int E>
static int nl_rk4_TE(const void* field, void* b1, double dz, const NLArgs& a, double* linemax, hipStream_t st)
{
    using R = typename Pol::R;
    const size_t sm = sizeof(LineSmem<Pol, T, E>) + RK4Rows<Pol, T>::bytes;
    auto kf = k_nl_rk4<Pol, T, E>;
    HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(kf, dim3(a.nr), dim3(Shape<Pol, T, E>::NT), sm, st, (const C2<R>*)field, (C2<R>*)b1, dz, a, linemax);
    HIPCHK(hipGetLastError());
    return 0;
}

// samples per thread: the default, or 8/16 when a.nl_e asks for it (FP32/mixed, N_T >= 1024)
template <class Pol, int T>
static int nl_rk4_T(const void* field, void* b1, double dz, const NLArgs& a, double* linemax, hipStream_t st)
{
    if constexpr (sizeof(typename Pol::R) == 4 && T >= 1024) {
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

}  // namespace gs
