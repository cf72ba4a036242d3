// gs_dev.h -- device-side building blocks shared by the line kernels
// (gs_line.hip) and the QDHT / linear-step kernels (gs_qdht.hip).
#pragma once

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

// MFMA operand / result vectors
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// split2: two floats -> packed bf16 hi pair and lo pair (round to nearest
// even, v_cvt_pk_bf16_f32), hi + lo = v to about 2^-17 relative
__device__ __forceinline__ void split2(float v0, float v1, uint32_t* hi, uint32_t* lo)
{
    const f32x2_t v = {v0, v1};
    const bf16x2_t h = __builtin_convertvector(v, bf16x2_t);
    const bf16x2_t l = __builtin_convertvector(v - __builtin_convertvector(h, f32x2_t), bf16x2_t);
    *hi = __builtin_bit_cast(uint32_t, h);
    *lo = __builtin_bit_cast(uint32_t, l);
}

struct PolF64 { using R = double; using S = double; using W = double; using P = double; static constexpr bool mfma_fft = false; };
struct PolF32 { using R = float;  using S = float;  using W = float;  using P = float;  static constexpr bool mfma_fft = false; };
struct PolMix { using R = float;  using S = double; using W = float;  using P = double; static constexpr bool mfma_fft = false; };
// the mixed policy with its line FFTs on the matrix cores (split-BF16 radix-16
// passes, NLArgs::mfma_fwd / mfma_inv); a separate instantiation, so the MFMA
// code's registers do not weigh on the default (VALU) line kernel
struct PolMixMfma : PolMix { static constexpr bool mfma_fft = true; };


#define HIPCHK(x)                                  \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return (int)e_;      \
    } while (0)

// ------------------------------------------------------------------ small DFTs
// X[q] = sum_r x[r] e^{SIGN 2 pi i r q / RADIX}
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
    } else if constexpr (RADIX == 8) {
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
    } else {
        static_assert(RADIX == 16, "radix");
        // 16 = 4 x 4: r = r1 + 4 r2, q = q2 + 4 q1; inner radix-4 over r2, twiddle w16^{r1 q2}, outer over r1
        const R h = (R)0.70710678118654752440084436210485;
        const R c1 = (R)0.92387953251128675612818318939679, s1 = (R)0.38268343236508977172845998403040;
        const R sg = (R)SIGN;
        C2<R> y[4][4];
#pragma unroll
        for (int r1 = 0; r1 < 4; ++r1) {
            C2<R> t[4] = {v[r1], v[r1 + 4], v[r1 + 8], v[r1 + 12]};
            dft_small<4, SIGN>(t);
#pragma unroll
            for (int q2 = 0; q2 < 4; ++q2) y[r1][q2] = t[q2];
        }
        const C2<R> w1 = {c1, sg * s1}, w2 = {h, sg * h}, w3 = {s1, sg * c1}, w6 = {-h, sg * h}, w9 = {-c1, -sg * s1};
        y[1][1] = cmul(y[1][1], w1);
        y[1][2] = cmul(y[1][2], w2);
        y[1][3] = cmul(y[1][3], w3);
        y[2][1] = cmul(y[2][1], w2);
        y[2][2] = cmuli<SIGN>(y[2][2]);
        y[2][3] = cmul(y[2][3], w6);
        y[3][1] = cmul(y[3][1], w3);
        y[3][2] = cmul(y[3][2], w6);
        y[3][3] = cmul(y[3][3], w9);
#pragma unroll
        for (int q2 = 0; q2 < 4; ++q2) {
            C2<R> t[4] = {y[0][q2], y[1][q2], y[2][q2], y[3][q2]};
            dft_small<4, SIGN>(t);
#pragma unroll
            for (int q1 = 0; q1 < 4; ++q1) v[q2 + 4 * q1] = t[q1];
        }
    }
}

// v[r] *= w^r for r = 1..RADIX-1.  The powers come from a product tree
// (w^r = w^(r/2) w^(r - r/2), depth <= 4 for RADIX <= 16), so a pass needs
// one twiddle load instead of RADIX-1 and the rounding error stays at a few ulp.
template <int RADIX, class R>
__device__ __forceinline__ void twiddle_pow(C2<R>* v, C2<R> w)
{
    C2<R> p[RADIX];
    p[1] = w;
#pragma unroll
    for (int r = 2; r < RADIX; ++r) p[r] = cmul(p[r / 2], p[r - r / 2]);
#pragma unroll
    for (int r = 1; r < RADIX; ++r) v[r] = cmul(v[r], p[r]);
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
template <class V> __device__ __forceinline__ AffR<V> aff_id(AffR<V>*) { return {(V)1, (V)0}; }
template <class V> __device__ __forceinline__ AffC<V> aff_id(AffC<V>*) { return {(V)1, (V)0, (V)0, (V)0}; }

// DPP lane moves (VALU, no LDS traffic): lanes whose source lane does not
// exist, or whose row is masked off, receive `id` instead.
//   row_shr:n 0x110+n (within rows of 16), row_bcast:15 0x142 (lane 15 of a
//   row to the next row), row_bcast:31 0x143 (lane 31 to rows 2 and 3)
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f(float id, float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, id), __builtin_bit_cast(int, v),
                                                                  CTRL, ROWMASK, 0xf, false));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f(double id, double v)
{
    const long long i = __builtin_bit_cast(long long, id), x = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp((int)i, (int)x, CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(i >> 32), (int)(x >> 32), CTRL, ROWMASK, 0xf, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL, int ROWMASK, class V>
__device__ __forceinline__ AffR<V> dpp_aff(AffR<V> v)
{
    return {dpp_f<CTRL, ROWMASK>((V)1, v.a), dpp_f<CTRL, ROWMASK>((V)0, v.b)};
}
template <int CTRL, int ROWMASK, class V>
__device__ __forceinline__ AffC<V> dpp_aff(AffC<V> v)
{
    return {dpp_f<CTRL, ROWMASK>((V)1, v.ar), dpp_f<CTRL, ROWMASK>((V)0, v.ai), dpp_f<CTRL, ROWMASK>((V)0, v.br),
            dpp_f<CTRL, ROWMASK>((V)0, v.bi)};
}
// inclusive wave scan (lane i gets lanes 0..i composed, earlier lanes first)
template <class A>
__device__ __forceinline__ A wave_scan_incl(A v)
{
    v = compose(dpp_aff<0x111, 0xf>(v), v);
    v = compose(dpp_aff<0x112, 0xf>(v), v);
    v = compose(dpp_aff<0x114, 0xf>(v), v);
    v = compose(dpp_aff<0x118, 0xf>(v), v);
    v = compose(dpp_aff<0x142, 0xa>(v), v);
    v = compose(dpp_aff<0x143, 0xc>(v), v);
    return v;
}

// Exclusive block scan of affine maps (composition, earlier threads first).
// lds must hold NW+1 entries.  Returns the prefix; *total = all threads composed.
// value of lane `l` (wave-uniform) broadcast to the whole wave
__device__ __forceinline__ float lane_bcast(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double lane_bcast(double v, int l)
{
    const long long x = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)x, l), hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <class V> __device__ __forceinline__ AffR<V> lane_bcast(AffR<V> v, int l) { return {lane_bcast(v.a, l), lane_bcast(v.b, l)}; }
template <class V> __device__ __forceinline__ AffC<V> lane_bcast(AffC<V> v, int l)
{
    return {lane_bcast(v.ar, l), lane_bcast(v.ai, l), lane_bcast(v.br, l), lane_bcast(v.bi, l)};
}

// Exclusive block scan of affine maps (composition, earlier threads first).
// lds must hold NW+1 entries.  Returns the prefix; *total = all threads composed.
// Wave totals go through LDS once; every wave then scans the NW totals in its
// first lanes (DPP, NW <= 16) and picks its own prefix with a lane broadcast.
template <int NT, class A>
__device__ A block_scan_excl(A v, A* lds, A* total)
{
    constexpr int NW = NT / 64;
    static_assert(NW <= 16, "wave totals are scanned within one DPP row");
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const A id = aff_id((A*)nullptr);
    const A inc = wave_scan_incl(v);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    A t = lane < NW ? lds[lane] : id;
    if constexpr (NW > 1) t = compose(dpp_aff<0x111, 0xf>(t), t);
    if constexpr (NW > 2) t = compose(dpp_aff<0x112, 0xf>(t), t);
    if constexpr (NW > 4) t = compose(dpp_aff<0x114, 0xf>(t), t);
    if constexpr (NW > 8) t = compose(dpp_aff<0x118, 0xf>(t), t);
    const A wpre = wid > 0 ? lane_bcast(t, wid - 1) : id;
    *total = lane_bcast(t, NW - 1);
    A exc = dpp_aff<0x138, 0xf>(inc);      // wave_shr:1, lane 0 gets the identity
    const A res = compose(wpre, exc);
    __syncthreads();                       // lds may be reused by the next scan
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

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_max(double v)
{
    return fmax(v, dpp_f<CTRL, ROWMASK>(v, v));   // masked-off lanes see their own value
}
// maximum over the wave, valid in lane 63 (DPP max-scan)
__device__ __forceinline__ double wave_max_last(double v)
{
    v = dpp_max<0x111, 0xf>(v);
    v = dpp_max<0x112, 0xf>(v);
    v = dpp_max<0x114, 0xf>(v);
    v = dpp_max<0x118, 0xf>(v);
    v = dpp_max<0x142, 0xa>(v);
    v = dpp_max<0x143, 0xc>(v);
    return v;
}

// block-wide max of two values at once (one barrier pair); lds: 2*NW doubles
template <int NT>
__device__ void block_max2(double& a, double& b, double* lds)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    a = wave_max_last(a);
    b = wave_max_last(b);
    if (lane == 63) { lds[wid] = a; lds[NW + wid] = b; }
    __syncthreads();
    double x = lane < NW ? lds[lane] : lds[0], y = lane < NW ? lds[NW + lane] : lds[NW];
    x = wave_max_last(x);
    y = wave_max_last(y);
    a = lane_bcast(x, 63);
    b = lane_bcast(y, 63);
    __syncthreads();
}

// a thread's status bits into the plan's DevStatus; a line whose plasma scan
// took the chunk-sequential fallback (only thread 0 reports that bit) also
// counts once in fallbacks
__device__ __forceinline__ void report_flags(DevStatus* st, int flags)
{
    if (!flags) return;
    atomicOr(&st->flags, flags);
    if (flags & kFlagScanSlow) atomicAdd(&st->fallbacks, 1);
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

// ------------------------------------------------------------------ linear step
// e^{i ph} for the linear-step phase, in the policy's phase precision P.
// MIXED reduces ph to r in [-pi/4, pi/4] and a quadrant q (pi/2 split in two
// doubles) and evaluates sin r and cos r from their Taylor series in FP64
// (terms to r^13 and r^14: the first omitted terms are below 3e-14 on the
// interval), then rotates by q.  FP64 and FP32 use sincos / sincosf.
template <class Pol>
__device__ __forceinline__ void lstep_phasor(typename Pol::P ph, typename Pol::P* sn, typename Pol::P* cs)
{
    using R = typename Pol::R;
    using P = typename Pol::P;
    if constexpr (sizeof(R) == 8) {
        sincos(ph, sn, cs);
    } else if constexpr (sizeof(P) == 4) {
        sincosf(ph, sn, cs);
    } else {
        const double q = rint(ph * 0.63661977236758134);            // 2/pi
        double r = fma(-q, 1.5707963267948966, ph);                  // pi/2, high part
        r = fma(-q, 6.123233995736766e-17, r);                       // pi/2, low part
        const double r2 = r * r;
        double sp = 1.0 / 6227020800.0, cp = 1.0 / 87178291200.0;          // 1/13!, 1/14!
        sp = fma(sp, -r2, 1.0 / 39916800.0);                                // Horner from the top
        sp = fma(sp, -r2, 1.0 / 362880.0);
        sp = fma(sp, -r2, 1.0 / 5040.0);
        sp = fma(sp, -r2, 1.0 / 120.0);
        sp = fma(sp, -r2, 1.0 / 6.0);
        const double sv = fma(-r * r2, sp, r);
        cp = fma(cp, -r2, 1.0 / 479001600.0);
        cp = fma(cp, -r2, 1.0 / 3628800.0);
        cp = fma(cp, -r2, 1.0 / 40320.0);
        cp = fma(cp, -r2, 1.0 / 720.0);
        cp = fma(cp, -r2, 1.0 / 24.0);
        cp = fma(cp, -r2, 0.5);
        const double cv = fma(-r2, cp, 1.0);
        const int iq = (int)(((long long)q) & 3);
        *sn = iq == 0 ? sv : (iq == 1 ? cv : (iq == 2 ? -sv : -cv));
        *cs = iq == 0 ? cv : (iq == 1 ? -sv : (iq == 2 ? -cv : sv));
    }
}

// 32 well-mixed bits from a point index and the bits of its inputs (the
// finaliser of MurmurHash3); deterministic, so reruns are bit-identical
__host__ __device__ __forceinline__ uint32_t mix_bits(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// x rounded to one of its two FP32 neighbours, the farther one with
// probability |x - nearest| / (neighbour spacing): the low 29 bits of u are
// added to the 29 FP64 mantissa bits below FP32 precision and those bits are
// then cleared (truncation of the magnitude), so the rounding error has zero
// mean.  A float state updated by steps whose increments fall below half an
// ulp (round to nearest drops them every step) then drifts as a random walk
// instead of coherently.  Exact for every x in the FP32 normal range.
__host__ __device__ __forceinline__ float round_unbiased(double x, uint32_t u)
{
    unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    b = (b + (u & 0x1fffffffu)) & ~0x1fffffffull;
    return (float)__builtin_bit_cast(double, b);
}

// one point of calculate_Lstep: B2 += dz B1k g; FIELD = B2 e^{i dz rate} amp; B2 = FIELD,
// with amp = e^{dz Im k(omega)} of the point's column; B2 lives in fk (LstepArgs).
// MIXED (FP32 state, FP64 phase) updates in FP64 and stores the state with
// unbiased rounding: at small dz the per-step B1 increment is below half an
// FP32 ulp of B2 on most points, and round-to-nearest would drop it every step.
// The rounding draw hashes the step's seed (its STEP_N) with the point index
// and the point's inputs, so a point whose inputs repeat from step to step
// (no nonlinearity, sub-ulp phase rotation) still gets a fresh draw each step.
template <class Pol>
__device__ __forceinline__ C2<typename Pol::R> lstep_value(size_t o, C2<typename Pol::R> b, typename Pol::R yr,
                                                           typename Pol::R yi, typename Pol::R g, typename Pol::P rate,
                                                           typename Pol::P amp, double dz, uint32_t seed)
{
    using R = typename Pol::R;
    using P = typename Pol::P;
    P sn, cs;
    lstep_phasor<Pol>((P)dz * rate, &sn, &cs);
    if constexpr (sizeof(P) > sizeof(R)) {
        const double dg = dz * (double)g;
        const double br = fma(dg, (double)yr, (double)b.x), bi = fma(dg, (double)yi, (double)b.y);
        const double fx = amp * (br * cs - bi * sn), fy = amp * (br * sn + bi * cs);
        const uint32_t h = mix_bits(((uint32_t)o * 0x9e3779b1u + mix_bits(seed * 0xc2b2ae3du + 0x165667b1u)) ^
                                    __float_as_uint(b.x) ^ (__float_as_uint(yi) << 7) ^ (__float_as_uint(yr) * 0x27d4eb2fu));
        return C2<R>{round_unbiased(fx, h), round_unbiased(fy, (h >> 16) | (h << 16))};
    } else {
        const R dg = (R)dz * g;
        const R br = b.x + dg * yr, bi = b.y + dg * yi;
        const R a = (R)amp, c = (R)cs, s = (R)sn;
        return C2<R>{a * (br * c - bi * s), a * (br * s + bi * c)};
    }
}

template <class Pol>
__device__ __forceinline__ void lstep_point(size_t o, typename Pol::R yr, typename Pol::R yi, typename Pol::P amp,
                                            const LstepArgs& la)
{
    using R = typename Pol::R;
    using P = typename Pol::P;
    C2<R>* fk = (C2<R>*)la.fk;
    const R g = ((const R*)la.gfac)[o];
    if (g < (R)0) { fk[o] = C2<R>{0, 0}; return; }
    fk[o] = lstep_value<Pol>(o, fk[o], yr, yi, g, ((const P*)la.rate)[o], amp, la.dz, la.seed);
}

// lstep_point over NB points with all table and state loads issued before any
// store (the stores could alias later loads, so a per-point loop would
// serialise one memory round trip per point)
template <class Pol, int NB>
__device__ __forceinline__ void lstep_batch(const size_t (&o)[NB], const typename Pol::R (&yr)[NB],
                                            const typename Pol::R (&yi)[NB], typename Pol::P amp, const LstepArgs& la)
{
    using R = typename Pol::R;
    using P = typename Pol::P;
    C2<R>* fk = (C2<R>*)la.fk;
    R g[NB];
    P rt[NB];
    C2<R> b[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        g[j] = ((const R*)la.gfac)[o[j]];
        rt[j] = ((const P*)la.rate)[o[j]];
        b[j] = fk[o[j]];
    }
    // a real branch, not a select: whole waves of masked columns skip the update
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        C2<R> v = {0, 0};
        if (g[j] >= (R)0) v = lstep_value<Pol>(o[j], b[j], yr[j], yi[j], g[j], rt[j], amp, la.dz, la.seed);
        fk[o[j]] = v;
    }
}

// e^{dz Im k(omega)} of a column, in the phase precision
template <class Pol>
__device__ __forceinline__ typename Pol::P lstep_amp(int w, const LstepArgs& la)
{
    using P = typename Pol::P;
    return (P)exp((P)la.dz * (P)la.k2[w]);
}

template <class Pol>
__global__ void k_lstep(const C2<typename Pol::R>* __restrict__ b1k, int nt, size_t n, LstepArgs la)
{
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const C2<typename Pol::R> y = b1k[o];
    lstep_point<Pol>(o, y.x, y.y, lstep_amp<Pol>((int)(o % nt), la), la);
}

}  // namespace gs
