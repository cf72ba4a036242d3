// gs_qdht.hip -- the QDHT (fhatha.cpp:218-230) as a real GEMM on MFMA with
// the linear step (propagator.cu:570-625) fused into the epilogue, the
// stand-alone linear step, and small helper kernels (reductions, precision
// conversions, average spectrum).
//
//  Y[R][2T] = A[R][R] X[R][2T], A = diag(m1) C diag(1/m1), R padded to the
//  tile.  Complex columns are split into re/im planes when staged into LDS,
//  so one lane ends up owning both parts of an output and can apply the
//  linear step in registers.
//   FP32   v_mfma_f32_32x32x2_f32, tiles 128 x 64 or 128 x 32 complex, K step
//          16, one or two K groups per workgroup
//   MIXED  split-BF16 (FP32 accumulate) on v_mfma_f32_32x32x16_bf16: the
//          march's transforms on narrow 128 x 32 tiles ("t32", two bf16 parts
//          and 3 products into the linear step, three parts and 6 products
//          back, one or two K groups), full transforms of the API on
//          128 x 128 tiles ("t128"); K step 32, software-pipelined over
//          double-buffered LDS
//   FP64   v_mfma_f64_16x16x4_f64, tile 64 x 32 complex, K step 16
//  The march's transforms can skip the 32-column blocks the linear step
//  zeroes (QCols); tile shapes that give the same bits (K groups, tile
//  width, tile order over the XCDs) are chosen by the plan's autotuner.
#include "gs_dev.h"

#include <cstdlib>

namespace gs {

// ------------------------------------------------------------------ QDHT GEMMs
typedef double v4d __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// blockIdx -> tile.  Dispatch is round-robin over the 8 XCDs, each with its own
// L2: block b runs on XCD b & 7, which gets one contiguous run of tiles.
// rows = false: all row tiles of a column tile on one XCD (its B columns are
// read into that L2 once); rows = true: all column tiles of a row tile (its
// rows of the operator A are), the better choice when A is larger than an L2
// and the column tiles are few.  Same arithmetic either way.
__device__ __forceinline__ void tile_of(int b, int mt, int ntl, int* mi, int* ni, bool rows = false)
{
    const int tiles = mt * ntl;
    int t = b;
    if ((tiles & 7) == 0) t = (b & 7) * (tiles >> 3) + (b >> 3);
    if (rows) {
        *mi = t / ntl;
        *ni = t % ntl;
    } else {
        *mi = t % mt;
        *ni = t / mt;
    }
}

// First column of column tile ni: the listed 32-column block of a
// column-skipping transform, or ni * BNC for the full transform.
__device__ __forceinline__ int col0_of(const QCols& qc, int ni, int BNC) { return qc.ctile ? qc.ctile[ni] : ni * BNC; }

// Blocks past the computed tiles of a column-skipping QDHT back: block z
// clears rows [m0, m0 + BM) of up to 8 skipped 32-column blocks of Y, when
// asked to (zero_dead, or the device controller's dead_dirty of an accepted
// attempt; the caller has already returned for a rejected one).
template <class T>
__device__ __forceinline__ void zero_skipped(T* __restrict__ Y, int NC, int mt, int BM, int z, const QCols& qc,
                                             const DevCtl* ctl)
{
    if (!(ctl ? ctl->dead_dirty : qc.zero_dead)) return;
    const int m0 = (z % mt) * BM, j0 = (z / mt) * 8, j1 = min(j0 + 8, qc.nzt);
    for (int j = j0; j < j1; ++j) {
        const int c0 = qc.ztile[j];
        for (int i = threadIdx.x; i < BM * 32; i += blockDim.x) Y[(size_t)(m0 + i / 32) * NC + c0 + (i & 31)] = T{0, 0};
    }
}

// FP32: tile 128 rows x BNC complex columns, K step 16, 4 waves: 2 x 2 waves
// of 64 rows x 32 columns (BNC = 64, the full transform) or 4 x 1 waves of
// 32 rows x 32 columns (BNC = 32, the column-skipping transform: ~900 live
// columns at N_T = 4096 still give 29 x 8 tiles).  Each output sums its k
// terms in the same order at either width, so the two agree bit for bit.
// FP32: 128 rows x BNC complex columns per group of 4 waves, K step 16, A and B
// staged in LDS.  As in the split-BF16 narrow tiles every output is (sum over
// the first K/2 rows) + (sum over the last K/2 rows), each half in k order:
// KS = 1 group sums both halves in turn, KS = 2 groups (512 threads) one half
// each, added through LDS, each group writing half of the rows.  Same bits
// for every (BNC, KS).  sm holds smem_f32<BNC, KS>() floats.
template <int BNC, int KS>
constexpr int smem_f32()
{
    constexpr int tiles = KS * 16 * ((128 + 32) + 2 * (BNC + 32));                 // As, Br, Bi per group
    constexpr int xch = (KS - 1) * 2 * (BNC == 64 ? 2 : 1) * 2 * 8 * 256;          // half the sums, both ways
    return tiles > xch ? tiles : xch;
}
template <bool LSTEP, int BNC, int KS = 1>
__device__ __forceinline__ void qdht_f32_body(int bid, const float* __restrict__ At, const C2<float>* __restrict__ X,
                                              C2<float>* __restrict__ Y, int K, int NC, LstepArgs la, const QCols& qc,
                                              float* __restrict__ sm)
{
    constexpr int BM = 128, BK = 16, PA = 32, PB = 32;
    constexpr int RB = BNC == 64 ? 2 : 1;      // 32-row blocks per wave
    static_assert(BNC == 64 || BNC == 32, "tile width");
    static_assert(KS == 1 || KS == 2, "one or two K groups");
    const DevCtl* ctl = la.ctl;
    if (ctl) {
        if (ctl->state != 0 || !ctl->accept) return;
        la.dz = ctl->dz_lin;
        la.seed = (uint32_t)ctl->steps;
    }
    const int M = K;
    const int mt = M / BM, nct = qc.ctile ? qc.nct : NC / BNC;
    if (bid >= mt * nct) {
        if constexpr (!LSTEP) zero_skipped(Y, NC, mt, BM, bid - mt * nct, qc, ctl);
        return;
    }
    int mi, ni;
    tile_of(bid, mt, nct, &mi, &ni, qc.order != 0);
    const int m0 = mi * BM, n0 = col0_of(qc, ni, BNC);
    const int tid = threadIdx.x & 255, lane = tid & 63, w = tid >> 6;
    const int ks = KS == 1 ? 0 : (int)(threadIdx.x >> 8);
    const int wm = BNC == 64 ? (w >> 1) * 64 : w * 32, wn = BNC == 64 ? (w & 1) * 32 : 0;
    const int lr = tid >> 4, lc = (tid & 15);
    typedef float TA[BM + PA];
    typedef float TB[BNC + PB];
    TA* As = reinterpret_cast<TA*>(sm + ks * 16 * ((BM + PA) + 2 * (BNC + PB)));
    TB* Br = reinterpret_cast<TB*>(reinterpret_cast<float*>(As) + BK * (BM + PA));
    TB* Bi = reinterpret_cast<TB*>(reinterpret_cast<float*>(Br) + BK * (BNC + PB));

    v16f acc[RB][2], first[RB][2];
#pragma unroll
    for (int a = 0; a < RB; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][c][r] = 0.f;

    constexpr int NB4 = BNC / 32;       // float4 (two complex values) of B per thread
    float4 ra0, ra1, rb[NB4];
    auto gload = [&](int k0) {
        const float* pa = At + (size_t)(k0 + lr) * M + m0 + lc * 8;
        ra0 = *(const float4*)pa;
        ra1 = *(const float4*)(pa + 4);
        const float* pb = (const float*)(X + (size_t)(k0 + lr) * NC + n0 + lc * 2 * NB4);
#pragma unroll
        for (int j = 0; j < NB4; ++j) rb[j] = *(const float4*)(pb + 4 * j);
    };
    auto sstore = [&]() {
        *(float4*)&As[lr][lc * 8] = ra0;
        *(float4*)&As[lr][lc * 8 + 4] = ra1;
#pragma unroll
        for (int j = 0; j < NB4; ++j) {
            const int c = lc * 2 * NB4 + 2 * j;
            Br[lr][c] = rb[j].x; Bi[lr][c] = rb[j].y;
            Br[lr][c + 1] = rb[j].z; Bi[lr][c + 1] = rb[j].w;
        }
    };
    const int kh = K / 2, kbeg = ks * (K / KS), kend = kbeg + K / KS;
    gload(kbeg);
    sstore();
    __syncthreads();
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        if constexpr (KS == 1) {
            if (k0 == kh) {
#pragma unroll
                for (int a = 0; a < RB; ++a)
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        first[a][c] = acc[a][c];
#pragma unroll
                        for (int r = 0; r < 16; ++r) acc[a][c][r] = 0.f;
                    }
            }
        }
        const bool more = k0 + BK < kend;
        if (more) gload(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 2) {
            const int kr = kk + (lane >> 5), cl = lane & 31;
            const float br = Br[kr][wn + cl], bi = Bi[kr][wn + cl];
#pragma unroll
            for (int rb2 = 0; rb2 < RB; ++rb2) {
                const float a = As[kr][wm + 32 * rb2 + cl];
                acc[rb2][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, br, acc[rb2][0], 0, 0, 0);
                acc[rb2][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bi, acc[rb2][1], 0, 0, 0);
            }
        }
        __syncthreads();
        if (more) {
            sstore();
            __syncthreads();
        }
    }
    if constexpr (KS == 1) {
#pragma unroll
        for (int a = 0; a < RB; ++a)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[a][c][r] = first[a][c][r] + acc[a][c][r];
    }
    if constexpr (KS == 2) {
        // group g keeps rows r >> 3 == g of each 32-row block; the other half of its
        // sums goes to the partner through LDS (lane-contiguous), then each adds the
        // partner's: y = first half + second half in both groups
        constexpr int NX = RB * 2 * 8;     // floats handed over per lane
        if (ks == 0) {
#pragma unroll
            for (int a = 0; a < RB; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int j = 0; j < 8; ++j) sm[((NX + (a * 2 + c) * 8 + j) * 256) + tid] = acc[a][c][8 + j];
        } else {
#pragma unroll
            for (int a = 0; a < RB; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int j = 0; j < 8; ++j) sm[(((a * 2 + c) * 8 + j) * 256) + tid] = acc[a][c][j];
        }
        __syncthreads();
        if (ks == 0) {
#pragma unroll
            for (int a = 0; a < RB; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int j = 0; j < 8; ++j) acc[a][c][j] += sm[(((a * 2 + c) * 8 + j) * 256) + tid];
        } else {
#pragma unroll
            for (int a = 0; a < RB; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int j = 0; j < 8; ++j) acc[a][c][8 + j] += sm[((NX + (a * 2 + c) * 8 + j) * 256) + tid];
        }
    }
    const int col = n0 + wn + (lane & 31);
    const float amp = LSTEP ? lstep_amp<PolF32>(col, la) : 1.f;
#pragma unroll
    for (int rb2 = 0; rb2 < RB; ++rb2)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            if (KS == 2 && (r >> 3) != ks) continue;
            const int row = m0 + wm + rb2 * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const size_t o = (size_t)row * NC + col;
            if (LSTEP) lstep_point<PolF32>(o, acc[rb2][0][r], acc[rb2][1][r], amp, la);
            else Y[o] = C2<float>{acc[rb2][0][r], acc[rb2][1][r]};
        }
}
template <bool LSTEP, int BNC, int KS>
__global__ __launch_bounds__(256 * KS) void k_qdht_f32(const float* __restrict__ At, const C2<float>* __restrict__ X,
                                                       C2<float>* __restrict__ Y, int K, int NC, LstepArgs la, QCols qc)
{
    __shared__ __align__(16) float sm[smem_f32<BNC, KS>()];
    qdht_f32_body<LSTEP, BNC, KS>(blockIdx.x, At, X, Y, K, NC, la, qc, sm);
}

// BF16 3-pass: same tile, K step 32, operands k-contiguous in LDS (split2 in gs_dev.h).

// BF16 3-pass, large tile: 128 rows x 128 complex columns per workgroup,
// 4 waves of 64 x 64 complex, K step 32 (two MFMA k-slices of 16).
// The kernel is bound by LDS bandwidth, so only B goes through LDS: the
// QDHT operator A is stored in MFMA fragment order (one coalesced 1 KiB
// load per 32 x 16 fragment) and read from L2 straight into registers one K
// step ahead.  B (the field, FP32) is loaded one column per lane (16 k rows,
// coalesced across lanes) two K steps ahead, split into bf16 hi/lo and stored
// k-contiguous in a double-buffered LDS tile; one barrier per K step.
namespace t128 {
constexpr int BM = 128, BNC = 128, BK = 32, LD = BK + 8;
constexpr int PLANE = BNC * LD;         // elements of one [128 columns][LD] plane
constexpr int BUF = 4 * PLANE;          // B hi re, B hi im, B lo re, B lo im
constexpr size_t SMEM = (size_t)2 * BUF * sizeof(uint16_t);
// A fragment index: [m tile][k step][k slice (2)][32-row block (4)][hi/lo][lane][8]
__host__ __device__ constexpr size_t afrag(int mt, int nk, int step, int kk, int rbb, int hl)
{
    return ((((((size_t)mt * nk + step) * 2 + kk) * 4 + rbb) * 2 + hl) * 64) * 8;
}
// global loads of one K step: 16 complex B values of one column
__device__ __forceinline__ void gload_b(const C2<float>* gX, int NC, int step, int nk, float (&xr)[16], float (&xi)[16])
{
    step = step < nk ? step : nk - 1;      // past the end: reload the last step (never used)
    const C2<float>* xp = gX + (size_t)step * BK * NC;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const C2<float> v = xp[(size_t)j * NC];
        xr[j] = v.x;
        xi[j] = v.y;
    }
}
// the wave's A fragments of one K step: [kk][rb][hi/lo]
__device__ __forceinline__ void gload_a(const uint16_t* A, int mi, int nk, int step, int rbb0, int lane, v8bf (&a)[2][2][2])
{
    step = step < nk ? step : nk - 1;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int hl = 0; hl < 2; ++hl)
                a[kk][rb][hl] = *(const v8bf*)(A + afrag(mi, nk, step, kk, rbb0 + rb, hl) + lane * 8);
}
// split of one K step of B into the LDS buffer b
__device__ __forceinline__ void sstore_b(uint16_t* b, int oB, const float (&xr)[16], const float (&xi)[16])
{
    uint32_t hr[8], lr[8], hi[8], li[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        split2(xr[2 * j], xr[2 * j + 1], &hr[j], &lr[j]);
        split2(xi[2 * j], xi[2 * j + 1], &hi[j], &li[j]);
    }
    *(u32x4*)(b + oB) = u32x4{hr[0], hr[1], hr[2], hr[3]};
    *(u32x4*)(b + oB + 8) = u32x4{hr[4], hr[5], hr[6], hr[7]};
    *(u32x4*)(b + oB + PLANE) = u32x4{hi[0], hi[1], hi[2], hi[3]};
    *(u32x4*)(b + oB + PLANE + 8) = u32x4{hi[4], hi[5], hi[6], hi[7]};
    *(u32x4*)(b + oB + 2 * PLANE) = u32x4{lr[0], lr[1], lr[2], lr[3]};
    *(u32x4*)(b + oB + 2 * PLANE + 8) = u32x4{lr[4], lr[5], lr[6], lr[7]};
    *(u32x4*)(b + oB + 3 * PLANE) = u32x4{li[0], li[1], li[2], li[3]};
    *(u32x4*)(b + oB + 3 * PLANE + 8) = u32x4{li[4], li[5], li[6], li[7]};
}
}  // namespace t128

template <bool LSTEP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_qdht_bf16x3_t128(const uint16_t* __restrict__ Afr,
                                                          const C2<float>* __restrict__ X, C2<float>* __restrict__ Y,
                                                          int K, int NC, LstepArgs la)
{
    using namespace t128;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    uint16_t* const sm = reinterpret_cast<uint16_t*>(smem_raw);
    if (la.ctl) {
        if (la.ctl->state != 0 || !la.ctl->accept) return;
        la.dz = la.ctl->dz_lin;
        la.seed = (uint32_t)la.ctl->steps;
    }

    const int M = K;
    int mi, ni;
    tile_of(blockIdx.x, M / BM, NC / BNC, &mi, &ni);
    const int m0 = mi * BM, n0 = ni * BNC;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
    const int bcol = tid & 127, bk = (tid >> 7) * 16;     // B loader: column, k half
    const int cl = lane & 31, kq = 8 * (lane >> 5);
    const int oB = bcol * LD + bk;                        // + PLANE per B plane
    const int rB = (wn + cl) * LD + kq;                   // + plane*PLANE, + cb*32*LD, + kk
    const int rbb0 = wm / 32;
    const C2<float>* gX = X + (size_t)bk * NC + n0 + bcol;
    const int nk = K / BK;

    v16f acc[2][2][2];    // [rb][part][cb]
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[a][b][c][r] = 0.f;

    // the B operands of one K step come out of LDS before the first MFMA
    struct OpsB {
        v8bf h[2][2][2], l[2][2][2];   // [kk][pt][cb]
    };
    auto read_b = [&](const uint16_t* b, OpsB& o) {
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
#pragma unroll
                for (int cb = 0; cb < 2; ++cb) {
                    o.h[kk][pt][cb] = *(const v8bf*)(b + rB + pt * PLANE + cb * 32 * LD + 16 * kk);
                    o.l[kk][pt][cb] = *(const v8bf*)(b + rB + (2 + pt) * PLANE + cb * 32 * LD + 16 * kk);
                }
    };
    auto mfma_ops = [&](const v8bf (&a)[2][2][2], const OpsB& o) {
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int pt = 0; pt < 2; ++pt)
#pragma unroll
                    for (int cb = 0; cb < 2; ++cb)
                        acc[rb][pt][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[kk][rb][1], o.h[kk][pt][cb], acc[rb][pt][cb], 0, 0, 0);
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int pt = 0; pt < 2; ++pt)
#pragma unroll
                    for (int cb = 0; cb < 2; ++cb)
                        acc[rb][pt][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[kk][rb][0], o.l[kk][pt][cb], acc[rb][pt][cb], 0, 0, 0);
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int pt = 0; pt < 2; ++pt)
#pragma unroll
                    for (int cb = 0; cb < 2; ++cb)
                        acc[rb][pt][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[kk][rb][0], o.h[kk][pt][cb], acc[rb][pt][cb], 0, 0, 0);
        }
    };
    // interleave the 48 MFMAs of a step with the split VALU, the LDS writes
    // and the global loads of the other stages
    auto schedule = [&]() {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
            __builtin_amdgcn_sched_group_barrier(0x002, 6, 0);   // VALU
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);   // DS write
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x020, 2, 0);   // VMEM read
        }
    };

    uint16_t* const buf0 = sm;
    uint16_t* const buf1 = sm + BUF;
    float xr[16], xi[16], nxr[16], nxi[16];
    v8bf a0[2][2][2], a1[2][2][2];
    gload_b(gX, NC, 0, nk, xr, xi);
    sstore_b(buf0, oB, xr, xi);
    gload_b(gX, NC, 1, nk, xr, xi);
    gload_a(Afr, mi, nk, 0, rbb0, lane, a0);
    __syncthreads();
    OpsB ops;
    for (int k = 0; k < nk; k += 2) {
        read_b(buf0, ops);
        __builtin_amdgcn_sched_barrier(0);
        gload_b(gX, NC, k + 2, nk, nxr, nxi);
        gload_a(Afr, mi, nk, k + 1, rbb0, lane, a1);
        mfma_ops(a0, ops);
        sstore_b(buf1, oB, xr, xi);
        schedule();
        __syncthreads();
        read_b(buf1, ops);
        __builtin_amdgcn_sched_barrier(0);
        gload_b(gX, NC, k + 3, nk, xr, xi);
        gload_a(Afr, mi, nk, k + 2, rbb0, lane, a0);
        mfma_ops(a1, ops);
        sstore_b(buf0, oB, nxr, nxi);
        schedule();
        __syncthreads();
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const int col = n0 + wn + cb * 32 + (lane & 31);
        const double amp = LSTEP ? lstep_amp<PolMix>(col, la) : 1.0;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                size_t o[8];
                float yr[8], yi[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int r = 8 * h + j;
                    const int row = m0 + wm + rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    o[j] = (size_t)row * NC + col;
                    yr[j] = acc[rb][0][cb][r];
                    yi[j] = acc[rb][1][cb][r];
                }
                if constexpr (LSTEP) {
                    lstep_batch<PolMix, 8>(o, yr, yi, amp, la);
                } else {
#pragma unroll
                    for (int j = 0; j < 8; ++j) Y[o[j]] = C2<float>{yr[j], yi[j]};
                }
            }
    }
}
constexpr size_t kBf16T128Smem = t128::SMEM;

// BF16 3-pass, narrow tile of the column-skipping transforms: 128 rows x 32
// complex columns per workgroup, 4 waves of 32 rows x 32 columns, K step 32
// (two MFMA k-slices of 16).  At N_T = 4096 about 900 of the columns are live,
// which gives 29 x 8 tiles for 256 CUs where 128-column tiles would give 64.
// The operator fragments are the t128 ones (each wave reads its 32-row block
// from L2 into registers one K step ahead); B is split once per workgroup
// into bf16 hi/lo planes in a double-buffered LDS tile shared by the four
// waves.  Every output takes its MFMA terms in the t128 order (per k slice
// lo*hi, hi*lo, hi*hi), so the two kernels agree bit for bit.
namespace t32 {
constexpr int BM = 128, BNC = 32, BK = 32, LD = BK + 8;
constexpr int PLANE = BNC * LD;         // one [32 columns][LD] plane
// a buffer holds NS bf16 parts of the K step's B, re and im: plane 2 part + (0 re, 1 im)
template <int NS> constexpr int buf_elems() { return 2 * NS * PLANE; }
// global loads of one K step: 4 consecutive k rows of one column
__device__ __forceinline__ void gload_b(const C2<float>* gX, int NC, int step, int nk, float (&xr)[4], float (&xi)[4])
{
    step = step < nk ? step : nk - 1;      // past the end: reload the last step (never used)
    const C2<float>* xp = gX + (size_t)step * BK * NC;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const C2<float> v = xp[(size_t)j * NC];
        xr[j] = v.x;
        xi[j] = v.y;
    }
}
// the wave's A fragments of one K step: [kk][part] of its 32-row block; parts 0, 1
// (hi, lo) interleaved as in t128, part 2 (the remainder) in a plane of its own
// behind them (A + 2 K^2)
template <int NS>
__device__ __forceinline__ void gload_a(const uint16_t* A, int K, int mi, int nk, int step, int rbb, int lane,
                                        v8bf (&a)[2][NS])
{
    step = step < nk ? step : nk - 1;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
        for (int hl = 0; hl < 2; ++hl) a[kk][hl] = *(const v8bf*)(A + t128::afrag(mi, nk, step, kk, rbb, hl) + lane * 8);
        if constexpr (NS == 3)
            a[kk][2] = *(const v8bf*)(A + (size_t)2 * K * K + t128::afrag(mi, nk, step, kk, rbb, 0) / 2 + lane * 8);
    }
}
// the K step's B split into NS bf16 parts (x = hi + lo (+ rest), each the
// nearest bf16 of what the earlier parts leave) into LDS buffer b
template <int NS>
__device__ __forceinline__ void sstore_b(uint16_t* b, int oB, const float (&xr)[4], const float (&xi)[4])
{
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int pt = 0; pt < 2; ++pt) {
        const float(&x)[4] = pt ? xi : xr;
        uint32_t h[2], l[2], r2[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            split2(x[2 * j], x[2 * j + 1], &h[j], &l[j]);
            if constexpr (NS == 3) {
                const f32x2_t v = {x[2 * j], x[2 * j + 1]};
                const f32x2_t rem = v - __builtin_convertvector(__builtin_bit_cast(bf16x2_t, h[j]), f32x2_t) -
                                    __builtin_convertvector(__builtin_bit_cast(bf16x2_t, l[j]), f32x2_t);
                r2[j] = __builtin_bit_cast(uint32_t, __builtin_convertvector(rem, bf16x2_t));
            }
        }
        *(u32x2*)(b + oB + pt * PLANE) = u32x2{h[0], h[1]};
        *(u32x2*)(b + oB + (2 + pt) * PLANE) = u32x2{l[0], l[1]};
        if constexpr (NS == 3) *(u32x2*)(b + oB + (4 + pt) * PLANE) = u32x2{r2[0], r2[1]};
    }
}
}  // namespace t32

// NS = 2: split-BF16 3-pass (lo*hi, hi*lo, hi*hi); NS = 3: both operands in three
// bf16 parts (to ~2^-24, the FP32 accuracy of the field), six products
// (mid*mid, hi*rest, rest*hi, mid*hi, hi*mid, hi*hi, smallest first).
// 128 rows x 32 complex columns per workgroup of 256 x KS threads.  Every
// output is (sum over the first K/2 rows) + (sum over the last K/2 rows), each
// half summed in k order: with KS = 1 one group of 4 waves sums both halves
// in turn (the first kept aside), with KS = 2 two groups sum one half each and
// add them through LDS, each group writing half of the rows.  Both give the
// same bits, so KS is a free choice of tile shape (two waves per SIMD on the
// few, ~1 per CU, tiles of a column-skipping transform).  sm holds
// smem_t32<NS, KS>() elements.
template <int NS, int KS>
constexpr int smem_t32()
{
    // double-buffered B per group; the exchange of 2 x 16 x 256 floats reuses it
    return 2 * KS * t32::buf_elems<NS>() > (KS - 1) * 2 * 16 * 256 * 2 ? 2 * KS * t32::buf_elems<NS>()
                                                                       : (KS - 1) * 2 * 16 * 256 * 2;
}
template <bool LSTEP, int NS = 2, int KS = 1>
__device__ __forceinline__ void qdht_t32_body(int bid, const uint16_t* __restrict__ Afr, const C2<float>* __restrict__ X,
                                              C2<float>* __restrict__ Y, int K, int NC, LstepArgs la, const QCols& qc,
                                              uint16_t* __restrict__ sm)
{
    using namespace t32;
    static_assert(KS == 1 || KS == 2, "one or two K ranges");
    constexpr int BUF = buf_elems<NS>();
    const DevCtl* ctl = la.ctl;
    if (ctl) {
        if (ctl->state != 0 || !ctl->accept) return;
        la.dz = ctl->dz_lin;
        la.seed = (uint32_t)ctl->steps;
    }
    const int M = K;
    const int mt = M / BM, nct = qc.ctile ? qc.nct : NC / BNC;
    if (bid >= mt * nct) {
        if constexpr (!LSTEP) zero_skipped(Y, NC, mt, BM, bid - mt * nct, qc, ctl);
        return;
    }
    int mi, ni;
    tile_of(bid, mt, nct, &mi, &ni, qc.order != 0);
    const int m0 = mi * BM, n0 = col0_of(qc, ni, BNC);
    const int tid = threadIdx.x & 255, lane = tid & 63, w = tid >> 6;   // within the K group
    const int ks = KS == 1 ? 0 : (int)(threadIdx.x >> 8);                // the K group
    const int bcol = tid & 31, bq = tid >> 5;               // B loader: column, k quarter (4 rows)
    const int oB = bcol * LD + 4 * bq;                       // + plane * PLANE
    const int rB = (lane & 31) * LD + 8 * (lane >> 5);      // + plane * PLANE + 16 kk
    const int nka = K / BK, nk = nka / KS, kb = ks * nk;      // K steps: all, per group, group's first
    const C2<float>* gX = X + (size_t)(4 * bq + kb * BK) * NC + n0 + bcol;

    v16f acc[2];    // [re, im]
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[a][r] = 0.f;

    struct OpsB {
        v8bf p[NS][2][2];   // [part][kk][pt]
    };
    auto read_b = [&](const uint16_t* b, OpsB& o) {
#pragma unroll
        for (int q = 0; q < NS; ++q)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk)
#pragma unroll
                for (int pt = 0; pt < 2; ++pt) o.p[q][kk][pt] = *(const v8bf*)(b + rB + (2 * q + pt) * PLANE + 16 * kk);
    };
    auto mfma_ops = [&](const v8bf (&a)[2][NS], const OpsB& o) {
        auto mm = [&](int kk, int qa, int qb) {
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
                acc[pt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[kk][qa], o.p[qb][kk][pt], acc[pt], 0, 0, 0);
        };
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            if constexpr (NS == 3) {
                mm(kk, 1, 1);
                mm(kk, 0, 2);
                mm(kk, 2, 0);
            }
            mm(kk, 1, 0);
            mm(kk, 0, 1);
            mm(kk, 0, 0);
        }
    };
    // interleave the MFMAs of a step with the split VALU, the LDS writes and the global loads
    auto schedule = [&]() {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, NS - 1, 0);   // MFMA
            __builtin_amdgcn_sched_group_barrier(0x002, 4 * (NS - 1), 0);   // VALU
            __builtin_amdgcn_sched_group_barrier(0x008, NS - 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x200, NS - 1, 0);   // DS write
            __builtin_amdgcn_sched_group_barrier(0x008, NS - 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x020, NS, 0);       // VMEM read
        }
    };

    uint16_t* const buf0 = sm + ks * 2 * BUF;
    uint16_t* const buf1 = buf0 + BUF;
    float xr[4], xi[4], nxr[4], nxi[4];
    v8bf a0[2][NS], a1[2][NS];
    v16f first[2];     // KS = 1: the first half's sums
    gload_b(gX, NC, 0, nk, xr, xi);
    sstore_b<NS>(buf0, oB, xr, xi);
    gload_b(gX, NC, 1, nk, xr, xi);
    gload_a<NS>(Afr, K, mi, nka, kb, w, lane, a0);
    __syncthreads();
    OpsB ops;
    for (int k = 0; k < nk; k += 2) {
        if constexpr (KS == 1) {
            if (k == nk / 2) {
#pragma unroll
                for (int pt = 0; pt < 2; ++pt) {
                    first[pt] = acc[pt];
#pragma unroll
                    for (int r = 0; r < 16; ++r) acc[pt][r] = 0.f;
                }
            }
        }
        read_b(buf0, ops);
        __builtin_amdgcn_sched_barrier(0);
        gload_b(gX, NC, k + 2, nk, nxr, nxi);
        gload_a<NS>(Afr, K, mi, nka, kb + k + 1, w, lane, a1);
        mfma_ops(a0, ops);
        sstore_b<NS>(buf1, oB, xr, xi);
        schedule();
        __syncthreads();
        read_b(buf1, ops);
        __builtin_amdgcn_sched_barrier(0);
        gload_b(gX, NC, k + 3, nk, xr, xi);
        gload_a<NS>(Afr, K, mi, nka, kb + k + 2, w, lane, a0);
        mfma_ops(a1, ops);
        sstore_b<NS>(buf0, oB, nxr, nxi);
        schedule();
        __syncthreads();
    }
    if constexpr (KS == 1) {
#pragma unroll
        for (int pt = 0; pt < 2; ++pt)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[pt][r] = first[pt][r] + acc[pt][r];
    }
    if constexpr (KS == 2) {
        // group g keeps rows h = g of the lane's tile; it hands the other half of its
        // partial sums to the partner group (lane-contiguous, conflict-free) and adds
        // the partner's: y = range 0 + range 1 in both groups
        float* xch = reinterpret_cast<float*>(sm);
        // (fixed register indices in both branches: no dynamic indexing of acc)
        if (ks == 0) {
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
#pragma unroll
                for (int j = 0; j < 8; ++j) xch[((16 + pt * 8 + j) * 256) + tid] = acc[pt][8 + j];
        } else {
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
#pragma unroll
                for (int j = 0; j < 8; ++j) xch[((pt * 8 + j) * 256) + tid] = acc[pt][j];
        }
        __syncthreads();
        if (ks == 0) {
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[pt][j] += xch[((pt * 8 + j) * 256) + tid];
        } else {
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[pt][8 + j] += xch[((16 + pt * 8 + j) * 256) + tid];
        }
    }
    const int col = n0 + (lane & 31);
    const double amp = LSTEP ? lstep_amp<PolMix>(col, la) : 1.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (KS == 2 && h != ks) continue;
        size_t o[8];
        float yr[8], yi[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int r = 8 * h + j;
            const int row = m0 + 32 * w + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            o[j] = (size_t)row * NC + col;
            yr[j] = acc[0][r];
            yi[j] = acc[1][r];
        }
        if constexpr (LSTEP) {
            lstep_batch<PolMix, 8>(o, yr, yi, amp, la);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) Y[o[j]] = C2<float>{yr[j], yi[j]};
        }
    }
}
template <bool LSTEP, int KS>
__global__ __launch_bounds__(256 * KS) void k_qdht_bf16x3_t32(const uint16_t* __restrict__ Afr, const C2<float>* __restrict__ X,
                                                              C2<float>* __restrict__ Y, int K, int NC, LstepArgs la, QCols qc)
{
    __shared__ __align__(16) uint16_t sm[smem_t32<2, KS>()];
    qdht_t32_body<LSTEP, 2, KS>(blockIdx.x, Afr, X, Y, K, NC, la, qc, sm);
}
template <bool LSTEP, int KS>
__global__ __launch_bounds__(256 * KS) void k_qdht_bf16x6_t32(const uint16_t* __restrict__ Afr, const C2<float>* __restrict__ X,
                                                              C2<float>* __restrict__ Y, int K, int NC, LstepArgs la, QCols qc)
{
    __shared__ __align__(16) uint16_t sm[smem_t32<3, KS>()];
    qdht_t32_body<LSTEP, 3, KS>(blockIdx.x, Afr, X, Y, K, NC, la, qc, sm);
}

// FP64: tile 64 rows x 32 complex columns, K step 16, 4 waves (2x2)
template <bool LSTEP>
__device__ __forceinline__ void qdht_f64_body(int bid, const double* __restrict__ At, const C2<double>* __restrict__ X,
                                              C2<double>* __restrict__ Y, int K, int NC, LstepArgs la, const QCols& qc)
{
    constexpr int BM = 64, BNC = 32, BK = 16, PA = 8, PB = 8;
    const DevCtl* ctl = la.ctl;
    if (ctl) {
        if (ctl->state != 0 || !ctl->accept) return;
        la.dz = ctl->dz_lin;
        la.seed = (uint32_t)ctl->steps;
    }
    __shared__ double As[BK][BM + PA];
    __shared__ double Br[BK][BNC + PB];
    __shared__ double Bi[BK][BNC + PB];
    const int M = K;
    const int mt = M / BM, nct = qc.ctile ? qc.nct : NC / BNC;
    if (bid >= mt * nct) {
        if constexpr (!LSTEP) zero_skipped(Y, NC, mt, BM, bid - mt * nct, qc, ctl);
        return;
    }
    int mi, ni;
    tile_of(bid, mt, nct, &mi, &ni, qc.order != 0);
    const int m0 = mi * BM, n0 = col0_of(qc, ni, BNC);
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
    const int col = n0 + wn + (lane & 15);
    const double amp = LSTEP ? lstep_amp<PolF64>(col, la) : 1.0;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // v_mfma_f64_16x16x4_f64 result layout: row = lane/16 + 4 r, column = lane%16
            const int row = m0 + wm + rb * 16 + (lane >> 4) + 4 * r;
            const size_t o = (size_t)row * NC + col;
            if (LSTEP) lstep_point<PolF64>(o, acc[rb][0][r], acc[rb][1][r], amp, la);
            else Y[o] = C2<double>{acc[rb][0][r], acc[rb][1][r]};
        }
}
template <bool LSTEP>
__global__ __launch_bounds__(256) void k_qdht_f64(const double* __restrict__ At, const C2<double>* __restrict__ X,
                                                  C2<double>* __restrict__ Y, int K, int NC, LstepArgs la, QCols qc)
{
    qdht_f64_body<LSTEP>(blockIdx.x, At, X, Y, K, NC, la, qc);
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

// propagator.cu:139-160 on the device: reject if dz^2 maxNL2 > tol^2 maxA2 (the
// attempt counts in STEP_N and n_reject), dz_new = 0.5 tol sqrt(maxA2 / maxNL2),
// collapse below dz_floor, clamp to the plane; the NaN / blast status of the
// attempt ends the run as the reference's throws do
__device__ __forceinline__ void controller_body(int nr, const double* __restrict__ lm, DevStatus* status, DevCtl* ctl)
{
    __shared__ double red[64];
    if (ctl->state != 0) return;
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < nr; i += blockDim.x) {
        a = fmax(a, lm[2 * i]);
        b = fmax(b, lm[2 * i + 1]);
    }
    const double A = block_max<1024>(a, red);
    const double B = block_max<1024>(b, red + 32);
    if (threadIdx.x != 0) return;
    status->maxA2 = A;
    status->maxNL2 = B;
    if (status->flags & kFlagNaN) { ctl->state = 2; ctl->err = 2; return; }
    if (status->flags & kFlagBlast) { ctl->state = 2; ctl->err = 3; return; }
    const double dz = ctl->dz, tol = ctl->tol;
    const bool repeat = dz * dz * B > tol * tol * A;
    if (repeat) {
        ctl->steps += 1;
        ctl->rejects += 1;
    }
    double dzn = 0.5 * tol * sqrt(A / B);
    if (dzn < ctl->dz_floor || !(dzn > 0)) {
        ctl->dz = dzn;
        ctl->state = 2;
        ctl->err = 1;
        return;
    }
    dzn = fmin(dzn, ctl->z_plane - ctl->z);
    ctl->dz = dzn;
    ctl->dz_lin = dzn;
    ctl->accept = repeat ? 0 : 1;
}
__global__ void k_controller(int nr, const double* __restrict__ lm, DevStatus* status, DevCtl* ctl)
{
    controller_body(nr, lm, status, ctl);
}

// after the linear half: CURRENT_Z += ZSTEP, STEP_N++ (solver_main.cpp:217-220);
// stop at the plane
__device__ __forceinline__ void finish_body(DevCtl* ctl)
{
    if (ctl->state != 0 || !ctl->accept) return;
    ctl->dead_dirty = 0;     // the accepted attempt's QDHT back has cleared the skipped columns
    ctl->z += ctl->dz_lin;
    ctl->steps += 1;
    ctl->accepted += 1;
    if (ctl->z >= ctl->z_plane) ctl->state = 1;
}
__global__ void k_finish(DevCtl* ctl) { finish_body(ctl); }

// ------------------------------------------------------------------ batched march
// Phases of the step attempts of several runs in one launch each (gs_march_many):
// block b works for run brun[b] as that run's block b - b0; each run's block
// range starts on a multiple of 8, so tile_of's XCD grouping holds within a
// run, and blocks past a run's count (padding) return at once.  The bodies
// are those of the single-run kernels, so every run's results are bit for bit
// those of its own march.
template <bool LSTEP, int KS>
__global__ __launch_bounds__(256 * KS) void k_qdht_t32_batch(const GemmRun* __restrict__ runs, const int* __restrict__ brun)
{
    constexpr int SM = smem_t32<3, KS>() > smem_t32<2, KS>() ? smem_t32<3, KS>() : smem_t32<2, KS>();
    __shared__ __align__(16) uint16_t sm[SM];
    const GemmRun& g = runs[brun[blockIdx.x]];
    const int bid = (int)blockIdx.x - g.b0;
    if (bid >= g.nb) return;
    if (g.qc.split3) qdht_t32_body<LSTEP, 3, KS>(bid, (const uint16_t*)g.A, (const C2<float>*)g.X, (C2<float>*)g.Y, g.K, g.NC, g.la, g.qc, sm);
    else qdht_t32_body<LSTEP, 2, KS>(bid, (const uint16_t*)g.A, (const C2<float>*)g.X, (C2<float>*)g.Y, g.K, g.NC, g.la, g.qc, sm);
}
template <bool LSTEP, int KS>
__global__ __launch_bounds__(256 * KS) void k_qdht_f32_batch(const GemmRun* __restrict__ runs, const int* __restrict__ brun)
{
    __shared__ __align__(16) float sm[smem_f32<32, KS>()];
    const GemmRun& g = runs[brun[blockIdx.x]];
    const int bid = (int)blockIdx.x - g.b0;
    if (bid >= g.nb) return;
    qdht_f32_body<LSTEP, 32, KS>(bid, (const float*)g.A, (const C2<float>*)g.X, (C2<float>*)g.Y, g.K, g.NC, g.la, g.qc, sm);
}
template <bool LSTEP>
__global__ __launch_bounds__(256) void k_qdht_f64_batch(const GemmRun* __restrict__ runs, const int* __restrict__ brun)
{
    const GemmRun& g = runs[brun[blockIdx.x]];
    const int bid = (int)blockIdx.x - g.b0;
    if (bid >= g.nb) return;
    qdht_f64_body<LSTEP>(bid, (const double*)g.A, (const C2<double>*)g.X, (C2<double>*)g.Y, g.K, g.NC, g.la, g.qc);
}
__global__ void k_controller_batch(const CtlRun* __restrict__ runs)
{
    const CtlRun& r = runs[blockIdx.x];
    controller_body(r.nr, r.linemax, r.status, r.ctl);
}
__global__ void k_finish_batch(DevCtl* ctls, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) finish_body(ctls + i);
}

int qdht_blocks(Prec p, int rp, int nt, const QCols* cols, bool back, bool gated)
{
    const int bm = p == kFP64 ? 64 : 128, mt = rp / bm;
    if (!cols || !cols->ctile) return mt * (nt / 32);
    return mt * cols->nct + ((back && (cols->zero_dead || gated)) ? mt * ((cols->nzt + 7) / 8) : 0);
}

// K groups of the narrow split-BF16 tiles: the caller's choice (the plan's
// autotuned one), unless GS_QDHT_KSPLIT (1 or 2) fixes it; both give the same bits
static int t32_groups(int want)
{
    static const int env = [] {
        const char* e = getenv("GS_QDHT_KSPLIT");
        return e ? (atoi(e) == 1 ? 1 : 2) : 0;
    }();
    return env ? env : (want == 1 ? 1 : 2);
}

int launch_qdht_batch(Prec p, bool lstep, const GemmRun* runs, const int* brun, int nblocks, hipStream_t st, int ks)
{
    const dim3 grid(nblocks), blk(256);
    if (p == kFP64) {
        if (lstep) hipLaunchKernelGGL(k_qdht_f64_batch<true>, grid, blk, 0, st, runs, brun);
        else hipLaunchKernelGGL(k_qdht_f64_batch<false>, grid, blk, 0, st, runs, brun);
    } else if (p == kFP32) {
        if (ks == 1) {
            if (lstep) hipLaunchKernelGGL((k_qdht_f32_batch<true, 1>), grid, blk, 0, st, runs, brun);
            else hipLaunchKernelGGL((k_qdht_f32_batch<false, 1>), grid, blk, 0, st, runs, brun);
        } else {
            const dim3 b2(512);
            if (lstep) hipLaunchKernelGGL((k_qdht_f32_batch<true, 2>), grid, b2, 0, st, runs, brun);
            else hipLaunchKernelGGL((k_qdht_f32_batch<false, 2>), grid, b2, 0, st, runs, brun);
        }
    } else if (t32_groups(ks) == 2) {
        const dim3 b2(512);
        if (lstep) hipLaunchKernelGGL((k_qdht_t32_batch<true, 2>), grid, b2, 0, st, runs, brun);
        else hipLaunchKernelGGL((k_qdht_t32_batch<false, 2>), grid, b2, 0, st, runs, brun);
    } else {
        if (lstep) hipLaunchKernelGGL((k_qdht_t32_batch<true, 1>), grid, blk, 0, st, runs, brun);
        else hipLaunchKernelGGL((k_qdht_t32_batch<false, 1>), grid, blk, 0, st, runs, brun);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int launch_controller_batch(const CtlRun* runs, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_controller_batch, dim3(n), dim3(1024), 0, st, runs);
    HIPCHK(hipGetLastError());
    return 0;
}

int launch_finish_batch(DevCtl* ctls, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_finish_batch, dim3((n + 63) / 64), dim3(64), 0, st, ctls, n);
    HIPCHK(hipGetLastError());
    return 0;
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

// the full BIGBUFFER2 from FIELD_k (unmasked points) and the masked points' values
template <class R>
__global__ void k_b2_merge(const C2<R>* __restrict__ fk, const C2<R>* __restrict__ b2, C2<R>* __restrict__ out,
                           const R* __restrict__ g, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = g[i] < (R)0 ? b2[i] : fk[i];
}
// FIELD_k (the live B2) from a full BIGBUFFER2: zero on the masked points
template <class R>
__global__ void k_b2_split(const C2<R>* __restrict__ b2, C2<R>* __restrict__ fk, const R* __restrict__ g, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) fk[i] = g[i] < (R)0 ? C2<R>{0, 0} : b2[i];
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

int launch_qdht(Prec p, int rp, int nt, const void* amat, const void* in, void* out,
                const LstepArgs* ls, hipStream_t st, const DevCtl* gate, const QCols* cols)
{
    LstepArgs la = ls ? *ls : LstepArgs{};
    if (!ls) la.ctl = gate;
    QCols qc = cols ? *cols : QCols{};
    // blocks of a column-skipping QDHT back that clear the skipped columns: present when the
    // host asks for it, or always in the device-driven march (they read DevCtl::dead_dirty)
    // cols with no column list: every column, through the narrow tiles (for split3)
    auto zblocks = [&](int mt) {
        if (!cols || !qc.ctile || ls || !(qc.zero_dead || gate)) return 0;
        return mt * ((qc.nzt + 7) / 8);
    };
    if (cols && (nt % 32 || (qc.ctile && qc.nct < 1))) return -1001;
    const int ncb = cols && qc.ctile ? qc.nct : nt / 32;     // 32-column blocks computed
    if (p == kFP64) {
        if (rp % 64 || nt % 32) return -1001;
        const int mt = rp / 64;
        const dim3 grid(mt * ncb + zblocks(mt));
        if (ls) hipLaunchKernelGGL(k_qdht_f64<true>, grid, dim3(256), 0, st, (const double*)amat, (const C2<double>*)in, (C2<double>*)out, rp, nt, la, qc);
        else hipLaunchKernelGGL(k_qdht_f64<false>, grid, dim3(256), 0, st, (const double*)amat, (const C2<double>*)in, (C2<double>*)out, rp, nt, la, qc);
    } else if (p == kFP32) {
        if (rp % 128 || nt % 64) return -1001;
        const int mt = rp / 128;
        if (cols && (qc.ctile || !qc.wide)) {
            const dim3 grid(mt * ncb + zblocks(mt));
            const bool two = qc.ks != 1;
            auto kt = two ? k_qdht_f32<true, 32, 2> : k_qdht_f32<true, 32, 1>;
            auto kf = two ? k_qdht_f32<false, 32, 2> : k_qdht_f32<false, 32, 1>;
            hipLaunchKernelGGL(ls ? kt : kf, grid, dim3(two ? 512 : 256), 0, st, (const float*)amat, (const C2<float>*)in, (C2<float>*)out, rp, nt, la, qc);
        } else {
            const dim3 grid(mt * (nt / 64));
            const bool two = !cols || qc.ks != 1;
            auto kt = two ? k_qdht_f32<true, 64, 2> : k_qdht_f32<true, 64, 1>;
            auto kf = two ? k_qdht_f32<false, 64, 2> : k_qdht_f32<false, 64, 1>;
            hipLaunchKernelGGL(ls ? kt : kf, grid, dim3(two ? 512 : 256), 0, st, (const float*)amat, (const C2<float>*)in, (C2<float>*)out, rp, nt, la, qc);
        }
    } else {
        if (rp % 128 || nt % 128) return -1001;
        const int mt = rp / 128;
        if (cols) {
            const dim3 grid(mt * ncb + zblocks(mt));
            const int ks = t32_groups(qc.ks);
            auto k3t = ks == 2 ? k_qdht_bf16x3_t32<true, 2> : k_qdht_bf16x3_t32<true, 1>;
            auto k3f = ks == 2 ? k_qdht_bf16x3_t32<false, 2> : k_qdht_bf16x3_t32<false, 1>;
            auto k6t = ks == 2 ? k_qdht_bf16x6_t32<true, 2> : k_qdht_bf16x6_t32<true, 1>;
            auto k6f = ks == 2 ? k_qdht_bf16x6_t32<false, 2> : k_qdht_bf16x6_t32<false, 1>;
            auto kf = qc.split3 ? (ls ? k6t : k6f) : (ls ? k3t : k3f);
            hipLaunchKernelGGL(kf, grid, dim3(256 * ks), 0, st, (const uint16_t*)amat, (const C2<float>*)in, (C2<float>*)out, rp, nt, la, qc);
        } else {
            const dim3 grid(mt * (nt / 128));
            auto kt = k_qdht_bf16x3_t128<true>;
            auto kf = k_qdht_bf16x3_t128<false>;
            HIPCHK(hipFuncSetAttribute((const void*)kt, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kBf16T128Smem));
            HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kBf16T128Smem));
            if (ls) hipLaunchKernelGGL(kt, grid, dim3(256), kBf16T128Smem, st, (const uint16_t*)amat, (const C2<float>*)in, (C2<float>*)out, rp, nt, la);
            else hipLaunchKernelGGL(kf, grid, dim3(256), kBf16T128Smem, st, (const uint16_t*)amat, (const C2<float>*)in, (C2<float>*)out, rp, nt, la);
        }
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

int launch_controller(int nr, const double* linemax, DevStatus* status, DevCtl* ctl, hipStream_t st)
{
    hipLaunchKernelGGL(k_controller, dim3(1), dim3(1024), 0, st, nr, linemax, status, ctl);
    HIPCHK(hipGetLastError());
    return 0;
}

int launch_finish(DevCtl* ctl, hipStream_t st)
{
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(1), 0, st, ctl);
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

int launch_b2_merge(Prec p, const void* fk, const void* b2, void* out, const void* gfac, size_t n, hipStream_t st)
{
    const dim3 grid((unsigned)((n + 255) / 256));
    if (p == kFP64)
        hipLaunchKernelGGL(k_b2_merge<double>, grid, dim3(256), 0, st, (const C2<double>*)fk, (const C2<double>*)b2, (C2<double>*)out, (const double*)gfac, n);
    else
        hipLaunchKernelGGL(k_b2_merge<float>, grid, dim3(256), 0, st, (const C2<float>*)fk, (const C2<float>*)b2, (C2<float>*)out, (const float*)gfac, n);
    HIPCHK(hipGetLastError());
    return 0;
}

int launch_b2_split(Prec p, const void* b2, void* fk, const void* gfac, size_t n, hipStream_t st)
{
    const dim3 grid((unsigned)((n + 255) / 256));
    if (p == kFP64)
        hipLaunchKernelGGL(k_b2_split<double>, grid, dim3(256), 0, st, (const C2<double>*)b2, (C2<double>*)fk, (const double*)gfac, n);
    else
        hipLaunchKernelGGL(k_b2_split<float>, grid, dim3(256), 0, st, (const C2<float>*)b2, (C2<float>*)fk, (const float*)gfac, n);
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
