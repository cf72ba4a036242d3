// gs_plan.cpp -- the plan object and the C-ABI of include/gsolver_amd.h.
//
// The host side replays the reference's control flow: process_input
// (set-up), the propagator() accept/reject controller (propagator.cu:125-266),
// the main loop with dump planes (solver_main.cpp:202-225) and the .udf
// writers (dump_types.cpp:25-72); every per-step array operation runs on the
// GPU through gs_device.hip.  There is no CPU fallback: without a device the
// plan cannot be created.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/gsolver_amd.h"
#include "gs_device.h"
#include "gs_host.h"

using gs::cd;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& m)
{
    g_err = m;
    return code;
}

#define HCK(x)                                                                                    \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) throw gs::Error(GS_E_HIP, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)
#define LCK(x)                                                                                    \
    do {                                                                                          \
        int e_ = (x);                                                                             \
        if (e_ != 0) throw gs::Error(GS_E_HIP, "kernel launch failed (" + std::to_string(e_) + ") at " #x); \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    void alloc(size_t bytes)
    {
        if (p) HCK(hipFree(p));
        p = nullptr;
        n = bytes;
        if (bytes) HCK(hipMalloc(&p, bytes));
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    template <class T> T* as() const { return (T*)p; }
};

void upload(DevBuf& b, const void* src, size_t bytes)
{
    b.alloc(bytes);
    HCK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
}

uint16_t bf16_rn_host(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
float bf16_to_float(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

struct gs_plan {
    gs::Setup s;
    gs::Prec prec = gs::kFP32;
    int dev = 0;             // the device current at plan creation (the plan's memory and stream)
    int rp = 0;              // padded rows
    size_t relem = 0;        // bytes per complex element in the field precision
    hipStream_t st = nullptr;
    hipEvent_t ev[5] = {};

    DevBuf field, b1, b2, fk, tmp;       // [rp][T] in field precision; fk = FIELD_k = B2 on the
                                         // unmasked points (the live state), b2 = B2's masked points
    DevBuf amat, at64;                   // QDHT operators
    DevBuf tw, tw64, zpow, pf, kerr, ionrate, k2, lrate, lgfac, avgw, dft16, mfma_tw;
    DevBuf linemax, status, diag, et, rhot, scratch64a, scratch64b;
    std::vector<DevBuf> wfilt;
    gs::NLArgs nl{};
    gs::LstepArgs ls{};

    // column skipping of the march's two QDHTs: the 32-column blocks that hold a
    // live omega column (ctile) and the others (ztile); FIELD keeps values in the
    // skipped blocks until a QDHT back clears them (dead_dirty)
    DevBuf ctile, ztile;
    int nct = 0, nzt = 0;
    bool qskip = true;
    bool dead_dirty = true;
    // mixed: which of the march's transforms split their operands in three bf16 parts
    // (GS_MIX_QDHT x6back, the default: the QDHT back; x6: both; x3: neither)
    bool split3_fwd = false, split3_back = false;
    // the column selection of one of the march's QDHTs: the live blocks, or every
    // column when skipping is off (through the narrow tiles when split3 asks for them;
    // nullptr: the plain full-width kernels)
    gs::QCols qcols(bool zero_dead, bool back) const
    {
        const int s3 = (back ? split3_back : split3_fwd) ? 1 : 0;
        if (!qskip)
            return gs::QCols{nullptr, 0, nullptr, 0, 0, s3, qks[1], qorder[1], prec == gs::kFP32 && !f32_narrow_full ? 1 : 0};
        return gs::QCols{ctile.as<int>(), nct, ztile.as<int>(), nzt, zero_dead ? 1 : 0, s3, qks[0], qorder[0], 0};
    }
    // GEMM tile shapes (autotune_qdht; every choice gives the same bits), for the
    // column-skipping [0] and the all-column [1] march: K groups of the mixed
    // narrow tiles, the tile order over the XCDs, and whether FP32 all-column
    // transforms use the narrow 32-column tiles
    int qks[2] = {2, 2};
    int qorder[2] = {0, 0};
    bool f32_narrow_full = false;
    // the march's transforms go through the column-list kernels when columns are
    // skipped, and always in mixed precision (its narrow tiles sum in their own
    // order, so the all-column march matches the skipping one bit for bit)
    // the march's transforms always take the plan's column list and tile shape
    bool march_cols(bool) const { return true; }

    double ZSTEP = 0, CURRENT_Z = 0;
    long STEP_N = 0, n_reject = 0;
    int n_Z = 0;
    bool dumps_started = false;   // a march call of this plan has written a dump plane
    double timings[4] = {0, 0, 0, 0};

    // device-driven march: controller state, its pinned host mirror, and the
    // captured graph of one step attempt
    DevBuf ctl;
    gs::DevCtl* ctl_host = nullptr;
    hipGraphExec_t slot_exec = nullptr;

    // reverse mode: FIELD before each step of the segment being reversed
    // (ckpt), FIELD and FIELD_k (the live BIGBUFFER2) at the start of every
    // segment (snaps), the
    // replay buffers that recompute a segment, the step sizes, the QDHT
    // transpose (row-major A, so the transposed-operand GEMM computes A^T x)
    // and the cotangent buffers.  A march whose per-step checkpoints fit in
    // adj_budget bytes is one segment (no recomputation).
    DevBuf ckpt, snaps, rfield, rb1, rfk, adjmat, gA, gB, gT1, gT2, stage;
    std::vector<double> adj_rk, adj_lin;
    int adj_seg = 0;                       // steps per segment
    double adj_budget = -1;                // bytes of per-step checkpoints kept without recomputation
                                           // (< 0: 60 % of the device memory free at the forward)
    long adj_seed0 = 0;                    // STEP_N at the first recorded step (rounding seeds of the replay)
    bool adj_dirty0 = false;               // FIELD's skipped columns were non-zero at the first recorded step
    double batch_nodes = 0;                // kernels per step attempt of the last gs_march_many over this plan
    double last_dz_nl = 0;     // NL step size of the last accepted gs_step attempt

    void* buf(int id)
    {
        switch (id) {
        case GS_FIELD: return field.p;
        case GS_B1: return b1.p;
        case GS_B2: return b2.p;
        case GS_FK: return fk.p;
        }
        throw gs::Error(GS_E_ARG, "unknown buffer id");
    }
    size_t field_bytes() const { return (size_t)rp * s.NT * relem; }
};

namespace {

void sync(gs_plan* p) { HCK(hipStreamSynchronize(p->st)); }

void check_status(gs_plan* p, gs::DevStatus* out)
{
    gs::DevStatus h;
    HCK(hipMemcpyAsync(&h, p->status.p, sizeof h, hipMemcpyDeviceToHost, p->st));
    sync(p);
    if (h.flags & gs::kFlagNaN) {
        HCK(hipMemsetAsync(p->status.p, 0, sizeof h, p->st));
        throw gs::Error(GS_E_PHYSICS, "float_type photoionization_function(...): Intensity is nan!");
    }
    if (h.flags & gs::kFlagBlast) {
        HCK(hipMemsetAsync(p->status.p, 0, sizeof h, p->st));
        throw gs::Error(GS_E_PHYSICS, "Attention!!!! Intensity is too big! Keep away, blast is possible.");
    }
    if (out) *out = h;
}

// batched FP64 line FFT on device scratch (init-time transforms)
void fft64_lines(gs_plan* p, const double* host_in, double* host_out, int nlines, int sign, double scale)
{
    const size_t bytes = (size_t)nlines * p->s.NT * 16;
    if (p->scratch64a.n < bytes) p->scratch64a.alloc(bytes);
    if (p->scratch64b.n < bytes) p->scratch64b.alloc(bytes);
    HCK(hipMemcpyAsync(p->scratch64a.p, host_in, bytes, hipMemcpyHostToDevice, p->st));
    LCK(gs::launch_fft_lines(gs::kFP64, p->s.NT, p->scratch64a.p, p->scratch64b.p, nlines, sign, scale, nullptr, p->tw64.p, p->st));
    HCK(hipMemcpyAsync(host_out, p->scratch64b.p, bytes, hipMemcpyDeviceToHost, p->st));
    sync(p);
}

void upload_field64(gs_plan* p, DevBuf& dst, const double* host)   // host [R][T] cplx -> device [rp][T]
{
    const size_t n = (size_t)p->s.NR * p->s.NT;
    if (p->scratch64a.n < n * 16) p->scratch64a.alloc((size_t)p->rp * p->s.NT * 16);
    HCK(hipMemcpyAsync(p->scratch64a.p, host, n * 16, hipMemcpyHostToDevice, p->st));
    HCK(hipMemsetAsync(dst.p, 0, p->field_bytes(), p->st));
    LCK(gs::launch_convert(p->prec, p->scratch64a.as<double>(), dst.p, n, 1, p->st));
    sync(p);
}

void download_field64(gs_plan* p, const void* src, double* host, size_t nlines)
{
    const size_t n = nlines * p->s.NT;
    if (p->scratch64a.n < n * 16) p->scratch64a.alloc((size_t)p->rp * p->s.NT * 16);
    LCK(gs::launch_convert(p->prec, p->scratch64a.as<double>(), const_cast<void*>(src), n, 0, p->st));
    HCK(hipMemcpyAsync(host, p->scratch64a.p, n * 16, hipMemcpyDeviceToHost, p->st));
    sync(p);
}

// split-BF16 operator M (row-major [m][k], rp x rp) of the mixed GEMM: hi = bf16(a),
// lo = bf16(a - hi), in the MFMA fragment order of the GEMM: [m/128][k/32][k slice
// (k%32)/16][32-row block][hi, lo][lane][8], lane = m%32 + 32 ((k%16)/8), element
// k%8 (v_mfma_f32_32x32x16_bf16 A operand)
// A third plane of the same fragment order without the hi/lo interleave, at
// 2 rp^2, holds the rest bf16(a - hi - lo) for the three-part split of the
// column-skipping GEMM.
void upload_bf16_fragments(DevBuf& dst, const std::vector<double>& M, int rp)
{
    std::vector<uint16_t> fr((size_t)3 * rp * rp);
    const int nk = rp / 32;
    for (int m = 0; m < rp; ++m)
        for (int k = 0; k < rp; ++k) {
            const double ad = M[(size_t)m * rp + k];
            const float a = (float)ad;
            const uint16_t h = bf16_rn_host(a), l = bf16_rn_host(a - bf16_to_float(h));
            const uint16_t r = bf16_rn_host((float)(ad - bf16_to_float(h) - bf16_to_float(l)));
            const int mt = m / 128, rbb = (m % 128) / 32, step = k / 32, kk = (k % 32) / 16;
            const int lane = m % 32 + 32 * ((k % 16) / 8), e = k % 8;
            const size_t base = (((((size_t)mt * nk + step) * 2 + kk) * 4 + rbb) * 2) * 64 * 8;
            fr[base + lane * 8 + e] = h;
            fr[base + 64 * 8 + lane * 8 + e] = l;
            fr[(size_t)2 * rp * rp + base / 2 + lane * 8 + e] = r;
        }
    upload(dst, fr.data(), fr.size() * 2);
}

// The forward radix-16 DFT Y_q = sum_r x_r e^{-2 pi i q r/16} as the real 32 x 32
// A operand of v_mfma_f32_32x32x16_bf16 (rows Re Y_0..15, Im Y_0..15; k slice 0
// takes Re x_r, slice 1 Im x_r), split into bf16 hi = bf16(a), lo = bf16(a - hi),
// in fragment order [k slice][hi, lo][lane][8]: lane = m + 32 (r / 8), element r % 8
void upload_dft16_fragments(DevBuf& dst)
{
    std::vector<uint16_t> fr(2 * 2 * 2 * 64 * 8);
    for (int sg = 0; sg < 2; ++sg)          // sign -1 (forward), +1 (inverse)
        for (int sl = 0; sl < 2; ++sl)
            for (int lane = 0; lane < 64; ++lane)
                for (int e = 0; e < 8; ++e) {
                    const int m = lane & 31, r = 8 * (lane >> 5) + e, q = m & 15;
                    const double ang = 2 * M_PI * ((q * r) & 15) / 16.0;
                    const double c = std::cos(ang), sn = (sg ? 1.0 : -1.0) * std::sin(ang);   // e^{+-i ang} = c + i sn
                    // rows < 16: Re Y = C Re x - Sn Im x; rows >= 16: Im Y = Sn Re x + C Im x
                    const double a = m < 16 ? (sl == 0 ? c : -sn) : (sl == 0 ? sn : c);
                    const uint16_t h = bf16_rn_host((float)a), l = bf16_rn_host((float)a - bf16_to_float(h));
                    fr[(((sg * 2 + sl) * 2 + 0) * 64 + lane) * 8 + e] = h;
                    fr[(((sg * 2 + sl) * 2 + 1) * 64 + lane) * 8 + e] = l;
                }
    upload(dst, fr.data(), fr.size() * 2);
}

void build_plan(gs_plan* p)
{
    const gs::Setup& s = p->s;
    const int nt = s.NT, nr = s.NR;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        throw gs::Error(GS_E_NODEVICE, "no GPU visible: gsolver_amd has no CPU fallback");
    if (nt < 256 || nt > 4096 || (nt & (nt - 1)))
        throw gs::Error(GS_E_ARG, "N_T must be a power of two between 256 and 4096 on the GPU path");
    p->rp = (nr + 127) / 128 * 128;
    const int rp = p->rp;
    const bool dbl = p->prec == gs::kFP64;
    p->relem = dbl ? 16 : 8;
    HCK(hipGetDevice(&p->dev));
    {
        // address of this device's fixed scratch of gs_nl_response_noalloc, looked up once per
        // device; a failure is reported by that entry point (which looks it up again), not here
        void* scr = nullptr;
        (void)gs::noalloc_scratch(&scr);
    }
    HCK(hipStreamCreate(&p->st));
    for (auto& e : p->ev) HCK(hipEventCreate(&e));

    // ---- operators: A = diag(m1) C diag(1/m1), padded; stored transposed (At[k][m] = A[m][k])
    std::vector<double> A((size_t)rp * rp, 0.0);
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < nr; ++j) A[(size_t)i * rp + j] = s.qm1[i] * s.qC[(size_t)i * nr + j] / s.qm1[j];
    std::vector<double> At((size_t)rp * rp);
    for (int i = 0; i < rp; ++i)
        for (int j = 0; j < rp; ++j) At[(size_t)j * rp + i] = A[(size_t)i * rp + j];
    upload(p->at64, At.data(), At.size() * 8);
    if (p->prec == gs::kFP64) {
        p->amat.alloc(0);
    } else if (p->prec == gs::kFP32) {
        std::vector<float> f(At.begin(), At.end());
        upload(p->amat, f.data(), f.size() * 4);
    } else {
        upload_bf16_fragments(p->amat, A, rp);
    }

    // ---- tables
    std::vector<double> tw64(2 * nt);
    for (int k = 0; k < nt; ++k) {
        // e^{-2 pi i k / T}, evaluated by octant symmetry for accuracy
        const double ang = 2.0 * M_PI * k / nt;
        tw64[2 * k] = std::cos(ang);
        tw64[2 * k + 1] = -std::sin(ang);
        if (4 * k == nt) { tw64[2 * k] = 0; tw64[2 * k + 1] = -1; }
        if (2 * k == nt) { tw64[2 * k] = -1; tw64[2 * k + 1] = 0; }
        if (4 * k == 3 * nt) { tw64[2 * k] = 0; tw64[2 * k + 1] = 1; }
    }
    upload(p->tw64, tw64.data(), tw64.size() * 8);
    std::vector<double> zp(2 * (nt + 1)), pf(2 * nt), kerr(s.kerr);
    for (int k = 0; k <= nt; ++k) {
        const cd z = std::pow(s.raman_z, (double)k);
        zp[2 * k] = z.real();
        zp[2 * k + 1] = z.imag();
    }
    for (int k = 0; k < nt; ++k) {
        pf[2 * k] = s.plasma_func[k].real();
        pf[2 * k + 1] = s.plasma_func[k].imag();
    }
    if (dbl) {
        upload(p->tw, tw64.data(), tw64.size() * 8);
        upload(p->zpow, zp.data(), zp.size() * 8);
        upload(p->pf, pf.data(), pf.size() * 8);
        upload(p->kerr, kerr.data(), kerr.size() * 8);
    } else {
        std::vector<float> a(tw64.begin(), tw64.end()), b(zp.begin(), zp.end()), c(pf.begin(), pf.end()), d(kerr.begin(), kerr.end());
        upload(p->tw, a.data(), a.size() * 4);
        upload(p->zpow, b.data(), b.size() * 4);
        upload(p->pf, c.data(), c.size() * 4);
        upload(p->kerr, d.data(), d.size() * 4);
    }
    upload(p->ionrate, s.ionrate_ln.data(), s.ionrate_ln.size() * 8);
    {
        // linear-step tables (propagator.cu:570-625): rate = -HO - s k1, g = 1/(1+s),
        // s = sqrtHO(-kt2/k1^2); g = -1 where FIELD is zeroed (omega outside
        // [OMEGA_MIN, OMEGA_MAX], kt2 > MAX_KT2 Re(k^2)) and on padding rows
        std::vector<double> k2(nt), rate((size_t)rp * nt, 0.0), g((size_t)rp * nt, -1.0);
        for (int w = 0; w < nt; ++w) k2[w] = s.wavenum[w].imag();
        for (int i = 0; i < nr; ++i) {
            const double kt = 2 * M_PI * s.qx[i] * s.qV / s.XMAX, kt2 = kt * kt;
            for (int w = 0; w < nt; ++w) {
                if (s.omega[w] < s.OMEGA_MIN || s.omega[w] > s.OMEGA_MAX) continue;
                if (kt2 > 0.25 * (s.wavenum[w] * s.wavenum[w]).real()) continue;
                const double k1 = s.wavenum[w].real();
                const double sq = gs::sqrtho(-kt2 / k1 / k1);
                rate[(size_t)i * nt + w] = -s.ho[w] - sq * k1;
                g[(size_t)i * nt + w] = 1.0 / (1.0 + sq);
            }
        }
        upload(p->k2, k2.data(), k2.size() * 8);
        // 32-column blocks with at least one unmasked point: the columns the march's QDHTs compute
        std::vector<int> ct, zt;
        for (int c0 = 0; c0 < nt; c0 += 32) {
            bool live = false;
            for (int i = 0; i < nr && !live; ++i)
                for (int w = c0; w < c0 + 32 && !live; ++w) live = g[(size_t)i * nt + w] >= 0;
            (live ? ct : zt).push_back(c0);
        }
        p->nct = (int)ct.size();
        p->nzt = (int)zt.size();
        if (p->nct > 0) upload(p->ctile, ct.data(), ct.size() * sizeof(int));
        if (p->nzt > 0) upload(p->ztile, zt.data(), zt.size() * sizeof(int));
        p->qskip = p->nct > 0 && !std::getenv("GS_QDHT_ALL_COLUMNS");
        if (p->prec == gs::kMIXED) {
            const char* m = std::getenv("GS_MIX_QDHT");
            const std::string mode = m ? m : "x6back";
            if (mode != "x3" && mode != "x6back" && mode != "x6")
                throw gs::Error(GS_E_ARG, "GS_MIX_QDHT must be x3, x6back or x6");
            p->split3_back = mode != "x3";
            p->split3_fwd = mode == "x6";
        }
        if (p->prec == gs::kFP32) {
            std::vector<float> rf(rate.begin(), rate.end());
            upload(p->lrate, rf.data(), rf.size() * 4);
        } else {
            upload(p->lrate, rate.data(), rate.size() * 8);
        }
        if (p->prec == gs::kFP64) {
            upload(p->lgfac, g.data(), g.size() * 8);
        } else {
            std::vector<float> gf(g.begin(), g.end());
            upload(p->lgfac, gf.data(), gf.size() * 4);
        }
    }
    {
        // average_spectrum weights: x0 (x1 - x0), then x_r (x_r - x_{r-1})
        std::vector<double> wts(nr);
        for (int r = 0; r < nr; ++r) {
            const double x1 = r == 0 ? s.qx[0] : s.qx[r - 1];
            const double x2 = r == 0 ? s.qx[1] : s.qx[r];
            wts[r] = (r == 0 ? x1 : x2) * (x2 - x1);
        }
        upload(p->avgw, wts.data(), wts.size() * 8);
    }
    for (const auto& d : s.dumps) {
        p->wfilt.emplace_back();
        if (!d.wfilter.empty()) upload(p->wfilt.back(), d.wfilter.data(), d.wfilter.size() * 8);
    }

    // ---- state buffers
    for (DevBuf* b : {&p->field, &p->b1, &p->b2, &p->fk, &p->tmp}) {
        b->alloc(p->field_bytes());
        HCK(hipMemset(b->p, 0, p->field_bytes()));
    }
    p->linemax.alloc((size_t)rp * 2 * 8);
    p->status.alloc(sizeof(gs::DevStatus));
    HCK(hipMemset(p->status.p, 0, sizeof(gs::DevStatus)));
    p->diag.alloc((size_t)rp * 3 * 8);

    // ---- kernel argument blocks
    gs::NLArgs& a = p->nl;
    a.nr = nr;
    a.inv_nt = 1.0 / nt;
    a.dt = s.TSTEP * 1e20;
    a.neutral = s.neutral_density;
    a.inv_neutral = 1.0 / s.neutral_density;
    a.inv_recomb = 1.0 / s.recomb_tau;
    a.U = s.ion_potential;
    a.pond = s.ponderomotive;
    a.sigma = s.avalanche;
    a.rho_init = s.rho_init;
    a.ionrate = p->ionrate.as<double>();
    a.use_kerr = s.n2 != 0;
    a.use_raman = s.raman_fraction > 0.001;
    a.one_m_fr = 1 - s.raman_fraction;
    a.frM = s.raman_fraction * s.raman_M;
    const cd omzT = 1.0 - std::pow(s.raman_z, (double)nt);
    a.onemzT_re = omzT.real();
    a.onemzT_im = omzT.imag();
    a.zpow = p->zpow.p;
    a.pf = p->pf.p;
    a.kerr = p->kerr.p;
    a.tw = p->tw.p;
    if (p->prec == gs::kMIXED) {
        upload_dft16_fragments(p->dft16);
        // lane-order twiddles of the MFMA radix-16 passes at Ns = 16, 256, ... (while 16 Ns
        // divides N_T): [pass][group j][r] = e^{-2 pi i r (j mod Ns) / (16 Ns)}
        std::vector<float> t;
        for (int Ns = 16; nt % (16 * Ns) == 0; Ns *= 16)
            for (int j = 0; j < nt / 16; ++j)
                for (int r = 0; r < 16; ++r) {
                    const size_t m = (size_t)r * (j & (Ns - 1)) * (nt / (16 * Ns));
                    t.push_back((float)tw64[2 * m]);
                    t.push_back((float)tw64[2 * m + 1]);
                }
        if (!t.empty()) upload(p->mfma_tw, t.data(), t.size() * 4);
    }
    a.dft16 = p->dft16.p;
    a.mfma_tw = p->mfma_tw.p;
    // which mixed line FFTs run on the matrix cores (GS_MIX_FFT: valu, the default; kerr: the
    // Kerr/Raman forward transform; fwd: every forward transform; kerr_inv / all: those and
    // the inverse transform of non-ionizing lines).  VALU is the faster and the more
    // accurate here; the MFMA variants are kept for measurement.
    a.mfma_fwd = a.mfma_inv = 0;
    if (p->prec == gs::kMIXED) {
        const char* m = std::getenv("GS_MIX_FFT");
        const std::string mode = m ? m : "valu";
        if (mode != "valu" && mode != "kerr" && mode != "fwd" && mode != "kerr_inv" && mode != "all")
            throw gs::Error(GS_E_ARG, "GS_MIX_FFT must be valu, kerr, fwd, kerr_inv or all");
        a.mfma_fwd = mode == "valu" ? 0 : (mode == "fwd" || mode == "all") ? 2 : 1;
        a.mfma_inv = (mode == "kerr_inv" || mode == "all") ? 1 : 0;
    }
    a.status = p->status.as<gs::DevStatus>();
    // samples per thread of the RK4 line kernel (tuning knob; 0 = built-in default)
    a.nl_e = 0;
    if (const char* e = std::getenv("GS_NL_E")) a.nl_e = std::atoi(e);
    gs::LstepArgs& l = p->ls;
    l.k2 = p->k2.as<double>();
    l.rate = p->lrate.p;
    l.gfac = p->lgfac.p;
    l.fk = p->fk.p;

    // ---- initial field: FFT of E(r,t), spectral phase, /FIELD_DENOM (pulse_gg.cpp:100-121, process_input.cpp:58)
    const size_t n = (size_t)nr * nt;
    std::vector<double> f64(2 * n), sp(2 * n);
    std::memcpy(f64.data(), s.field_t.data(), n * 16);
    fft64_lines(p, f64.data(), sp.data(), nr, -1, 1.0);
    for (int r = 0; r < nr; ++r)
        for (int w = 0; w < nt; ++w) {
            const size_t o = (size_t)r * nt + w;
            const cd v = cd(sp[2 * o], sp[2 * o + 1]) * gs::init_spectral_factor(s, r, w) / 1e8;
            sp[2 * o] = v.real();
            sp[2 * o + 1] = v.imag();
        }
    // init_zstep_kerr (process_input.cpp:734-753): max |IFFT(FIELD)/T|^2
    fft64_lines(p, sp.data(), f64.data(), nr, +1, 1.0 / nt);
    double maxI = 0;
    for (size_t i = 0; i < n; ++i) maxI = std::max(maxI, f64[2 * i] * f64[2 * i] + f64[2 * i + 1] * f64[2 * i + 1]);
    p->ZSTEP = s.ZNET[1] - s.ZNET[0];
    const double kNL = s.OMEGA0 / 299792458.0 * s.n2 * maxI;
    if (p->ZSTEP * kNL > s.MAX_TOLERANCE) p->ZSTEP = s.MAX_TOLERANCE / kNL;
    p->CURRENT_Z = s.CURRENT_Z0;
    upload_field64(p, p->field, sp.data());
    // BIGBUFFER2 = QDHT(FIELD) in FP64
    {
        const size_t bytes = (size_t)rp * nt * 16;
        if (p->scratch64a.n < bytes) p->scratch64a.alloc(bytes);
        if (p->scratch64b.n < bytes) p->scratch64b.alloc(bytes);
        HCK(hipMemsetAsync(p->scratch64a.p, 0, bytes, p->st));
        HCK(hipMemcpyAsync(p->scratch64a.p, sp.data(), n * 16, hipMemcpyHostToDevice, p->st));
        LCK(gs::launch_qdht(gs::kFP64, rp, nt, p->at64.p, p->scratch64a.p, p->scratch64b.p, nullptr, p->st));
        LCK(gs::launch_convert(p->prec, p->scratch64b.as<double>(), p->b2.p, (size_t)rp * nt, 1, p->st));
        LCK(gs::launch_b2_split(p->prec, p->b2.p, p->fk.p, p->lgfac.p, (size_t)rp * nt, p->st));
        sync(p);
    }
    p->s.field_t.clear();
    p->s.field_t.shrink_to_fit();
}

// ---- autotuner of the line kernel's samples per thread
// For FP32 / mixed lines of N_T >= 1024 the RK4 kernel has two shapes (8 or 16
// samples per thread: 512 or 256 threads per line at N_T = 4096, different
// register and LDS footprints).  The first plan of a (precision, N_T, N_R) in
// the process times both on its initial field and keeps 16 when it beats 8 by
// 5 %; GS_NL_E overrides (4 is also accepted).
void autotune_nl(gs_plan* p)
{
    static std::mutex mu;
    static std::map<std::tuple<int, int, int>, int> cache;
    gs::NLArgs& a = p->nl;
    if (a.nl_e != 0 || p->prec == gs::kFP64 || p->s.NT < 1024) return;
    const auto key = std::make_tuple((int)p->prec, p->s.NT, p->s.NR);
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            a.nl_e = it->second;
            return;
        }
    }
    hipEvent_t e0, e1;
    HCK(hipEventCreate(&e0));
    HCK(hipEventCreate(&e1));
    // candidates alternate (warm-up round first), each keeps its best time; the
    // default shape (8) is replaced only by a clear (5 %) win
    const int cand[2] = {8, 16};
    float tmin[2] = {1e30f, 1e30f};
    for (int rep = 0; rep < 6; ++rep)
        for (int c = 0; c < 2; ++c) {
            a.nl_e = cand[c];
            HCK(hipEventRecord(e0, p->st));
            LCK(gs::launch_nl_rk4(p->prec, p->s.NT, p->field.p, p->b1.p, p->ZSTEP, a, p->linemax.as<double>(), p->st));
            HCK(hipEventRecord(e1, p->st));
            sync(p);
            float t = 0;
            HCK(hipEventElapsedTime(&t, e0, e1));
            if (rep > 0) tmin[c] = std::min(tmin[c], t);
        }
    int best = cand[0];
    float tb = 0.95f * tmin[0];
    for (int c = 1; c < 2; ++c)
        if (tmin[c] < tb) {
            tb = tmin[c];
            best = cand[c];
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HCK(hipMemsetAsync(p->status.p, 0, sizeof(gs::DevStatus), p->st));
    HCK(hipMemsetAsync(p->b1.p, 0, p->field_bytes(), p->st));
    sync(p);
    a.nl_e = best;
    std::lock_guard<std::mutex> g(mu);
    cache[key] = best;
}

// ---- autotuner of the QDHT tile shapes
// Tile shapes that give the same bits, timed on the plan's own operator and
// initial field (FIELD -> BIGBUFFER1, cleared afterwards) by the first plan of
// a (precision, N_R, N_T) in the process, separately for the column-skipping
// and the all-column transform (back, the one without the linear step):
// the tile order over the XCDs (all row tiles of a column tile per XCD, or all
// column tiles of a row tile) in every precision; mixed and FP32, one or two
// K groups per workgroup (256 or 512 threads); FP32 all-column, 64- or
// 32-column tiles.  The default (the first candidate) is replaced only by a
// clear (3 %) win.  GS_QDHT_KSPLIT fixes the mixed K groups for the process.
struct QShape {
    int ks = 2, order = 0;
    bool narrow = false;
};
void autotune_qdht(gs_plan* p)
{
    static std::mutex mu;
    static std::map<std::tuple<int, int, int>, std::pair<QShape, QShape>> cache;
    const auto key = std::make_tuple((int)p->prec, p->s.NR, p->s.NT);
    auto apply = [&](const QShape& a, const QShape& b) {
        p->qks[0] = a.ks;
        p->qorder[0] = a.order;
        p->qks[1] = b.ks;
        p->qorder[1] = b.order;
        p->f32_narrow_full = b.narrow;
    };
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            apply(it->second.first, it->second.second);
            return;
        }
    }
    hipEvent_t e0, e1;
    HCK(hipEventCreate(&e0));
    HCK(hipEventCreate(&e1));
    const int rp = p->rp, nt = p->s.NT;
    const void* A = p->prec == gs::kFP64 ? p->at64.p : p->amat.p;
    const int s3 = p->prec == gs::kMIXED && p->split3_back ? 1 : 0;
    QShape best[2];
    for (int mode = 0; mode < 2; ++mode) {     // 0: the live column blocks, 1: every column
        if (mode == 0 && p->nct == 0) continue;
        std::vector<QShape> cand;
        for (int order : {0, 1}) {
            if (p->prec == gs::kMIXED || (p->prec == gs::kFP32 && mode == 0)) {
                for (int ks : {2, 1}) cand.push_back(QShape{ks, order, false});
            } else if (p->prec == gs::kFP32) {
                for (bool nar : {false, true})
                    for (int ks : {2, 1}) cand.push_back(QShape{ks, order, nar});
            } else {
                cand.push_back(QShape{2, order, false});
            }
        }
        // candidates alternate (warm-up round first); each keeps its best time
        std::vector<float> tmin(cand.size(), 1e30f);
        for (int rep = 0; rep < 5; ++rep)
            for (size_t c = 0; c < cand.size(); ++c) {
                const QShape& q = cand[c];
                const gs::QCols qc = mode == 0
                    ? gs::QCols{p->ctile.as<int>(), p->nct, p->ztile.as<int>(), p->nzt, 0, s3, q.ks, q.order, 0}
                    : gs::QCols{nullptr, 0, nullptr, 0, 0, s3, q.ks, q.order, p->prec == gs::kFP32 && !q.narrow ? 1 : 0};
                HCK(hipEventRecord(e0, p->st));
                LCK(gs::launch_qdht(p->prec, rp, nt, A, p->field.p, p->b1.p, nullptr, p->st, nullptr, &qc));
                HCK(hipEventRecord(e1, p->st));
                sync(p);
                float t = 0;
                HCK(hipEventElapsedTime(&t, e0, e1));
                if (rep > 0) tmin[c] = std::min(tmin[c], t);
            }
        size_t b = 0;
        float tb = 0.97f * tmin[0];
        for (size_t c = 1; c < cand.size(); ++c)
            if (tmin[c] < tb) {
                tb = tmin[c];
                b = c;
            }
        best[mode] = cand[b];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HCK(hipMemsetAsync(p->b1.p, 0, p->field_bytes(), p->st));
    sync(p);
    apply(best[0], best[1]);
    std::lock_guard<std::mutex> g(mu);
    cache[key] = std::make_pair(best[0], best[1]);
}

// ---- one NL attempt + controller pieces
void nl_step(gs_plan* p, double dz, double* maxA2, double* maxNL2)
{
    HCK(hipEventRecord(p->ev[0], p->st));
    LCK(gs::launch_nl_rk4(p->prec, p->s.NT, p->field.p, p->b1.p, dz, p->nl, p->linemax.as<double>(), p->st));
    HCK(hipEventRecord(p->ev[1], p->st));
    LCK(gs::launch_reduce_max(p->s.NR, p->linemax.as<double>(), p->status.as<gs::DevStatus>(), p->st));
    HCK(hipEventRecord(p->ev[4], p->st));
    gs::DevStatus h;
    check_status(p, &h);
    *maxA2 = h.maxA2;
    *maxNL2 = h.maxNL2;
}

// the two QDHTs of an accepted step: B1 -> k-space with the linear step fused
// (FIELD_k), FIELD_k -> FIELD; with column skipping only the live blocks are
// computed, and the first QDHT back after FIELD was set clears the others
void enqueue_linear(gs_plan* p, const gs::LstepArgs& ls, hipEvent_t mid)
{
    const void* A = p->prec == gs::kFP64 ? p->at64.p : p->amat.p;
    const gs::QCols fwd = p->qcols(false, false), back = p->qcols(p->dead_dirty, true);
    LCK(gs::launch_qdht(p->prec, p->rp, p->s.NT, A, p->b1.p, nullptr, &ls, p->st, nullptr,
                        p->march_cols(false) ? &fwd : nullptr));
    if (mid) HCK(hipEventRecord(mid, p->st));
    LCK(gs::launch_qdht(p->prec, p->rp, p->s.NT, A, p->fk.p, p->field.p, nullptr, p->st, nullptr,
                        p->march_cols(true) ? &back : nullptr));
    p->dead_dirty = false;
}

void linear_half(gs_plan* p, double dz)
{
    p->ls.dz = dz;
    p->ls.seed = (uint32_t)p->STEP_N;
    HCK(hipEventRecord(p->ev[2], p->st));
    enqueue_linear(p, p->ls, p->ev[3]);
}

void finish_timings(gs_plan* p, hipEvent_t end)
{
    float a = 0, b = 0, c = 0, d = 0;
    sync(p);
    (void)hipEventElapsedTime(&a, p->ev[0], p->ev[1]);
    (void)hipEventElapsedTime(&d, p->ev[1], p->ev[4]);
    (void)hipEventElapsedTime(&b, p->ev[2], p->ev[3]);
    (void)hipEventElapsedTime(&c, p->ev[3], end);
    p->timings[0] = a;
    p->timings[1] = b;
    p->timings[2] = c;
    p->timings[3] = d;
}

// propagator(), propagator.cu:125-266 (apertures unsupported)
void propagator(gs_plan* p)
{
    const auto& Z = p->s.ZNET;
    const double tol = p->s.MAX_TOLERANCE;
    for (;;) {
        double maxA2, maxNL2;
        p->last_dz_nl = p->ZSTEP;
        nl_step(p, p->ZSTEP, &maxA2, &maxNL2);
        bool repeat = false;
        if (p->ZSTEP * p->ZSTEP * maxNL2 > tol * tol * maxA2) { repeat = true; p->STEP_N++; p->n_reject++; }
        p->ZSTEP = 0.5 * tol * std::sqrt(maxA2 / maxNL2);
        const int nz = std::max(p->n_Z, 1);
        if (p->ZSTEP < 1e-10 * (Z[nz] - Z[nz - 1]) || !(p->ZSTEP > 0))
            throw gs::Error(GS_E_PHYSICS, "Minimum Z step size reached, collapse is possible!");
        p->ZSTEP = std::min(p->ZSTEP, Z[nz] - p->CURRENT_Z);
        if (!repeat) {
            linear_half(p, p->ZSTEP);
            hipEvent_t end;
            HCK(hipEventCreate(&end));
            HCK(hipEventRecord(end, p->st));
            finish_timings(p, end);
            (void)hipEventDestroy(end);
            return;
        }
    }
}

// ---- device-driven march
// one step attempt: NL RK4 at ctl.dz, controller, linear half at ctl.dz_lin
// when accepted, bookkeeping; every kernel reads its dz and gate from ctl
void enqueue_slot(gs_plan* p)
{
    gs::DevCtl* c = p->ctl.as<gs::DevCtl>();
    gs::NLArgs a = p->nl;
    a.ctl = c;
    LCK(gs::launch_nl_rk4(p->prec, p->s.NT, p->field.p, p->b1.p, 0.0, a, p->linemax.as<double>(), p->st));
    LCK(gs::launch_controller(p->s.NR, p->linemax.as<double>(), p->status.as<gs::DevStatus>(), c, p->st));
    gs::LstepArgs l = p->ls;
    l.ctl = c;
    const void* A = p->prec == gs::kFP64 ? p->at64.p : p->amat.p;
    // the QDHT back carries its clearing blocks; they act only when DevCtl::dead_dirty is set
    const gs::QCols fwd = p->qcols(false, false), back = p->qcols(false, true);
    LCK(gs::launch_qdht(p->prec, p->rp, p->s.NT, A, p->b1.p, nullptr, &l, p->st, nullptr, p->march_cols(false) ? &fwd : nullptr));
    LCK(gs::launch_qdht(p->prec, p->rp, p->s.NT, A, p->fk.p, p->field.p, nullptr, p->st, c, p->march_cols(true) ? &back : nullptr));
    LCK(gs::launch_finish(c, p->st));
}

// the slot captured once into a graph, so a step costs one graph launch
void launch_slot(gs_plan* p)
{
    if (!p->slot_exec && !std::getenv("GS_NO_GRAPH")) {
        hipGraph_t g = nullptr;
        HCK(hipStreamBeginCapture(p->st, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_slot(p);
        } catch (...) {
            (void)hipStreamEndCapture(p->st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HCK(hipStreamEndCapture(p->st, &g));
        const hipError_t e = hipGraphInstantiate(&p->slot_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HCK(e);
    }
    if (p->slot_exec) HCK(hipGraphLaunch(p->slot_exec, p->st));
    else enqueue_slot(p);
}

// the plan's host mirror <- a controller state read back from the device
void pull_state(gs_plan* p, const gs::DevCtl& h)
{
    p->ZSTEP = h.dz;
    p->CURRENT_Z = h.z;
    p->STEP_N = (long)h.steps;
    p->n_reject = (long)h.rejects;
    p->dead_dirty = h.dead_dirty != 0;
}

// host mirror <- device controller state
void pull_ctl(gs_plan* p)
{
    HCK(hipMemcpyAsync(p->ctl_host, p->ctl.p, sizeof(gs::DevCtl), hipMemcpyDeviceToHost, p->st));
    sync(p);
    pull_state(p, *p->ctl_host);
}

// arm the controller for the plane ZNET[n_Z] from the host state
// controller state for the plane ZNET[n_Z] from the plan's host state
void fill_ctl(const gs_plan* p, gs::DevCtl& h)
{
    const auto& Z = p->s.ZNET;
    const int nz = std::max(p->n_Z, 1);
    h = gs::DevCtl{};
    h.dz = p->ZSTEP;
    h.z = p->CURRENT_Z;
    h.z_plane = Z[p->n_Z];
    h.dz_floor = 1e-10 * (Z[nz] - Z[nz - 1]);
    h.tol = p->s.MAX_TOLERANCE;
    h.steps = p->STEP_N;
    h.rejects = p->n_reject;
    h.dead_dirty = p->dead_dirty ? 1 : 0;
}

void arm_ctl(gs_plan* p)
{
    if (!p->ctl.p) {
        p->ctl.alloc(sizeof(gs::DevCtl));
        HCK(hipHostMalloc((void**)&p->ctl_host, sizeof(gs::DevCtl), 0));
    }
    fill_ctl(p, *p->ctl_host);
    HCK(hipMemcpyAsync(p->ctl.p, p->ctl_host, sizeof(gs::DevCtl), hipMemcpyHostToDevice, p->st));
}


[[noreturn]] void ctl_error(gs_plan* p, int err)
{
    HCK(hipMemsetAsync(p->status.p, 0, sizeof(gs::DevStatus), p->st));
    sync(p);
    if (err == 2) throw gs::Error(GS_E_PHYSICS, "float_type photoionization_function(...): Intensity is nan!");
    if (err == 3) throw gs::Error(GS_E_PHYSICS, "Attention!!!! Intensity is too big! Keep away, blast is possible.");
    throw gs::Error(GS_E_PHYSICS, "Minimum Z step size reached, collapse is possible!");
}

// ---- .udf output (dump_types.cpp:25-72)
struct Udf {
    std::FILE* f = nullptr;
    int cplx = 0;
    Udf() = default;
    Udf(const Udf&) = delete;
    Udf& operator=(const Udf&) = delete;
    Udf(Udf&& o) noexcept : f(o.f), cplx(o.cplx) { o.f = nullptr; }
    ~Udf()
    {
        if (f) std::fclose(f);
    }
    // resume: continue a file of the same layout written by an earlier call
    // (its header and planes are kept; writes are positioned per plane)
    void open(const std::string& path, int iscomplex, int tN, int xN, int yN, int zN, bool resume)
    {
        cplx = iscomplex;
        const int h[6] = {8, iscomplex, tN, xN, yN, zN};
        if (resume && (f = std::fopen(path.c_str(), "r+b")) != nullptr) {
            char tag0[7];
            int h0[6];
            if (std::fread(tag0, 1, 7, f) == 7 && std::fread(h0, sizeof(int), 6, f) == 6 &&
                std::memcmp(tag0, "UDF1.0", 7) == 0 && std::memcmp(h0, h, sizeof h) == 0)
                return;
            std::fclose(f);   // another layout: start the file again
        }
        f = std::fopen(path.c_str(), "w+b");
        if (!f) throw gs::Error(GS_E_CONFIG, "void dump_type_empty_class :: init_dumpfile : Unable to open file!");
        const char tag[] = "UDF1.0";
        std::fwrite(tag, 1, sizeof tag, f);
        std::fwrite(h, sizeof(int), 6, f);
        std::fflush(f);
    }
    void write(long long elem_ofs, const double* data, size_t nelem)
    {
        const long long off = 31 + elem_ofs * (cplx ? 16 : 8);
        fseeko(f, off, SEEK_SET);
        std::fwrite(data, cplx ? 16 : 8, nelem, f);
        std::fflush(f);
    }
};

struct DumpSet {
    std::vector<Udf> files;
};

void open_dumps(gs_plan* p, const std::string& outdir, DumpSet* ds)
{
    const auto& s = p->s;
    const bool resume = p->dumps_started;   // continue the planes an earlier call wrote
    ds->files.resize(s.dumps.size());
    for (size_t i = 0; i < s.dumps.size(); ++i) {
        const auto& d = s.dumps[i];
        const std::string path = outdir + "/" + d.filename;
        if (d.type == "full") ds->files[i].open(path, 1, s.NT, s.NR, 1, s.NZ / d.nz_denom, resume);
        else if (d.type == "maxI" || d.type == "flux" || d.type == "plasma_max") ds->files[i].open(path, 0, 1, s.NR, 1, s.NZ, resume);
        else if (d.type == "field_axis") ds->files[i].open(path, 1, s.NT, 1, 1, s.NZ, resume);
        else ds->files[i].open(path, 0, s.NT, 1, 1, s.NZ, resume);   // average_spectrum, plasma_axis
    }
}

void do_dumps(gs_plan* p, DumpSet* ds)
{
    const auto& s = p->s;
    const int nr = s.NR, nt = s.NT, nz = p->n_Z;
    p->dumps_started = true;
    double* dflux = p->diag.as<double>();
    double* dmax = dflux + p->rp;
    double* drho = dmax + p->rp;
    bool have_plain = false;
    std::vector<double> flux(nr), maxI(nr), rho(nr);
    for (size_t i = 0; i < s.dumps.size(); ++i) {
        const auto& d = s.dumps[i];
        Udf& u = ds->files[i];
        bool filtered = false;
        for (double v : d.wfilter) filtered |= v != 1.0;
        if (d.type == "flux" || d.type == "maxI" || d.type == "plasma_max") {
            if (filtered) {
                LCK(gs::launch_line_diag(p->prec, nt, p->field.p, nr, p->wfilt[i].as<double>(), s.TSTEP, p->nl, dflux, dmax, drho,
                                         nullptr, nullptr, p->st));
                std::vector<double> o(nr);
                HCK(hipMemcpyAsync(o.data(), d.type == "flux" ? dflux : dmax, nr * 8, hipMemcpyDeviceToHost, p->st));
                check_status(p, nullptr);
                u.write((long long)nr * nz, o.data(), nr);
                have_plain = false;
                continue;
            }
            if (!have_plain) {
                LCK(gs::launch_line_diag(p->prec, nt, p->field.p, nr, nullptr, s.TSTEP, p->nl, dflux, dmax, drho, nullptr,
                                         nullptr, p->st));
                HCK(hipMemcpyAsync(flux.data(), dflux, nr * 8, hipMemcpyDeviceToHost, p->st));
                HCK(hipMemcpyAsync(maxI.data(), dmax, nr * 8, hipMemcpyDeviceToHost, p->st));
                HCK(hipMemcpyAsync(rho.data(), drho, nr * 8, hipMemcpyDeviceToHost, p->st));
                check_status(p, nullptr);
                have_plain = true;
            }
            const std::vector<double>& o = d.type == "flux" ? flux : (d.type == "maxI" ? maxI : rho);
            u.write((long long)nr * nz, o.data(), nr);
        } else if (d.type == "field_axis" || d.type == "full") {
            if (d.type == "full" && nz % d.nz_denom) continue;
            const int lines = d.type == "full" ? nr : 1;
            if (p->et.n < (size_t)lines * nt * p->relem) p->et.alloc((size_t)p->rp * nt * p->relem);
            LCK(gs::launch_fft_lines(p->prec, nt, p->field.p, p->et.p, lines, +1, 1.0 / nt, nullptr, p->tw.p, p->st));
            std::vector<double> h((size_t)lines * nt * 2);
            download_field64(p, p->et.p, h.data(), lines);
            if (d.type == "full") u.write((long long)nt * ((long long)nr * (nz / d.nz_denom)), h.data(), (size_t)lines * nt);
            else u.write((long long)nt * nz, h.data(), nt);
        } else if (d.type == "average_spectrum") {
            if (p->rhot.n < (size_t)nt * 8) p->rhot.alloc((size_t)std::max(nt, p->rp) * nt * 8);
            LCK(gs::launch_avg_spectrum(p->prec, nr, nt, p->field.p, p->avgw.as<double>(), p->rhot.as<double>(), p->st));
            std::vector<double> h(nt);
            HCK(hipMemcpyAsync(h.data(), p->rhot.p, nt * 8, hipMemcpyDeviceToHost, p->st));
            sync(p);
            for (double& v : h) v *= s.XMAX * s.XMAX / 2.0;
            u.write((long long)nt * nz, h.data(), nt);
        } else if (d.type == "plasma_axis") {
            if (p->rhot.n < (size_t)nt * 8) p->rhot.alloc((size_t)std::max(nt, p->rp) * nt * 8);
            LCK(gs::launch_line_diag(p->prec, nt, p->field.p, 1, nullptr, s.TSTEP, p->nl, nullptr, nullptr, nullptr, nullptr,
                                     p->rhot.as<double>(), p->st));
            std::vector<double> h(nt);
            HCK(hipMemcpyAsync(h.data(), p->rhot.p, nt * 8, hipMemcpyDeviceToHost, p->st));
            check_status(p, nullptr);
            u.write((long long)nt * nz, h.data(), nt);
        }
    }
}

template <class F>
int guarded(F&& f)
{
    try {
        return f();
    } catch (const gs::Error& e) {
        return fail(e.code, e.what());
    } catch (const std::exception& e) {
        return fail(GS_E_ARG, e.what());
    }
}

}  // namespace

// ---- batched march: the step attempts of several runs, one launch per phase
// (every run's NL lines; the controllers; every run's QDHT + linear step; every
// run's QDHT back; the bookkeeping), captured once into a graph.  Runs that
// sit at a plane or have failed are no-ops in it (their DevCtl::state).
struct BatchMarch {
    int n = 0;
    hipStream_t st = nullptr;
    DevBuf ctls, nlruns, ctlruns;
    gs::DevCtl* host = nullptr;     // pinned mirror of the n controllers
    struct NLGroup {
        gs::Prec prec;
        int nt, nl_e;
        bool mfma;
        DevBuf items;
        int nitems;
    };
    struct GemmLaunch {
        gs::Prec prec;
        bool lstep;
        DevBuf runs, brun;
        int nblocks;
        int ks;
    };
    std::vector<NLGroup> nlg;
    std::vector<GemmLaunch> gl;
    hipGraphExec_t exec = nullptr;
    size_t graph_nodes = 0;         // kernels in the captured attempt of all n runs

    BatchMarch(gs_plan* const* plans, int n_) : n(n_)
    {
        HCK(hipStreamCreate(&st));
        ctls.alloc(sizeof(gs::DevCtl) * n);
        HCK(hipHostMalloc((void**)&host, sizeof(gs::DevCtl) * n, 0));
        std::memset(host, 0, sizeof(gs::DevCtl) * n);
        gs::DevCtl* dctl = ctls.as<gs::DevCtl>();
        std::vector<gs::NLRun> nr(n);
        std::vector<gs::CtlRun> cr(n);
        for (int i = 0; i < n; ++i) {
            gs_plan* p = plans[i];
            nr[i] = gs::NLRun{p->field.p, p->b1.p, p->linemax.as<double>(), p->nl};
            nr[i].a.ctl = dctl + i;
            cr[i] = gs::CtlRun{p->s.NR, p->linemax.as<double>(), p->status.as<gs::DevStatus>(), dctl + i};
            // NL lines: one group per (precision, N_T, line shape)
            const bool mfma = p->nl.mfma_fwd || p->nl.mfma_inv;
            NLGroup* g = nullptr;
            for (auto& x : nlg)
                if (x.prec == p->prec && x.nt == p->s.NT && x.nl_e == p->nl.nl_e && x.mfma == mfma) g = &x;
            if (!g) {
                nlg.push_back(NLGroup{p->prec, p->s.NT, p->nl.nl_e, mfma, DevBuf(), 0});
                g = &nlg.back();
            }
            g->nitems += p->s.NR;
        }
        upload(nlruns, nr.data(), sizeof(gs::NLRun) * n);
        upload(ctlruns, cr.data(), sizeof(gs::CtlRun) * n);
        for (auto& g : nlg) {
            std::vector<int2> it;
            for (int i = 0; i < n; ++i) {
                gs_plan* p = plans[i];
                const bool mfma = p->nl.mfma_fwd || p->nl.mfma_inv;
                if (p->prec == g.prec && p->s.NT == g.nt && p->nl.nl_e == g.nl_e && mfma == g.mfma)
                    for (int r = 0; r < p->s.NR; ++r) it.push_back(int2{i, r});
            }
            upload(g.items, it.data(), sizeof(int2) * it.size());
        }
        // QDHTs: one launch per precision and direction; each run's blocks start on a multiple of 8
        for (gs::Prec pr : {gs::kFP64, gs::kFP32, gs::kMIXED})
            for (bool lstep : {true, false}) {
                std::vector<gs::GemmRun> runs;
                std::vector<int> brun;
                for (int i = 0; i < n; ++i) {
                    gs_plan* p = plans[i];
                    if (p->prec != pr) continue;
                    gs::GemmRun g{};
                    g.A = pr == gs::kFP64 ? p->at64.p : p->amat.p;
                    g.X = lstep ? p->b1.p : p->fk.p;
                    g.Y = lstep ? nullptr : p->field.p;
                    g.K = p->rp;
                    g.NC = p->s.NT;
                    if (lstep) g.la = p->ls;
                    g.la.ctl = dctl + i;
                    g.qc = p->qcols(false, !lstep);
                    g.nb = gs::qdht_blocks(pr, p->rp, p->s.NT, &g.qc, !lstep, true);
                    g.b0 = (int)brun.size();
                    const int padded = (g.nb + 7) / 8 * 8;
                    brun.insert(brun.end(), padded, (int)runs.size());
                    runs.push_back(g);
                }
                if (runs.empty()) continue;
                gl.push_back(GemmLaunch{pr, lstep, DevBuf(), DevBuf(), (int)brun.size(), runs[0].qc.ks});
                upload(gl.back().runs, runs.data(), sizeof(gs::GemmRun) * runs.size());
                upload(gl.back().brun, brun.data(), sizeof(int) * brun.size());
            }
        // capture one batched attempt
        hipGraph_t graph = nullptr;
        HCK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            enqueue();
        } catch (...) {
            (void)hipStreamEndCapture(st, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HCK(hipStreamEndCapture(st, &graph));
        (void)hipGraphGetNodes(graph, nullptr, &graph_nodes);
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HCK(e);
    }
    BatchMarch(const BatchMarch&) = delete;
    BatchMarch& operator=(const BatchMarch&) = delete;
    ~BatchMarch()
    {
        if (st) (void)hipStreamSynchronize(st);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (host) (void)hipHostFree(host);
        if (st) (void)hipStreamDestroy(st);
    }
    void enqueue()
    {
        for (auto& g : nlg)
            LCK(gs::launch_nl_rk4_batch(g.prec, g.nt, g.nl_e, g.mfma, nlruns.as<gs::NLRun>(), g.items.as<int2>(),
                                        g.nitems, st));
        LCK(gs::launch_controller_batch(ctlruns.as<gs::CtlRun>(), n, st));
        for (bool lstep : {true, false})
            for (auto& g : gl)
                if (g.lstep == lstep)
                    LCK(gs::launch_qdht_batch(g.prec, lstep, g.runs.as<gs::GemmRun>(), g.brun.as<int>(), g.nblocks, st, g.ks));
        LCK(gs::launch_finish_batch(ctls.as<gs::DevCtl>(), n, st));
    }
    void launch() { HCK(hipGraphLaunch(exec, st)); }
    void send() { HCK(hipMemcpyAsync(ctls.p, host, sizeof(gs::DevCtl) * n, hipMemcpyHostToDevice, st)); }
    void receive()
    {
        HCK(hipMemcpyAsync(host, ctls.p, sizeof(gs::DevCtl) * n, hipMemcpyDeviceToHost, st));
        HCK(hipStreamSynchronize(st));
    }
};

// ====================================================================== C-ABI
extern "C" {

const char* gs_last_error(void) { return g_err.c_str(); }

int gs_plan_create(const char* cfg_path, int nover, const char* const* over, int precision, const char* workdir, gs_plan** out)
{
    return guarded([&]() {
        if (!cfg_path || !out) return fail(GS_E_ARG, "null argument");
        if (precision < 0 || precision > 2) return fail(GS_E_ARG, "precision must be GS_FP64, GS_FP32 or GS_MIXED");
        std::vector<std::string> ov;
        for (int i = 0; i < nover; ++i) ov.emplace_back(over[i]);
        std::unique_ptr<gs_plan> p(new gs_plan);
        p->prec = (gs::Prec)precision;
        gs::Config cfg(cfg_path, ov);
        p->s = gs::build_setup(cfg, workdir ? workdir : "");
        build_plan(p.get());
        autotune_nl(p.get());
        autotune_qdht(p.get());
        *out = p.release();
        return 0;
    });
}

void gs_plan_destroy(gs_plan* p)
{
    if (!p) return;
    if (p->st) (void)hipStreamSynchronize(p->st);
    for (auto& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    if (p->slot_exec) (void)hipGraphExecDestroy(p->slot_exec);
    if (p->ctl_host) (void)hipHostFree(p->ctl_host);
    hipStream_t st = p->st;
    delete p;
    if (st) (void)hipStreamDestroy(st);
}

int gs_dims(gs_plan* p, int* nr, int* nt, int* nz)
{
    if (!p) return fail(GS_E_ARG, "null plan");
    if (nr) *nr = p->s.NR;
    if (nt) *nt = p->s.NT;
    if (nz) *nz = p->s.NZ;
    return 0;
}

static int get_scalar(gs_plan* p, const char* name, double* v);

int gs_get_scalar(gs_plan* p, const char* name, double* v)
{
    return guarded([&]() { return get_scalar(p, name, v); });
}

static int get_scalar(gs_plan* p, const char* name, double* v)
{
    if (!p || !name || !v) return fail(GS_E_ARG, "null argument");
    const std::string n(name);
    const auto& s = p->s;
    if (n == "ZSTEP") *v = p->ZSTEP;
    else if (n == "CURRENT_Z") *v = p->CURRENT_Z;
    else if (n == "NL_SAMPLES_PER_THREAD") *v = p->nl.nl_e;
    else if (n == "LAST_DZ_NL") *v = p->last_dz_nl;
    else if (n == "ADJOINT_STEPS") *v = (double)p->adj_rk.size();
    else if (n == "ADJOINT_SEGMENT") *v = p->adj_seg;
    else if (n == "ADJOINT_CHECKPOINT_BYTES") *v = p->adj_budget;
    else if (n == "STEP_N") *v = (double)p->STEP_N;
    else if (n == "n_Z") *v = p->n_Z;
    else if (n == "n_reject") *v = (double)p->n_reject;
    else if (n == "OMEGA0") *v = s.OMEGA0;
    else if (n == "OMEGA_MIN") *v = s.OMEGA_MIN;
    else if (n == "OMEGA_MAX") *v = s.OMEGA_MAX;
    else if (n == "GROUP_VELOCITY") *v = s.GROUP_VELOCITY;
    else if (n == "TSTEP") *v = s.TSTEP;
    else if (n == "TMIN") *v = s.TMIN;
    else if (n == "TMAX") *v = s.TMAX;
    else if (n == "XMAX") *v = s.XMAX;
    else if (n == "MAX_TOLERANCE") *v = s.MAX_TOLERANCE;
    else if (n == "NONLIN_REFRINDEX") *v = s.n2;
    else if (n == "AVALANCHE_CROSSSECTION") *v = s.avalanche;
    else if (n == "PONDEROMOTIVE_COEFFICIENT") *v = s.ponderomotive;
    else if (n == "IONIZATION_POTENTIAL") *v = s.ion_potential;
    else if (n == "RECOMBINATION_TAU") *v = s.recomb_tau;
    else if (n == "NEUTRAL_DENSITY") *v = s.neutral_density;
    else if (n == "qS") *v = s.qS;
    else if (n == "precision") *v = p->prec;
    else if (n == "QDHT_SKIP_COLUMNS") *v = p->qskip ? 1 : 0;
    else if (n == "QDHT_K_GROUPS") *v = p->qks[p->qskip ? 0 : 1];
    else if (n == "QDHT_TILE_ORDER") *v = p->qorder[p->qskip ? 0 : 1];
    else if (n == "BATCH_GRAPH_KERNELS") *v = p->batch_nodes;
    else if (n == "QDHT_F32_NARROW_FULL") *v = p->f32_narrow_full ? 1 : 0;
    else if (n == "SCAN_FALLBACKS") {
        // line launches whose plasma scan needed the chunk-sequential fallback, since plan creation
        gs::DevStatus h;
        HCK(hipMemcpyAsync(&h, p->status.p, sizeof h, hipMemcpyDeviceToHost, p->st));
        sync(p);
        *v = h.fallbacks;
    }
    else if (n == "QDHT_LIVE_BLOCKS") *v = p->nct;
    else if (n == "WAVENUMBER0_re") *v = s.WAVENUMBER0.real();
    else return fail(GS_E_ARG, "unknown scalar " + n);
    return 0;
}

int gs_set_scalar(gs_plan* p, const char* name, double v)
{
    if (!p || !name) return fail(GS_E_ARG, "null argument");
    const std::string n(name);
    if (n == "ZSTEP") p->ZSTEP = v;
    else if (n == "CURRENT_Z") p->CURRENT_Z = v;
    else if (n == "STEP_N") {
        // step attempts so far; it also seeds the mixed policy's unbiased rounding
        // of each step, so a march resumed from a saved state sets it back
        if (!(v >= 0 && v < 9.0e15) || v != (double)(long)v) return fail(GS_E_ARG, "STEP_N must be a non-negative integer");
        p->STEP_N = (long)v;
    }
    else if (n == "n_Z") {
        if (!(v >= 0 && v <= p->s.NZ - 1) || v != (double)(int)v)
            return fail(GS_E_ARG, "n_Z must be an integer plane index in 0..N_Z-1");
        p->n_Z = (int)v;
        if (p->n_Z == 0) p->dumps_started = false;   // a march from plane 0 starts its files again
    }
    else if (n == "MAX_TOLERANCE") p->s.MAX_TOLERANCE = v;
    else if (n == "QDHT_SKIP_COLUMNS") {
        // 1: the march's QDHTs compute only the 32-column blocks holding unmasked
        // points; 0: every column (the step graph is captured again)
        const bool on = v != 0 && p->nct > 0;
        if (on != p->qskip && p->slot_exec) {
            (void)hipGraphExecDestroy(p->slot_exec);
            p->slot_exec = nullptr;
        }
        p->qskip = on;
    }
    else if (n == "QDHT_K_GROUPS" || n == "QDHT_TILE_ORDER" || n == "QDHT_F32_NARROW_FULL") {
        // the march's GEMM tile shape (normally autotuned; the same bits either way):
        // K groups of the mixed / FP32 tiles in the current column mode (1 or 2),
        // the tile order over the XCDs in the current column mode (0 or 1), or
        // FP32 all-column transforms on the narrow tiles (0 or 1)
        const bool groups = n == "QDHT_K_GROUPS";
        if (groups ? !(v == 1 || v == 2) : !(v == 0 || v == 1))
            return fail(GS_E_ARG, n + (groups ? " must be 1 or 2" : " must be 0 or 1"));
        if (p->slot_exec) {
            (void)hipGraphExecDestroy(p->slot_exec);
            p->slot_exec = nullptr;
        }
        if (groups) p->qks[p->qskip ? 0 : 1] = (int)v;
        else if (n == "QDHT_TILE_ORDER") p->qorder[p->qskip ? 0 : 1] = (int)v;
        else p->f32_narrow_full = v != 0;
    }
    else if (n == "ADJOINT_CHECKPOINT_BYTES") {
        if (v != v) return fail(GS_E_ARG, "ADJOINT_CHECKPOINT_BYTES must be a number");
        p->adj_budget = v < 0 ? -1.0 : v;     // negative: 60 % of the free device memory
    }
    else return fail(GS_E_ARG, "scalar " + n + " is not settable");
    return 0;
}

static int copy_table(const gs::Setup& s, const char* name, double* out, size_t n);

int gs_get_table(gs_plan* p, const char* name, double* out, size_t n)
{
    if (!p || !name || !out) return fail(GS_E_ARG, "null argument");
    return copy_table(p->s, name, out, n);
}

int gs_setup_table(const char* cfg_path, int nover, const char* const* over, const char* name, double* out, size_t n)
{
    return guarded([&]() {
        if (!cfg_path || !name || !out) return fail(GS_E_ARG, "null argument");
        std::vector<std::string> ov;
        for (int i = 0; i < nover; ++i) ov.emplace_back(over[i]);
        gs::Config cfg(cfg_path, ov);
        const gs::Setup s = gs::build_setup(cfg, "");
        return copy_table(s, name, out, n);
    });
}

static int copy_table(const gs::Setup& s, const char* name, double* out, size_t n)
{
    const std::string nm(name);
    auto put = [&](const std::vector<double>& v) {
        if (n < v.size()) return fail(GS_E_ARG, "output too small for " + nm);
        std::memcpy(out, v.data(), v.size() * 8);
        return 0;
    };
    auto putc = [&](const std::vector<cd>& v) {
        if (n < 2 * v.size()) return fail(GS_E_ARG, "output too small for " + nm);
        std::memcpy(out, v.data(), v.size() * 16);
        return 0;
    };
    if (nm == "omega") return put(s.omega);
    if (nm == "wavenum") return putc(s.wavenum);
    if (nm == "ho") return put(s.ho);
    if (nm == "kerr") return put(s.kerr);
    if (nm == "plasma_func") return putc(s.plasma_func);
    if (nm == "raman_function") return putc(s.raman_function);
    if (nm == "ionrate_ln") return put(s.ionrate_ln);
    if (nm == "qdht_x") return put(s.qx);
    if (nm == "qdht_m1") return put(s.qm1);
    if (nm == "qdht_c") return put(s.qC);
    if (nm == "znet") return put(s.ZNET);
    return fail(GS_E_ARG, "unknown table " + nm);
}

// after a buffer was set through the API: FIELD_k is zero wherever the linear
// step zeroes FIELD (the skipped QDHT columns rely on it), and a new FIELD may
// hold values in the skipped columns until a QDHT back clears them
static void mark_written(gs_plan* p, int buffer)
{
    if (buffer == GS_FK)
        LCK(gs::launch_b2_split(p->prec, p->fk.p, p->fk.p, p->lgfac.p, (size_t)p->rp * p->s.NT, p->st));
    if (buffer == GS_FIELD) p->dead_dirty = true;
}

int gs_get(gs_plan* p, int buffer, double* host)
{
    return guarded([&]() {
        if (!p || !host) return fail(GS_E_ARG, "null argument");
        void* src = p->buf(buffer);
        if (buffer == GS_B2) {
            // B2 = FIELD_k on the unmasked points, its set-up values on the masked ones
            LCK(gs::launch_b2_merge(p->prec, p->fk.p, p->b2.p, p->tmp.p, p->lgfac.p, (size_t)p->rp * p->s.NT, p->st));
            src = p->tmp.p;
        }
        download_field64(p, src, host, p->s.NR);
        return 0;
    });
}

int gs_put(gs_plan* p, int buffer, const double* host)
{
    return guarded([&]() {
        if (!p || !host) return fail(GS_E_ARG, "null argument");
        DevBuf tmp;
        void* dst = p->buf(buffer);
        const size_t n = (size_t)p->s.NR * p->s.NT;
        upload(tmp, host, n * 16);
        HCK(hipMemsetAsync(dst, 0, p->field_bytes(), p->st));
        LCK(gs::launch_convert(p->prec, tmp.as<double>(), dst, n, 1, p->st));
        if (buffer == GS_B2)   // the unmasked part is the state the next linear step updates
            LCK(gs::launch_b2_split(p->prec, p->b2.p, p->fk.p, p->lgfac.p, (size_t)p->rp * p->s.NT, p->st));
        mark_written(p, buffer);
        sync(p);
        return 0;
    });
}

int gs_nl_step(gs_plan* p, double dz, double* maxA2, double* maxNL2)
{
    return guarded([&]() {
        if (!p) return fail(GS_E_ARG, "null plan");
        double a, b;
        nl_step(p, dz, &a, &b);
        if (maxA2) *maxA2 = a;
        if (maxNL2) *maxNL2 = b;
        return 0;
    });
}

int gs_nl_response(gs_plan* p, const double* in, double* out, int n_lines)
{
    return guarded([&]() {
        if (!p || !in || !out || n_lines < 1) return fail(GS_E_ARG, "bad argument");
        const int nt = p->s.NT;
        const size_t n = (size_t)n_lines * nt;
        DevBuf d64, din, dout;
        upload(d64, in, n * 16);
        din.alloc(n * p->relem);
        dout.alloc(n * p->relem);
        LCK(gs::launch_convert(p->prec, d64.as<double>(), din.p, n, 1, p->st));
        LCK(gs::launch_nl_response(p->prec, nt, din.p, dout.p, n_lines, p->nl, p->st));
        LCK(gs::launch_convert(p->prec, d64.as<double>(), dout.p, n, 0, p->st));
        HCK(hipMemcpyAsync(out, d64.p, n * 16, hipMemcpyDeviceToHost, p->st));
        check_status(p, nullptr);
        return 0;
    });
}

int gs_nl_response_noalloc(gs_plan* p, const double* in, double* out, int n_lines)
{
    return guarded([&]() {
        if (!p || !in || !out || n_lines < 1) return fail(GS_E_ARG, "bad argument");
        const int nt = p->s.NT;
        const size_t n = (size_t)n_lines * nt;
        if (n > gs::kNoAllocMaxSamples)
            return fail(GS_E_ARG, "gs_nl_response_noalloc: n_lines * N_T = " + std::to_string(n) +
                                      " exceeds the fixed scratch of " + std::to_string(gs::kNoAllocMaxSamples) +
                                      " samples");
        HCK(hipSetDevice(p->dev));
        void* base = nullptr;
        LCK(gs::noalloc_scratch(&base));
        // [FP64 host copy: 16 n][input, field precision][output, field precision], 256-byte aligned
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        char* sc = static_cast<char*>(base);
        double* d64 = reinterpret_cast<double*>(sc);
        void* din = sc + up(n * 16);
        void* dout = sc + up(n * 16) + up(n * p->relem);
        HCK(hipMemcpyAsync(d64, in, n * 16, hipMemcpyHostToDevice, p->st));
        LCK(gs::launch_convert(p->prec, d64, din, n, 1, p->st));
        LCK(gs::launch_nl_response(p->prec, nt, din, dout, n_lines, p->nl, p->st));
        LCK(gs::launch_convert(p->prec, d64, dout, n, 0, p->st));
        HCK(hipMemcpyAsync(out, d64, n * 16, hipMemcpyDeviceToHost, p->st));
        check_status(p, nullptr);
        return 0;
    });
}

int gs_qdht(gs_plan* p, int buffer)
{
    return guarded([&]() {
        if (!p) return fail(GS_E_ARG, "null plan");
        void* b = p->buf(buffer);
        const size_t n = (size_t)p->rp * p->s.NT;
        const void* A = p->prec == gs::kFP64 ? p->at64.p : p->amat.p;
        if (buffer == GS_B2) {
            // the full B2 in, the full B2 out, then its unmasked part back into FIELD_k
            LCK(gs::launch_b2_merge(p->prec, p->fk.p, p->b2.p, p->tmp.p, p->lgfac.p, n, p->st));
            LCK(gs::launch_qdht(p->prec, p->rp, p->s.NT, A, p->tmp.p, p->b2.p, nullptr, p->st));
            LCK(gs::launch_b2_split(p->prec, p->b2.p, p->fk.p, p->lgfac.p, n, p->st));
        } else {
            LCK(gs::launch_qdht(p->prec, p->rp, p->s.NT, A, b, p->tmp.p, nullptr, p->st));
            HCK(hipMemcpyAsync(b, p->tmp.p, p->field_bytes(), hipMemcpyDeviceToDevice, p->st));
        }
        mark_written(p, buffer);
        sync(p);
        return 0;
    });
}

int gs_linear_step(gs_plan* p, double dz)
{
    return guarded([&]() {
        if (!p) return fail(GS_E_ARG, "null plan");
        p->ls.dz = dz;
        p->ls.seed = (uint32_t)p->STEP_N;
        LCK(gs::launch_lstep(p->prec, p->rp, p->s.NT, p->b1.p, p->ls, p->st));
        sync(p);
        return 0;
    });
}

int gs_step(gs_plan* p)
{
    return guarded([&]() {
        if (!p) return fail(GS_E_ARG, "null plan");
        // the main loop's plane bookkeeping (solver_main.cpp:206-212), without the dumps
        const auto& Z = p->s.ZNET;
        while (p->n_Z < p->s.NZ - 1 && p->CURRENT_Z >= Z[p->n_Z]) p->n_Z++;
        if (p->CURRENT_Z >= Z[p->s.NZ - 1]) return fail(GS_E_ARG, "gs_step: the march is at its last plane");
        propagator(p);
        p->STEP_N++;
        p->CURRENT_Z += p->ZSTEP;
        return 0;
    });
}

int gs_fixed_step(gs_plan* p, double dz, double* maxA2, double* maxNL2)
{
    return guarded([&]() {
        if (!p) return fail(GS_E_ARG, "null plan");
        double a, b;
        nl_step(p, dz, &a, &b);
        linear_half(p, dz);
        hipEvent_t end;
        HCK(hipEventCreate(&end));
        HCK(hipEventRecord(end, p->st));
        finish_timings(p, end);
        (void)hipEventDestroy(end);
        p->ZSTEP = dz;
        p->CURRENT_Z += dz;
        p->STEP_N++;
        if (maxA2) *maxA2 = a;
        if (maxNL2) *maxNL2 = b;
        return 0;
    });
}

int gs_march(gs_plan* p, const char* outdir, long max_steps)
{
    return guarded([&]() {
        if (!p) return fail(GS_E_ARG, "null plan");
        DumpSet ds;
        open_dumps(p, outdir ? outdir : ".", &ds);
        const auto& Z = p->s.ZNET;
        long steps = 0;
        if (std::getenv("GS_HOST_MARCH")) {
            // host-driven controller: one synchronisation per step attempt
            for (;;) {
                if (p->CURRENT_Z >= Z[p->n_Z]) {
                    do_dumps(p, &ds);
                    if (p->n_Z == p->s.NZ - 1) return 0;
                    p->n_Z++;
                }
                if (max_steps > 0 && steps >= max_steps) return 1;
                if (p->n_Z == 0) throw gs::Error(GS_E_PHYSICS, "CURRENT_Z is below the first dump plane");
                propagator(p);
                p->STEP_N++;
                p->CURRENT_Z += p->ZSTEP;
                ++steps;
            }
        }
        // device-driven controller: step attempts are enqueued in chunks and the
        // host reads the controller state once per chunk; attempts after the
        // plane (or after an error) are no-ops
        constexpr long kChunk = 32;
        for (;;) {
            if (p->CURRENT_Z >= Z[p->n_Z]) {
                do_dumps(p, &ds);
                if (p->n_Z == p->s.NZ - 1) return 0;
                p->n_Z++;
            }
            if (max_steps > 0 && steps >= max_steps) return 1;
            if (p->n_Z == 0) throw gs::Error(GS_E_PHYSICS, "CURRENT_Z is below the first dump plane");
            arm_ctl(p);
            for (;;) {
                const long want = max_steps > 0 ? std::min(kChunk, max_steps - steps) : kChunk;
                for (long k = 0; k < want; ++k) launch_slot(p);
                pull_ctl(p);
                const gs::DevCtl& h = *p->ctl_host;
                steps += (long)h.accepted;
                if (h.state == 2) ctl_error(p, h.err);
                if (h.accepted > 0) {
                    // re-arm keeps the plane; clears the accepted counter
                    if (h.state == 1 || (max_steps > 0 && steps >= max_steps)) break;
                    arm_ctl(p);
                }
            }
        }
    });
}

int gs_march_many(gs_plan* const* plans, int n, const char* const* outdirs, int* rcs)
{
    return guarded([&]() {
        if (!plans || n <= 0) return fail(GS_E_ARG, "gs_march_many: no plans");
        for (int i = 0; i < n; ++i) {
            if (!plans[i]) return fail(GS_E_ARG, "gs_march_many: null plan");
            for (int j = 0; j < i; ++j)
                if (plans[j] == plans[i]) return fail(GS_E_ARG, "gs_march_many: a plan is listed twice");
        }
        BatchMarch b(plans, n);
        for (int i = 0; i < n; ++i) plans[i]->batch_nodes = (double)b.graph_nodes;
        std::vector<DumpSet> ds(n);
        std::vector<char> active(n, 1);
        int left = n;
        // the main loop's plane bookkeeping (solver_main.cpp:206-212) of run i; false once it is done
        auto at_plane = [&](int i) {
            gs_plan* p = plans[i];
            const auto& Z = p->s.ZNET;
            if (p->CURRENT_Z >= Z[p->n_Z]) {
                do_dumps(p, &ds[i]);
                if (p->n_Z == p->s.NZ - 1) {
                    active[i] = 0;
                    --left;
                    if (rcs) rcs[i] = 0;
                    return false;
                }
                p->n_Z++;
            }
            if (p->n_Z == 0) throw gs::Error(GS_E_PHYSICS, "CURRENT_Z is below the first dump plane");
            fill_ctl(p, b.host[i]);
            return true;
        };
        for (int i = 0; i < n; ++i) {
            if (rcs) rcs[i] = 1;
            sync(plans[i]);
            open_dumps(plans[i], outdirs && outdirs[i] ? outdirs[i] : ".", &ds[i]);
        }
        for (int i = 0; i < n; ++i) {
            if (!at_plane(i)) b.host[i].state = 1;     // a run already at its last plane stays idle
        }
        constexpr int kChunk = 32;
        while (left > 0) {
            b.send();
            for (int k = 0; k < kChunk; ++k) b.launch();
            b.receive();
            for (int i = 0; i < n; ++i) {
                if (!active[i]) continue;
                const gs::DevCtl h = b.host[i];
                pull_state(plans[i], h);
                if (h.state == 2) ctl_error(plans[i], h.err);
                if (h.state == 1 && !at_plane(i)) b.host[i].state = 1;
            }
        }
        return 0;
    });
}

int gs_line_diagnostics(gs_plan* p, double* flux, double* maxI, double* rhomax)
{
    return guarded([&]() {
        if (!p) return fail(GS_E_ARG, "null plan");
        const int nr = p->s.NR;
        double* d = p->diag.as<double>();
        LCK(gs::launch_line_diag(p->prec, p->s.NT, p->field.p, nr, nullptr, p->s.TSTEP, p->nl, d, d + p->rp, d + 2 * p->rp,
                                 nullptr, nullptr, p->st));
        if (flux) HCK(hipMemcpyAsync(flux, d, nr * 8, hipMemcpyDeviceToHost, p->st));
        if (maxI) HCK(hipMemcpyAsync(maxI, d + p->rp, nr * 8, hipMemcpyDeviceToHost, p->st));
        if (rhomax) HCK(hipMemcpyAsync(rhomax, d + 2 * p->rp, nr * 8, hipMemcpyDeviceToHost, p->st));
        check_status(p, nullptr);
        return 0;
    });
}

int gs_fft_lines(gs_plan* p, const double* in, double* out, int n_lines, int sign)
{
    return guarded([&]() {
        if (!p || !in || !out || n_lines < 1) return fail(GS_E_ARG, "bad argument");
        const size_t n = (size_t)n_lines * p->s.NT;
        DevBuf d64, din, dout;
        upload(d64, in, n * 16);
        din.alloc(n * p->relem);
        dout.alloc(n * p->relem);
        LCK(gs::launch_convert(p->prec, d64.as<double>(), din.p, n, 1, p->st));
        LCK(gs::launch_fft_lines(p->prec, p->s.NT, din.p, dout.p, n_lines, sign, 1.0, nullptr, p->tw.p, p->st));
        LCK(gs::launch_convert(p->prec, d64.as<double>(), dout.p, n, 0, p->st));
        HCK(hipMemcpyAsync(out, d64.p, n * 16, hipMemcpyDeviceToHost, p->st));
        sync(p);
        return 0;
    });
}

int gs_fixed_steps(gs_plan* p, double dz, int n, double* ms4)
{
    return guarded([&]() {
        if (!p) return fail(GS_E_ARG, "null plan");
        if (n <= 0) return fail(GS_E_ARG, "gs_fixed_steps: n must be positive");
        p->ls.dz = dz;
        if (!ms4) {
            // no timings asked for: the kernels alone, no event markers between them
            for (int k = 0; k < n; ++k) {
                LCK(gs::launch_nl_rk4(p->prec, p->s.NT, p->field.p, p->b1.p, dz, p->nl, p->linemax.as<double>(), p->st));
                LCK(gs::launch_reduce_max(p->s.NR, p->linemax.as<double>(), p->status.as<gs::DevStatus>(), p->st));
                p->ls.seed = (uint32_t)(p->STEP_N + k);
                enqueue_linear(p, p->ls, nullptr);
            }
            check_status(p, nullptr);
            for (int k = 0; k < n; ++k) p->CURRENT_Z += dz;
            p->ZSTEP = dz;
            p->STEP_N += n;
            return 0;
        }
        std::vector<hipEvent_t> ev((size_t)5 * n, nullptr);
        struct EvGuard {
            std::vector<hipEvent_t>& v;
            ~EvGuard() { for (auto e : v) if (e) (void)hipEventDestroy(e); }
        } guard{ev};
        for (auto& e : ev) HCK(hipEventCreate(&e));
        for (int k = 0; k < n; ++k) {
            hipEvent_t* e = &ev[(size_t)5 * k];
            HCK(hipEventRecord(e[0], p->st));
            LCK(gs::launch_nl_rk4(p->prec, p->s.NT, p->field.p, p->b1.p, dz, p->nl, p->linemax.as<double>(), p->st));
            HCK(hipEventRecord(e[1], p->st));
            LCK(gs::launch_reduce_max(p->s.NR, p->linemax.as<double>(), p->status.as<gs::DevStatus>(), p->st));
            HCK(hipEventRecord(e[2], p->st));
            p->ls.seed = (uint32_t)(p->STEP_N + k);
            enqueue_linear(p, p->ls, e[3]);
            HCK(hipEventRecord(e[4], p->st));
        }
        check_status(p, nullptr);     // synchronises; NaN / blast flags of any of the steps
        double t[4] = {0, 0, 0, 0};
        for (int k = 0; k < n; ++k) {
            const hipEvent_t* e = &ev[(size_t)5 * k];
            float a = 0, b = 0, c = 0, d = 0;
            (void)hipEventElapsedTime(&a, e[0], e[1]);
            (void)hipEventElapsedTime(&d, e[1], e[2]);
            (void)hipEventElapsedTime(&b, e[2], e[3]);
            (void)hipEventElapsedTime(&c, e[3], e[4]);
            t[0] += a; t[1] += b; t[2] += c; t[3] += d;
        }
        for (int i = 0; i < 4; ++i) p->timings[i] = t[i] / n;
        if (ms4)
            for (int i = 0; i < 4; ++i) ms4[i] = p->timings[i];
        for (int k = 0; k < n; ++k) p->CURRENT_Z += dz;
        p->ZSTEP = dz;
        p->STEP_N += n;
        return 0;
    });
}

int gs_adjoint_forward(gs_plan* p, int n, const double* dz_rk, const double* dz_lin)
{
    return guarded([&]() {
        if (!p || !dz_rk || !dz_lin) return fail(GS_E_ARG, "null argument");
        if (n <= 0) return fail(GS_E_ARG, "gs_adjoint_forward: n must be positive");
        for (int k = 0; k < n; ++k)
            if (!(dz_rk[k] > 0) || !(dz_lin[k] > 0)) return fail(GS_E_ARG, "gs_adjoint_forward: step sizes must be positive");
        const size_t fb = p->field_bytes();
        const int rp = p->rp, nr = p->s.NR, nt = p->s.NT;
        p->adj_rk.clear();
        p->adj_lin.clear();
        // one segment when every step's FIELD fits the budget; otherwise segments
        // of K ~ sqrt(2n) steps (memory ceil(n/K) x 2 + K fields, one recomputation)
        double budget = p->adj_budget;
        if (budget < 0) {
            size_t dev_free = 0, dev_total = 0;
            HCK(hipMemGetInfo(&dev_free, &dev_total));
            budget = 0.6 * (double)(dev_free + p->ckpt.n);     // the old checkpoints are freed first
        }
        int K = n;
        if ((double)n * (double)fb > budget) K = std::max(1, (int)std::ceil(std::sqrt(2.0 * n)));
        const int nseg = (n + K - 1) / K;
        if (p->ckpt.n != fb * (size_t)K) {
            p->ckpt.alloc(0);
            p->ckpt.alloc(fb * (size_t)K);
        }
        if (K < n) {
            if (p->snaps.n != 2 * fb * (size_t)nseg) p->snaps.alloc(2 * fb * (size_t)nseg);
            for (DevBuf* b : {&p->rfield, &p->rb1, &p->rfk})
                if (b->n != fb) {
                    b->alloc(fb);
                    HCK(hipMemset(b->p, 0, fb));
                }
        } else {
            p->snaps.alloc(0);
        }
        p->adj_seg = K;
        if (!p->adjmat.p) {
            // the transposed operator A^T = diag(1/m1) C^T diag(m1) for the plan's GEMM:
            // FP64 / FP32 take their operand transposed, so they get row-major A;
            // mixed gets the split-BF16 fragments of A^T
            std::vector<double> A((size_t)rp * rp, 0.0);
            for (int i = 0; i < nr; ++i)
                for (int j = 0; j < nr; ++j) A[(size_t)i * rp + j] = p->s.qm1[i] * p->s.qC[(size_t)i * nr + j] / p->s.qm1[j];
            if (p->prec == gs::kFP64) {
                upload(p->adjmat, A.data(), A.size() * 8);
            } else if (p->prec == gs::kFP32) {
                std::vector<float> f(A.begin(), A.end());
                upload(p->adjmat, f.data(), f.size() * 4);
            } else {
                std::vector<double> At((size_t)rp * rp);
                for (int i = 0; i < rp; ++i)
                    for (int j = 0; j < rp; ++j) At[(size_t)i * rp + j] = A[(size_t)j * rp + i];
                upload_bf16_fragments(p->adjmat, At, rp);
            }
            for (DevBuf* b : {&p->gA, &p->gB, &p->gT1, &p->gT2}) {
                b->alloc(fb);
                HCK(hipMemset(b->p, 0, fb));
            }
            p->stage.alloc(3 * fb);
        }
        const void* A = p->prec == gs::kFP64 ? p->at64.p : p->amat.p;
        gs::LstepArgs ls = p->ls;
        // the march's transforms (column list, splits and tile shape of gs_step);
        // the first QDHT back clears FIELD's skipped columns when they are not zero
        const bool dirty0 = p->dead_dirty;
        const gs::QCols qf = p->qcols(false, false), qb0 = p->qcols(dirty0, true), qb = p->qcols(false, true);
        for (int k = 0; k < n; ++k) {
            if (K == n) {
                HCK(hipMemcpyAsync((char*)p->ckpt.p + (size_t)k * fb, p->field.p, fb, hipMemcpyDeviceToDevice, p->st));
            } else if (k % K == 0) {
                char* sn = (char*)p->snaps.p + (size_t)(k / K) * 2 * fb;
                HCK(hipMemcpyAsync(sn, p->field.p, fb, hipMemcpyDeviceToDevice, p->st));
                HCK(hipMemcpyAsync(sn + fb, p->fk.p, fb, hipMemcpyDeviceToDevice, p->st));
            }
            LCK(gs::launch_nl_rk4(p->prec, nt, p->field.p, p->b1.p, dz_rk[k], p->nl, p->linemax.as<double>(), p->st));
            ls.dz = dz_lin[k];
            ls.seed = (uint32_t)(p->STEP_N + k);
            LCK(gs::launch_qdht(p->prec, rp, nt, A, p->b1.p, nullptr, &ls, p->st, nullptr, &qf));
            LCK(gs::launch_qdht(p->prec, rp, nt, A, p->fk.p, p->field.p, nullptr, p->st, nullptr, k ? &qb : &qb0));
        }
        check_status(p, nullptr);
        p->adj_rk.assign(dz_rk, dz_rk + n);
        p->adj_seed0 = p->STEP_N;
        p->adj_dirty0 = dirty0;
        p->dead_dirty = false;
        p->adj_lin.assign(dz_lin, dz_lin + n);
        for (int k = 0; k < n; ++k) p->CURRENT_Z += dz_lin[k];
        p->ZSTEP = dz_lin[n - 1];
        p->STEP_N += n;
        return 0;
    });
}

int gs_adjoint_backward(gs_plan* p, const double* gfield, const double* gb2, int tie_b2, double* gfield0, double* gb2_0)
{
    return guarded([&]() {
        if (!p || !gfield || !gfield0) return fail(GS_E_ARG, "null argument");
        const int n = (int)p->adj_rk.size();
        if (n == 0) return fail(GS_E_ARG, "gs_adjoint_backward: no steps recorded (call gs_adjoint_forward first)");
        const size_t fb = p->field_bytes();
        const int rp = p->rp, nt = p->s.NT, K = p->adj_seg;
        const gs::Prec pa = p->prec;     // the transposed GEMMs run in the plan's GEMM precision
        const void* A = p->prec == gs::kFP64 ? p->at64.p : p->amat.p;
        upload_field64(p, p->gA, gfield);
        if (gb2) upload_field64(p, p->gB, gb2);
        else HCK(hipMemsetAsync(p->gB.p, 0, fb, p->st));
        gs::LstepArgs la = p->ls;
        gs::LstepArgs lr = p->ls;        // the replay's linear step updates the replay buffers
        lr.fk = p->rfk.p;
        // the forward's transforms for the replay; the transposed ones skip the same
        // columns: the linear step masks them (its transpose drops their cotangent),
        // so the cotangents of skipped columns of FIELD_k and B1 are never read and
        // stay at the zeros set here
        const gs::QCols qf = p->qcols(false, false), qb0 = p->qcols(p->adj_dirty0, true), qb = p->qcols(false, true);
        HCK(hipMemsetAsync(p->gT1.p, 0, fb, p->st));
        for (int k1 = n; k1 > 0;) {
            const int k0 = (k1 - 1) / K * K;
            if (K < n) {
                // recompute the segment's FIELDs from its snapshot (same kernels, same order)
                const char* sn = (const char*)p->snaps.p + (size_t)(k0 / K) * 2 * fb;
                HCK(hipMemcpyAsync(p->rfield.p, sn, fb, hipMemcpyDeviceToDevice, p->st));
                HCK(hipMemcpyAsync(p->rfk.p, sn + fb, fb, hipMemcpyDeviceToDevice, p->st));
                for (int k = k0; k < k1; ++k) {
                    HCK(hipMemcpyAsync((char*)p->ckpt.p + (size_t)(k - k0) * fb, p->rfield.p, fb, hipMemcpyDeviceToDevice, p->st));
                    if (k + 1 == k1) break;    // the segment's last FIELD is all the reverse sweep needs
                    LCK(gs::launch_nl_rk4(p->prec, nt, p->rfield.p, p->rb1.p, p->adj_rk[k], p->nl, p->linemax.as<double>(), p->st));
                    lr.dz = p->adj_lin[k];
                    lr.seed = (uint32_t)(p->adj_seed0 + k);
                    LCK(gs::launch_qdht(p->prec, rp, nt, A, p->rb1.p, nullptr, &lr, p->st, nullptr, &qf));
                    LCK(gs::launch_qdht(p->prec, rp, nt, A, p->rfk.p, p->rfield.p, nullptr, p->st, nullptr, k ? &qb : &qb0));
                }
            }
            for (int k = k1 - 1; k >= k0; --k) {
                // FIELD_{k+1} = Q FIELD_k(k-space): cotangent of FIELD_k(k-space)
                LCK(gs::launch_qdht(pa, rp, nt, p->adjmat.p, p->gA.p, p->gT1.p, nullptr, p->st, nullptr, &qb));
                la.dz = p->adj_lin[k];
                LCK(gs::launch_lstep_adj(p->prec, rp, nt, p->gT1.p, p->gB.p, p->gT2.p, la, p->st));
                // B1(k-space) = Q B1
                LCK(gs::launch_qdht(pa, rp, nt, p->adjmat.p, p->gT2.p, p->gT1.p, nullptr, p->st, nullptr, &qf));
                LCK(gs::launch_nl_rk4_vjp(p->prec, nt, (const char*)p->ckpt.p + (size_t)(k - k0) * fb, p->gT1.p, p->gA.p,
                                          p->stage.p, (size_t)rp * nt, p->adj_rk[k], p->nl, p->st));
            }
            k1 = k0;
        }
        check_status(p, nullptr);
        if (gb2_0) download_field64(p, p->gB.p, gb2_0, p->s.NR);
        if (tie_b2) {
            // initial BIGBUFFER2 = QDHT(FIELD): its cotangent adds Q^T gB to FIELD's
            LCK(gs::launch_qdht(pa, rp, nt, p->adjmat.p, p->gB.p, p->gT1.p, nullptr, p->st));
            LCK(gs::launch_axpy(p->prec, p->gA.p, p->gT1.p, (size_t)rp * nt, p->st));
        }
        download_field64(p, p->gA.p, gfield0, p->s.NR);
        return 0;
    });
}

int gs_last_timings(gs_plan* p, double* ms4)
{
    if (!p || !ms4) return fail(GS_E_ARG, "null argument");
    for (int i = 0; i < 4; ++i) ms4[i] = p->timings[i];
    return 0;
}

}  // extern "C"
