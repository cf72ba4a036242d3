// gs_adjoint.hip -- reverse-mode derivative of the fixed-step z-march: the
// adjoint of the RK4 NL step per radial line, and the transposed linear step.
// The reference has no adjoint; these kernels are the exact transpose of the
// forward scheme of gs_line.hip / gs_qdht.hip (calculate_NLresponse,
// propagator.cu:629-783; the plasma Heun recurrence, ionization.cpp:9-95; the
// RK4 of propagator.cu:376-425; calculate_Lstep, propagator.cu:570-625) with
// the step sizes held fixed.  The QDHT transpose diag(1/m1) C^T diag(m1) runs
// on the forward GEMM kernels with its own operator.
//
// Cotangent of a complex value z of a real objective L: dL/dRe z + i dL/dIm z;
// a complex-linear map y = M x has g_x = M^H g_y.  Per RHS evaluation
// N = F(loss E) + PF F(rho E) - i K F(y E), E = F^-1(x)/T, the adjoint
//   - recomputes E, I, W(I), the plasma density rho (same Newton scan) and the
//     Raman-weighted intensity y,
//   - takes the unnormalised inverse FFTs of lam, conj(PF) lam and i K lam
//     (the cotangents of the three time signals),
//   - runs the transposed Raman convolution (the forward closed-form scan on
//     the time-reversed signal) and the transposed plasma recurrence (an
//     affine scan on the time-reversed cotangents),
//   - and returns F(gE)/T.
// Jumps of the forward (loss gate, W cut-offs, the clamp rho <= N) have zero
// derivative almost everywhere and contribute nothing.
#include "gs_line.h"

namespace gs {

// W(I) as photoionization() and dW/dI: ln W is linear in ln I between table
// nodes, so dW/dI = W (b - a) / (step I)
template <class W>
__device__ __forceinline__ W photoionization_d(W I, const double* __restrict__ tab, int* flags, W* dW)
{
    *dW = 0;
    if (I < (W)1e-3) return 0;
    const W lnI = log(I) + (W)36.841361487904734;
    if (lnI != lnI) { *flags |= kFlagNaN; return 0; }
    const W u = (lnI - (W)29.933606208922594) / (W)0.003304003148325;
    const int n = (int)floor(u);
    if (n > 8190) { *flags |= kFlagBlast; return 0; }
    if (n < 0) return 0;
    const W lnIn = (W)29.933606208922594 + (W)n * (W)0.003304003148325;
    const W a = (W)tab[n], b = (W)tab[n + 1];
    const W slope = (b - a) / (W)0.003304003148325;
    const W w = exp(a + (lnI - lnIn) * slope);
    *dW = w * slope / I;
    return w;
}

// partial derivatives of one Heun step (heun_step) from rho = x with the rates
// of the previous sample (Wp, cp) and of this sample (Wn, cn); all zero where
// the clamp at N is active
template <class S>
struct HeunD {
    S d, dWp, dcp, dWn, dcn;
};
template <class S>
__device__ __forceinline__ HeunD<S> heun_partials(S x, S Wp, S cp, S Wn, S cn, S dt, S N, S invN, S invtau)
{
    const bool in0 = !(x < 0 || x > N);
    const S k1 = dt * ((x > N ? (S)0 : Wp) + (in0 ? cp * x * ((S)1 - x * invN) : (S)0) - (x > 0 ? x * invtau : (S)0));
    const S xm = x + k1;
    const bool in1 = !(xm < 0 || xm > N);
    const S k2 = dt * ((xm > N ? (S)0 : Wn) + (in1 ? cn * xm * ((S)1 - xm * invN) : (S)0) - (xm > 0 ? xm * invtau : (S)0));
    HeunD<S> r{0, 0, 0, 0, 0};
    if (x + (S)0.5 * (k1 + k2) > N) return r;
    const S dk1 = dt * ((in0 ? cp * ((S)1 - (S)2 * x * invN) : (S)0) - (x >= 0 ? invtau : (S)0));
    const S q = dt * ((in1 ? cn * ((S)1 - (S)2 * xm * invN) : (S)0) - (xm >= 0 ? invtau : (S)0));
    const S h = (S)0.5 * dt;
    r.d = (S)1 + (S)0.5 * (dk1 + q * ((S)1 + dk1));
    r.dWp = x <= N ? h * ((S)1 + q) : (S)0;
    r.dcp = in0 ? h * x * ((S)1 - x * invN) * ((S)1 + q) : (S)0;
    r.dWn = xm <= N ? h : (S)0;
    r.dcn = in1 ? h * xm * ((S)1 - xm * invN) : (S)0;
    return r;
}

// time reversal of chunk-layout sequences across the line: v[n] <- v[T-1-n],
// staged through padded LDS buffers (one value per C2 slot); ends with a
// barrier so the buffers can be reused at once
template <int T, int E, class V, class B>
__device__ __forceinline__ void reverse_line(V (&v)[E], B* buf)
{
    static_assert(sizeof(V) <= sizeof(B), "slot too small");
    const int n0 = threadIdx.x * E;
#pragma unroll
    for (int e = 0; e < E; ++e) __builtin_memcpy(&buf[pidx(T - 1 - (n0 + e), E)], &v[e], sizeof(V));
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) __builtin_memcpy(&v[e], &buf[pidx(n0 + e, E)], sizeof(V));
    __syncthreads();
}
template <int T, int E, class V, class B>
__device__ __forceinline__ void reverse_line2(V (&v)[E], V (&u)[E], B* buf0, B* buf1)
{
    static_assert(sizeof(V) <= sizeof(B), "slot too small");
    const int n0 = threadIdx.x * E;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = pidx(T - 1 - (n0 + e), E);
        __builtin_memcpy(&buf0[i], &v[e], sizeof(V));
        __builtin_memcpy(&buf1[i], &u[e], sizeof(V));
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = pidx(n0 + e, E);
        __builtin_memcpy(&v[e], &buf0[i], sizeof(V));
        __builtin_memcpy(&u[e], &buf1[i], sizeof(V));
    }
    __syncthreads();
}

// Raman tail of the circular convolution, out[n] = fR M Im(s_n (1 - z^T) + z^{n+1} s_{T-1}),
// s_n = z s_{n-1} + u_n (the R * u part of nl_rhs without the (1-fR) u term)
template <class R, int T, int NT, int E>
__device__ __forceinline__ void raman_tail(const R (&u)[E], R (&out)[E], AffC<R>* scan, const NLArgs& a)
{
    const int n0 = threadIdx.x * E;
    const C2<R>* zp = (const C2<R>*)a.zpow;
    const C2<R> z = zp[1];
    C2<R> loc = {0, 0};
#pragma unroll
    for (int e = 0; e < E; ++e) {
        loc = cmul(z, loc);
        loc.x += u[e];
    }
    const C2<R> zE = zp[E];
    AffC<R> tot;
    const AffC<R> pre = block_scan_excl<NT>(AffC<R>{zE.x, zE.y, loc.x, loc.y}, scan, &tot);
    const C2<R> slast = {tot.br, tot.bi};
    const C2<R> omz = {(R)a.onemzT_re, (R)a.onemzT_im};
    C2<R> sv = {pre.br, pre.bi};
    C2<R> zq = zp[n0 + 1];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        sv = cmul(z, sv);
        sv.x += u[e];
        out[e] = (R)a.frM * (sv.x * omz.y + sv.y * omz.x + zq.x * slast.y + zq.y * slast.x);
        zq = cmul(zq, z);
    }
}

// Transposed plasma recurrence.  Forward: rho[n] = H(rho[n-1]; W[n-1], c[n-1], W[n], c[n])
// for n >= 1 (the rates of sample 0 never enter: E0 = 0).  With grho the direct
// cotangents of rho, mu[n] = grho[n] + d[n+1] mu[n+1] (d[n] = dH/drho of step n)
// is a suffix recurrence: it runs as a prefix affine scan on the time-reversed
// line.  Returns the cotangents of the rates W and c per sample.
template <class S, int T, int NT, int E, class Wt, class B>
__device__ __forceinline__ void plasma_vjp(const Wt (&Wv)[E], const Wt (&cv)[E], const S (&rho)[E], const S (&grho)[E],
                                           S (&gW)[E], S (&gc)[E], const NLArgs& a, const PlasmaScratch& ps, B* buf0,
                                           B* buf1)
{
    const int tid = threadIdx.x;
    const S dt = (S)a.dt, N = (S)a.neutral, invN = (S)a.inv_neutral, invtau = (S)a.inv_recomb;
    // previous sample's density and rates at the chunk boundary
    ps.W[tid] = (double)Wv[E - 1];
    ps.c[tid] = (double)cv[E - 1];
    ps.rho[tid] = (double)rho[E - 1];
    __syncthreads();
    const S Wm1 = tid > 0 ? (S)ps.W[tid - 1] : (S)0, cm1 = tid > 0 ? (S)ps.c[tid - 1] : (S)0;
    const S rm1 = tid > 0 ? (S)ps.rho[tid - 1] : (S)0;
    __syncthreads();
    HeunD<S> hd[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = tid * E + e;
        if (n == 0) {
            hd[e] = HeunD<S>{0, 0, 0, 0, 0};
            continue;
        }
        const S x = e == 0 ? rm1 : rho[e - 1];
        const bool first = n == 1;     // the step into n = 1 starts from E0 = 0
        const S Wp = first ? (S)0 : (e == 0 ? Wm1 : (S)Wv[e - 1]);
        const S cp = first ? (S)0 : (e == 0 ? cm1 : (S)cv[e - 1]);
        hd[e] = heun_partials<S>(x, Wp, cp, (S)Wv[e], (S)cv[e], dt, N, invN, invtau);
    }
    // multiplier of mu[n+1] in mu[n]: d[n+1], from the next sample (0 past the end)
    ps.W[tid] = (double)hd[0].d;
    __syncthreads();
    S dn[E];
#pragma unroll
    for (int e = 0; e < E - 1; ++e) dn[e] = hd[e + 1].d;
    dn[E - 1] = tid + 1 < NT ? (S)ps.W[tid + 1] : (S)0;
    // reversed index j = T-1-n, nu[j] = mu[T-1-j]: nu[j] = b[j] + m[j] nu[j-1] with
    // b[j] = grho[T-1-j], m[j] = dn[T-1-j] (m = 0 at j = 0), a prefix recurrence
    S b[E], m[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        b[e] = grho[e];
        m[e] = dn[e];
    }
    reverse_line2<T, E>(b, m, buf0, buf1);
    S A = 1, Bc = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        A = m[e] * A;
        Bc = m[e] * Bc + b[e];
    }
    AffR<S> tot;
    const AffR<S> pre = block_scan_excl<NT>(AffR<S>{A, Bc}, (AffR<S>*)ps.scan, &tot);
    S nu = pre.b;
    S mu[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        nu = b[e] + m[e] * nu;
        mu[e] = nu;
    }
    reverse_line<T, E>(mu, buf0);
    // cotangents of the rates: step n feeds W[n], c[n] and (n >= 2) W[n-1], c[n-1]
#pragma unroll
    for (int e = 0; e < E; ++e) {
        gW[e] = mu[e] * hd[e].dWn;
        gc[e] = mu[e] * hd[e].dcn;
    }
#pragma unroll
    for (int e = 1; e < E; ++e) {
        const int n = tid * E + e;
        if (n >= 2) {
            gW[e - 1] += mu[e] * hd[e].dWp;
            gc[e - 1] += mu[e] * hd[e].dcp;
        }
    }
    // e = 0: its previous sample is the last of the previous thread
    ps.W[tid] = (double)(tid > 0 ? mu[0] * hd[0].dWp : (S)0);
    ps.c[tid] = (double)(tid > 0 ? mu[0] * hd[0].dcp : (S)0);
    __syncthreads();
    if (tid + 1 < NT && (tid + 1) * E >= 2) {
        gW[E - 1] += (S)ps.W[tid + 1];
        gc[E - 1] += (S)ps.c[tid + 1];
    }
    __syncthreads();
}

// gx += J^H lam for the RHS of one line at the spectrum x (all strided layout)
template <class Pol, int T, int E>
__device__ __forceinline__ void nl_rhs_vjp(const C2<typename Pol::R> (&x)[E], const C2<typename Pol::R> (&lam)[E],
                                           C2<typename Pol::R> (&gx)[E], LineSmem<Pol, T, E>& sm, const NLArgs& a,
                                           int* flags)
{
    using R = typename Pol::R;
    using S = typename Pol::S;
    using W = typename Pol::W;
    constexpr int NT = Shape<Pol, T, E>::NT;
    const int tid = threadIdx.x, n0 = tid * E;
    const C2<R>* tw = (const C2<R>*)a.tw;

    // ---- forward quantities of the line
    C2<R> Ev[E];
    fft_inv<R, T, NT, E>(x, Ev, sm.b0, sm.b1, tw);
    W Ws[E], dWs[E], cs[E], dcs[E];
    R lossf[E], dlossf[E];
    S rho[E];
    bool anyw = false, anyloss = false, has_rho = false;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        Ev[e] = cscale(Ev[e], (R)a.inv_nt);
        const W I = (W)Ev[e].x * (W)Ev[e].x + (W)Ev[e].y * (W)Ev[e].y;
        Ws[e] = photoionization_d<W>(I, a.ionrate, flags, &dWs[e]);
        anyw |= Ws[e] > (W)0;
        const W den = (W)a.U + (W)a.pond * I;
        cs[e] = (W)a.sigma / den * I;
        dcs[e] = (W)a.sigma * (W)a.U / (den * den);
        const int n = n0 + e;
        const bool on = n >= 1 && n <= T - 2 && I > (W)29.933606208922594;
        lossf[e] = on ? (R)(-(W)0.5 * Ws[e] * den / I) : (R)0;
        dlossf[e] = on ? (R)(-(W)0.5 * (dWs[e] * den / I - Ws[e] * (W)a.U / (I * I))) : (R)0;
        anyloss |= lossf[e] != (R)0;
        rho[e] = 0;
    }
    const bool ionizing = __syncthreads_or(anyw) || a.rho_init != 0.0;
    if (ionizing) has_rho = plasma_scan2<S, E, NT>(Ws, cs, rho, a, scratch_of(sm), flags);
    const bool lossy = ionizing && __syncthreads_or(anyloss);
    R yv[E];
    if (a.use_kerr) {
        R Iv[E];
#pragma unroll
        for (int e = 0; e < E; ++e) Iv[e] = Ev[e].x * Ev[e].x + Ev[e].y * Ev[e].y;
        if (a.use_raman) {
            raman_tail<R, T, NT, E>(Iv, yv, (AffC<R>*)sm.scan, a);
#pragma unroll
            for (int e = 0; e < E; ++e) yv[e] += (R)a.one_m_fr * Iv[e];
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) yv[e] = Iv[e];
        }
    }

    // ---- cotangents of the time signals and of E, I
    C2<R> gE[E];
    R gI[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        gE[e] = C2<R>{0, 0};
        gI[e] = 0;
    }
    C2<R> sp[E], tl[E];
    if (lossy) {
        fft_inv<R, T, NT, E>(lam, tl, sm.b0, sm.b1, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            gE[e] = cadd(gE[e], cscale(tl[e], lossf[e]));
            gI[e] += (Ev[e].x * tl[e].x + Ev[e].y * tl[e].y) * dlossf[e];
        }
    }
    S grho[E];
    if (has_rho) {
        const C2<R>* pf = (const C2<R>*)a.pf;
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const C2<R> f = pf[tid + r * NT];
            sp[r] = cmul(C2<R>{f.x, -f.y}, lam[r]);
        }
        fft_inv<R, T, NT, E>(sp, tl, sm.b0, sm.b1, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            gE[e] = cadd(gE[e], cscale(tl[e], (R)rho[e]));
            grho[e] = (S)(Ev[e].x * tl[e].x + Ev[e].y * tl[e].y);
        }
    }
    if (a.use_kerr) {
        const R* kp = (const R*)a.kerr;
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const R k = kp[tid + r * NT];
            sp[r] = C2<R>{-k * lam[r].y, k * lam[r].x};     // i K lam
        }
        fft_inv<R, T, NT, E>(sp, tl, sm.b0, sm.b1, tw);
        R gy[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            gE[e] = cadd(gE[e], cscale(tl[e], yv[e]));
            gy[e] = Ev[e].x * tl[e].x + Ev[e].y * tl[e].y;
        }
        if (a.use_raman) {
            // transpose of the circular convolution: the same tail on the reversed signal
            R u[E], t[E];
#pragma unroll
            for (int e = 0; e < E; ++e) u[e] = gy[e];
            reverse_line<T, E>(u, sm.b0);
            raman_tail<R, T, NT, E>(u, t, (AffC<R>*)sm.scan, a);
            reverse_line<T, E>(t, sm.b0);
#pragma unroll
            for (int e = 0; e < E; ++e) gI[e] += (R)a.one_m_fr * gy[e] + t[e];
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) gI[e] += gy[e];
        }
    }
    if (has_rho) {
        S gW[E], gc[E];
        plasma_vjp<S, T, NT, E>(Ws, cs, rho, grho, gW, gc, a, scratch_of(sm), sm.b0, sm.b1);
#pragma unroll
        for (int e = 0; e < E; ++e) gI[e] += (R)(gW[e] * (S)dWs[e] + gc[e] * (S)dcs[e]);
    }
    // I = |E|^2, then E = IFFT(x)/T: gx += FFT(gE)/T
    R fac[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        gE[e] = cadd(gE[e], cscale(Ev[e], (R)2 * gI[e]));
        fac[e] = (R)a.inv_nt;
    }
    fft_fwd_acc<0, R, T, NT, E>(gE, fac, gx, sm.b0, sm.b1, tw, nullptr);
}

// Adjoint of the RK4 NL step of one line (k_nl_rk4): from the cotangent g of
// B1 = (k1 + 2 k2 + 2 k3 + k4)/6 at FIELD = A and step dz, gA = J^H g.
// The stage inputs X2 = A + dz/2 k1, X3 = A + dz/2 k2, X4 = A + dz k3 are
// recomputed and kept in global rows (stage, three planes of [rows][T]); each
// thread reads back only its own entries.
template <class Pol, int T, int E>
__global__ __launch_bounds__((Shape<Pol, T, E>::NT)) void k_nl_rk4_vjp(const C2<typename Pol::R>* __restrict__ field,
                                                                       const C2<typename Pol::R>* __restrict__ gb1,
                                                                       C2<typename Pol::R>* __restrict__ gfield,
                                                                       C2<typename Pol::R>* __restrict__ stage,
                                                                       size_t plane, double dz, NLArgs a)
{
    using R = typename Pol::R;
    constexpr int NT = Shape<Pol, T, E>::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    LineSmem<Pol, T, E>& sm = *reinterpret_cast<LineSmem<Pol, T, E>*>(smem_raw);
    const int tid = threadIdx.x;
    const size_t row = (size_t)blockIdx.x * T;
    const C2<R>* Ag = field + row;
    C2<R>* X2 = stage + row;
    C2<R>* X3 = X2 + plane;
    C2<R>* X4 = X3 + plane;
    const R h = (R)dz, hh = (R)(0.5 * dz);
    int flags = 0;

    C2<R> xin[E], k[E];
#pragma unroll
    for (int r = 0; r < E; ++r) xin[r] = Ag[tid + r * NT];
    nl_rhs<Pol, T, E>(xin, k, sm, a, &flags);
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const C2<R> av = Ag[tid + r * NT];
        xin[r] = cadd(av, cscale(k[r], hh));
        X2[tid + r * NT] = xin[r];
    }
    nl_rhs<Pol, T, E>(xin, k, sm, a, &flags);
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const C2<R> av = Ag[tid + r * NT];
        xin[r] = cadd(av, cscale(k[r], hh));
        X3[tid + r * NT] = xin[r];
    }
    nl_rhs<Pol, T, E>(xin, k, sm, a, &flags);
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const C2<R> av = Ag[tid + r * NT];
        X4[tid + r * NT] = cadd(av, cscale(k[r], h));
    }

    // reverse: lk is the cotangent of the stage output k_j, gx that of its input X_j
    C2<R> gA[E], lk[E], gx[E];
#pragma unroll
    for (int r = 0; r < E; ++r) gx[r] = gA[r] = C2<R>{0, 0};
    const C2<R>* grow = gb1 + row;
#pragma unroll
    for (int s = 3; s >= 0; --s) {
        const C2<R>* X = s == 3 ? X4 : (s == 2 ? X3 : (s == 1 ? X2 : Ag));
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const C2<R> g = grow[tid + r * NT];
            const R wg = (s == 1 || s == 2) ? (R)(1.0 / 3.0) : (R)(1.0 / 6.0);
            const R wx = s == 2 ? h : hh;      // X4 = A + h k3; X3, X2 = A + h/2 k
            lk[r] = s == 3 ? cscale(g, wg) : cadd(cscale(g, wg), cscale(gx[r], wx));
            xin[r] = X[tid + r * NT];
        }
#pragma unroll
        for (int r = 0; r < E; ++r) gx[r] = C2<R>{0, 0};
        nl_rhs_vjp<Pol, T, E>(xin, lk, gx, sm, a, &flags);
#pragma unroll
        for (int r = 0; r < E; ++r) gA[r] = s == 3 ? gx[r] : cadd(gA[r], gx[r]);
    }
#pragma unroll
    for (int r = 0; r < E; ++r) gfield[row + tid + r * NT] = gA[r];
    report_flags(a.status, flags);
}

// transpose of the linear step at one point: with gfk the cotangent of FIELD_k
// and gb2 that of BIGBUFFER2 after the step, gb2 <- conj(P) (gb2 + gfk) (the
// cotangent of BIGBUFFER2 before it) and gb1k <- dz g gb2, P = e^{dz(i rate + Im k)};
// masked points (g < 0: FIELD zeroed, BIGBUFFER2 untouched) pass gb2 through
template <class Pol>
__global__ void k_lstep_adj(const C2<typename Pol::R>* __restrict__ gfk, C2<typename Pol::R>* __restrict__ gb2,
                            C2<typename Pol::R>* __restrict__ gb1k, int nt, size_t n, LstepArgs la)
{
    using R = typename Pol::R;
    using P = typename Pol::P;
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const R g = ((const R*)la.gfac)[o];
    if (g < (R)0) {
        gb1k[o] = C2<R>{0, 0};
        return;
    }
    P sp, cp;
    lstep_phasor<Pol>((P)la.dz * ((const P*)la.rate)[o], &sp, &cp);
    const R sn = (R)sp, cs = (R)cp;
    const R amp = (R)lstep_amp<Pol>((int)(o % nt), la);
    const C2<R> v = cadd(gb2[o], gfk[o]);
    const C2<R> u = {amp * (cs * v.x + sn * v.y), amp * (cs * v.y - sn * v.x)};
    gb2[o] = u;
    gb1k[o] = cscale(u, (R)la.dz * g);
}

template <class R>
__global__ void k_axpy_c(C2<R>* __restrict__ y, const C2<R>* __restrict__ x, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = cadd(y[i], x[i]);
}

template <class Pol, int T>
static int nl_vjp_T(const void* field, const void* gb1, void* gfield, void* stage, size_t plane, double dz,
                    const NLArgs& a, hipStream_t st)
{
    using R = typename Pol::R;
    constexpr int E = DefaultE<Pol, T>::v;
    const size_t sm = sizeof(LineSmem<Pol, T, E>);
    auto kf = k_nl_rk4_vjp<Pol, T, E>;
    HIPCHK(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
    hipLaunchKernelGGL(kf, dim3(a.nr), dim3(Shape<Pol, T, E>::NT), sm, st, (const C2<R>*)field, (const C2<R>*)gb1,
                       (C2<R>*)gfield, (C2<R>*)stage, plane, dz, a);
    HIPCHK(hipGetLastError());
    return 0;
}

int launch_nl_rk4_vjp(Prec p, int nt, const void* field, const void* gb1, void* gfield, void* stage, size_t plane,
                      double dz, const NLArgs& a, hipStream_t st)
{
    if (plane < (size_t)a.nr * nt) return -1001;
    if (p == kFP64) GS_T_SWITCH(nt, (nl_vjp_T<PolF64, T_>(field, gb1, gfield, stage, plane, dz, a, st)));
    if (p == kFP32) GS_T_SWITCH(nt, (nl_vjp_T<PolF32, T_>(field, gb1, gfield, stage, plane, dz, a, st)));
    GS_T_SWITCH(nt, (nl_vjp_T<PolMix, T_>(field, gb1, gfield, stage, plane, dz, a, st)));
}

int launch_lstep_adj(Prec p, int rp, int nt, const void* gfk, void* gb2, void* gb1k, const LstepArgs& la, hipStream_t st)
{
    const size_t n = (size_t)rp * nt;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (p == kFP64)
        hipLaunchKernelGGL(k_lstep_adj<PolF64>, grid, dim3(256), 0, st, (const C2<double>*)gfk, (C2<double>*)gb2, (C2<double>*)gb1k, nt, n, la);
    else if (p == kFP32)
        hipLaunchKernelGGL(k_lstep_adj<PolF32>, grid, dim3(256), 0, st, (const C2<float>*)gfk, (C2<float>*)gb2, (C2<float>*)gb1k, nt, n, la);
    else
        hipLaunchKernelGGL(k_lstep_adj<PolMix>, grid, dim3(256), 0, st, (const C2<float>*)gfk, (C2<float>*)gb2, (C2<float>*)gb1k, nt, n, la);
    HIPCHK(hipGetLastError());
    return 0;
}

int launch_axpy(Prec p, void* y, const void* x, size_t n_complex, hipStream_t st)
{
    const dim3 grid((unsigned)((n_complex + 255) / 256));
    if (p == kFP64) hipLaunchKernelGGL(k_axpy_c<double>, grid, dim3(256), 0, st, (C2<double>*)y, (const C2<double>*)x, n_complex);
    else hipLaunchKernelGGL(k_axpy_c<float>, grid, dim3(256), 0, st, (C2<float>*)y, (const C2<float>*)x, n_complex);
    HIPCHK(hipGetLastError());
    return 0;
}

}  // namespace gs
