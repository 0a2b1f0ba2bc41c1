// CG and BiCGStab with every scalar on the device (the solvers of
// CG.tpp:14-113 for the non-Hermitian operators M and M^dag, and of
// BiCGStab.tpp:14-150 for any operator).
//
// Each recurrence scalar (alpha, beta, omega, chi, psi, phi, delta) lives in
// device memory, one complex value per slot (per RHS, or one for the whole
// batch in the reference's batch-global mode), and is formed by a one-thread-
// per-slot kernel from the inner products the previous kernels wrote; the
// vector updates read it from there and skip finished slots.  Nothing waits
// for the host inside an iteration: the host reads the done flags every
// NSLX_POLL iterations (default 16) and at maxit.  Finished slots are frozen
// by the masked updates, so the few iterations run past convergence before
// the next poll change nothing; the reported iteration count is the one at
// which the last slot converged, as the per-iteration host loop would report.
#include <algorithm>

#include "engine.hpp"

namespace nslx {
namespace {

struct Cz { double re, im; };
__device__ inline Cz ld(const double* p, int64_t i) { return {p[2 * i], p[2 * i + 1]}; }
__device__ inline void st(double* p, int64_t i, Cz v) { p[2 * i] = v.re; p[2 * i + 1] = v.im; }
__device__ inline Cz cmul(Cz a, Cz b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline Cz cdiv(Cz a, Cz b) {
    const double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// t2 = tol^2 |b|^2 (relative) or tol^2 (batch-global); done = |r|^2 <= t2 ? 0 : -1
__global__ void k_start(int64_t n, const double* bb, const double* rr, double tol2, int global, double* t2, int* done) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    t2[i] = global ? tol2 : tol2 * bb[2 * i];
    done[i] = rr[2 * i] <= t2[i] ? 0 : -1;
}

// CG: alpha = |r|^2 / <p, t>
__global__ void k_cg_alpha(int64_t n, const double* rr, const double* pt, double* al) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st(al, i, cdiv(Cz{rr[2 * i], 0.0}, ld(pt, i)));
}

// CG: beta = |r_new|^2 / |r|^2 for running slots (0 for finished ones); the
// slot finishes at iteration it once |r_new|^2 <= t2
__global__ void k_cg_beta(int64_t n, const double* rn, double* rr, const double* t2, double* be, int* done, int it) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (done[i] >= 0) { st(be, i, Cz{0.0, 0.0}); return; }
    st(be, i, Cz{rn[2 * i] / rr[2 * i], 0.0});
    rr[2 * i] = rn[2 * i];
    rr[2 * i + 1] = rn[2 * i + 1];
    if (rn[2 * i] <= t2[i]) done[i] = it;
}

// BiCGStab: phi = <y, As> / delta
__global__ void k_bi_phi(int64_t n, const double* yAs, const double* delta, double* phi) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st(phi, i, cdiv(ld(yAs, i), ld(delta, i)));
}

// omega = 1 / phi
__global__ void k_bi_om(int64_t n, const double* phi, double* om) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st(om, i, cdiv(Cz{1.0, 0.0}, ld(phi, i)));
}

// chi = <Aw, w> / |Aw|^2
__global__ void k_bi_chi(int64_t n, const double* aww, const double* awaw, double* chi) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st(chi, i, cdiv(ld(aww, i), Cz{awaw[2 * i], 0.0}));
}

// psi = -omega delta' / (delta chi); mpsi = -psi; pc = psi chi
__global__ void k_bi_psi(int64_t n, const double* om, const double* dp1, const double* delta, const double* chi,
                         double* mpsi, double* pc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Cz psi = cdiv(cmul(Cz{-1.0, 0.0}, cmul(ld(om, i), ld(dp1, i))), cmul(ld(delta, i), ld(chi, i)));
    st(mpsi, i, Cz{-psi.re, -psi.im});
    st(pc, i, cmul(psi, ld(chi, i)));
}

// running slots: phi = <y, As_new> / delta', delta = delta', |r|^2 recorded,
// finished at iteration it once |r|^2 <= t2
__global__ void k_bi_update(int64_t n, const double* yAs, const double* dp1, const double* rn, const double* t2,
                            double* phi, double* delta, double* rr, int* done, int it) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || done[i] >= 0) return;
    st(phi, i, cdiv(ld(yAs, i), ld(dp1, i)));
    st(delta, i, ld(dp1, i));
    rr[2 * i] = rn[2 * i];
    rr[2 * i + 1] = rn[2 * i + 1];
    if (rn[2 * i] <= t2[i]) done[i] = it;
}

inline unsigned g1(int64_t n) { return (unsigned)((n + 255) / 256); }

void copy_dev(void* dst, const void* src, size_t bytes, hipStream_t s) {
    NSLX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
}

}  // namespace

int solve_generic_dev(FermionMatrix& fm, int op, int method, const void* b, void* x, const void* x0, int64_t nb,
                      double tol, int tol_kind, int64_t maxit, nslx_solve_report* rep, hipStream_t s) {
    SolverWork& w = fm.sw;
    const int64_t len = fm.len();
    const size_t bytes = (size_t)nb * len * 16;
    const int global = tol_kind == NSLX_TOL_ABS_BATCH;
    const int64_t ns = global ? 1 : nb;
    const unsigned gs = g1(ns);
    w.ensure_scalars(nb, 1, s);
    w.r.ensure(bytes); w.p0.ensure(bytes); w.u.ensure(bytes);
    w.v1.ensure(bytes); w.v2.ensure(bytes); w.v3.ensure(bytes); w.v4.ensure(bytes);
    // 16 complex scalars per slot
    DevBuf<double> sc((size_t)32 * ns);
    DevBuf<double> t2((size_t)ns);
    double* S[16];
    for (int k = 0; k < 16; ++k) S[k] = sc.p + (size_t)2 * k * ns;
    double *rr = S[0], *bb = S[1], *d0 = S[2], *d1 = S[3], *al = S[4], *be = S[5], *phi = S[6], *delta = S[7],
           *om = S[8], *chi = S[9], *mpsi = S[10], *pc = S[11], *dp1 = S[12], *yAs = S[13];
    int* done = w.done.p;
    DevBuf<double> one((size_t)2 * ns);
    {
        std::vector<double> h(2 * ns, 0.0);
        for (int64_t i = 0; i < ns; ++i) h[2 * i] = 1.0;
        NSLX_HIP(hipMemcpyAsync(one.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
        NSLX_HIP(hipStreamSynchronize(s));
    }
    auto dot = [&](const void* a, const void* c, double* out) { dot_perb(1, a, c, nb, len, out, global, fm.vsc, s); };
    auto axpy = [&](const void* xx, const void* yy, void* zz, const double* cs, double sign, const int* dn) {
        axpy_perb(1, xx, yy, zz, cs, sign, nb, len, global, dn, s);
    };
    copy_dev(x, x0 ? x0 : b, bytes, s);
    dot(b, b, bb);
    // r = b - A x; stop test
    fm.apply(op, x, w.u.p, nb, NSLX_F64, s);
    axpy(b, w.u.p, w.r.p, one.p, -1.0, nullptr);
    dot(w.r.p, w.r.p, rr);
    hipLaunchKernelGGL(k_start, dim3(gs), dim3(256), 0, s, ns, bb, rr, tol * tol, global, t2.p, done);
    NSLX_LAUNCH_CHECK();

    const int poll = std::max(1, (int)env_double("NSLX_POLL", 16.0));
    std::vector<int> dh(ns);
    auto all_done = [&]() {
        NSLX_HIP(hipMemcpyAsync(dh.data(), done, ns * sizeof(int), hipMemcpyDeviceToHost, s));
        NSLX_HIP(hipStreamSynchronize(s));
        for (int v : dh) if (v < 0) return false;
        return true;
    };
    int64_t it = 0;
    bool finished = all_done();
    if (method == NSLX_CG) {
        void* p = w.p0.p;
        void* t = w.u.p;
        copy_dev(p, w.r.p, bytes, s);
        while (!finished && it < maxit) {
            ++it;
            fm.apply(op, p, t, nb, NSLX_F64, s);
            dot(p, t, d0);
            hipLaunchKernelGGL(k_cg_alpha, dim3(gs), dim3(256), 0, s, ns, rr, d0, al);
            NSLX_LAUNCH_CHECK();
            axpy(x, p, x, al, 1.0, done);
            axpy(w.r.p, t, w.r.p, al, -1.0, done);
            dot(w.r.p, w.r.p, d1);
            hipLaunchKernelGGL(k_cg_beta, dim3(gs), dim3(256), 0, s, ns, d1, rr, t2.p, be, done, (int)it);
            NSLX_LAUNCH_CHECK();
            axpy(w.r.p, p, p, be, 1.0, done);      // p = r + beta p for running slots
            if (it % poll == 0 || it == maxit) finished = all_done();
        }
    } else {
        // BiCGStab.tpp:38-117: y = s = r; As = A s; delta = <y,r>; phi = <y,As>/delta
        void* y = w.v1.p; void* sv = w.v2.p; void* As = w.v3.p; void* wv = w.v4.p; void* Aw = w.u.p;
        copy_dev(y, w.r.p, bytes, s);
        copy_dev(sv, w.r.p, bytes, s);
        fm.apply(op, sv, As, nb, NSLX_F64, s);
        dot(y, w.r.p, delta);
        dot(y, As, yAs);
        hipLaunchKernelGGL(k_bi_phi, dim3(gs), dim3(256), 0, s, ns, yAs, delta, phi);
        NSLX_LAUNCH_CHECK();
        while (!finished && it < maxit) {
            ++it;
            hipLaunchKernelGGL(k_bi_om, dim3(gs), dim3(256), 0, s, ns, phi, om);
            NSLX_LAUNCH_CHECK();
            axpy(w.r.p, As, wv, om, -1.0, done);          // w = r - omega As
            fm.apply(op, wv, Aw, nb, NSLX_F64, s);
            dot(Aw, wv, d0);
            dot(Aw, Aw, d1);
            hipLaunchKernelGGL(k_bi_chi, dim3(gs), dim3(256), 0, s, ns, d0, d1, chi);
            NSLX_LAUNCH_CHECK();
            axpy(wv, Aw, w.r.p, chi, -1.0, done);         // r = w - chi Aw
            axpy(x, sv, x, om, 1.0, done);                // x += omega s
            axpy(x, wv, x, chi, 1.0, done);               // x += chi w
            dot(y, w.r.p, dp1);
            hipLaunchKernelGGL(k_bi_psi, dim3(gs), dim3(256), 0, s, ns, om, dp1, delta, chi, mpsi, pc);
            NSLX_LAUNCH_CHECK();
            axpy(w.r.p, sv, sv, mpsi, 1.0, done);         // s = r - psi s
            axpy(sv, As, sv, pc, 1.0, done);              // s += psi chi As
            fm.apply(op, sv, As, nb, NSLX_F64, s);
            dot(y, As, yAs);
            dot(w.r.p, w.r.p, d1);
            hipLaunchKernelGGL(k_bi_update, dim3(gs), dim3(256), 0, s, ns, yAs, dp1, d1, t2.p, phi, delta, rr, done,
                               (int)it);
            NSLX_LAUNCH_CHECK();
            if (it % poll == 0 || it == maxit) finished = all_done();
        }
    }
    finished = all_done();
    std::vector<double> rrh(2 * ns);
    NSLX_HIP(hipMemcpyAsync(rrh.data(), rr, rrh.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));   // scalars are freed on return
    int64_t conv = 0, maxi = 0;
    for (int v : dh) if (v >= 0) { ++conv; maxi = std::max<int64_t>(maxi, v); }
    if (rep) {
        rep->iterations = finished ? maxi : it;
        rep->iterations_total = rep->iterations;
        rep->outer_steps = 0;
        rep->converged_rhs = global ? (finished ? nb : 0) : conv;
        double tsum = 0.0;
        for (int64_t i = 0; i < ns; ++i) tsum += rrh[2 * i];
        rep->rsqr = tsum;
    }
    return finished ? NSLX_OK : NSLX_NOT_CONVERGED;
}

}  // namespace nslx
