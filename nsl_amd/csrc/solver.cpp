// Krylov solvers on the device (reference: src/NSL/LinAlg/Solver/Impl/CG.tpp,
// BiCGStab.tpp, CGpreconditioned.tpp).
//
// * Fused CG on MM^dagger / M^dagger M: two hop launches per iteration carry
//   the whole recurrence (hop.hpp MODE_CGA/MODE_CGB); alpha, beta, |r|^2 and
//   the per-RHS stop flags stay on the device; the host polls a status word
//   every few iterations.  Iterations after convergence are no-ops, so the
//   returned x is the iterate of the converging step, as in CG.tpp:70-73.
// * Defect correction (precision != F64): the FP64 true residual drives an
//   inner CG in FP32 or BF16x3 on the normalised residual; x is accumulated
//   in FP64 (the inner/outer pattern of CGpreconditioned.tpp:14-120 with a
//   low-precision inner solve of the same operator).
// * BiCGStab and CG on the non-Hermitian single hops: device dots with the
//   scalars formed on the host (as the reference, one sync per inner product).
#include <algorithm>
#include <chrono>
#include <cmath>

#include "engine.hpp"

namespace nslx {

void SolverWork::ensure_scalars(int64_t nb, int64_t nparts, hipStream_t s) {
    const size_t n = (size_t)std::max<int64_t>(nb, 1);
    rr.ensure(n); alpha.ensure(n); beta.ensure(n); tol2.ensure(n); bn2.ensure(n); rn2.ensure(n);
    otol2.ensure(n); scal.ensure(16 * n + 16);
    part.ensure(n * std::max<int64_t>(nparts, 1) + 16);
    done.ensure(n); status.ensure(4);
    if (!counter.p) {
        counter.alloc(4);
        NSLX_HIP(hipMemsetAsync(counter.p, 0, 4 * sizeof(unsigned), s));
    }
    status_h.ensure(4);
    dbl_h.ensure(8 * n + 16);
    int_h.ensure(n + 4);
}

namespace {

inline int poll_every() {
    static const int v = (int)env_double("NSLX_POLL", 8.0);
    return std::max(1, v);
}

CgDev make_cg(SolverWork& w, int global) {
    CgDev c;
    c.rr = w.rr.p; c.alpha = w.alpha.p; c.beta = w.beta.p; c.tol2 = w.tol2.p; c.done = w.done.p;
    c.status = w.status.p; c.part = w.part.p; c.counter = w.counter.p; c.global = global;
    return c;
}

// Fused CG on A = outer(inner(.)).  tol2 must be set.  x holds x0 unless
// x0_zero.  Buffers r, p0, p1, u: nb*len elements of the storage type.
// Returns the number of launched iterations; per-RHS iteration counts in done_h.
int64_t cg_fused(FermionMatrix& fm, int prec, int inner, int outer, const void* b, void* x, bool x0_zero,
                 int64_t nb, int global, int64_t maxit, void* r, void* p0, void* p1, void* u, hipStream_t s,
                 std::vector<int>& done_h) {
    SolverWork& w = fm.sw;
    const int tile = fm.tile(prec, nb);
    const int nparts = hop_nparts(prec, tile, fm.nx, fm.nt);
    w.ensure_scalars(nb, nparts, s);
    HopArgs a = fm.hop_args(prec, nb, s);
    a.cg = make_cg(w, global);
    const int f64 = prec == NSLX_F64;

    if (x0_zero) {
        cg_init_zero(f64, b, x, r, p0, nb, fm.len(), a.cg, fm.vsc, s);
    } else {
        HopArgs h = a;
        h.X = x;
        h.Y = u;
        launch_hop(prec, inner, MODE_PLAIN, h, tile, s);
        HopArgs c = a;
        c.X = u;
        c.Bv = b;
        c.R = r;
        c.Pnew = p0;
        c.iter = 0;
        launch_hop(prec, outer, MODE_CGI, c, tile, s);
    }
    void* pb[2] = {p0, p1};
    const int chk = poll_every();
    int64_t it = 0;
    bool finished = false;
    // status after init
    NSLX_HIP(hipMemcpyAsync(w.status_h.p, w.status.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    if (w.status_h.p[0] == 0) finished = true;
    while (!finished && it < maxit) {
        ++it;
        HopArgs A = a;
        A.X = r;
        A.Pold = pb[(it - 1) & 1];
        A.Pnew = pb[it & 1];
        A.Y = u;
        A.iter = (int)it;
        launch_hop(prec, inner, MODE_CGA, A, tile, s);
        HopArgs B = a;
        B.X = u;
        B.Pold = pb[it & 1];
        B.Xs = x;
        B.R = r;
        B.iter = (int)it;
        launch_hop(prec, outer, MODE_CGB, B, tile, s);
        if (it % chk == 0 || it == maxit) {
            NSLX_HIP(hipMemcpyAsync(w.status_h.p, w.status.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
            NSLX_HIP(hipStreamSynchronize(s));
            if (w.status_h.p[0] == 0) finished = true;
        }
    }
    const int64_t nslot = global ? 1 : nb;
    done_h.assign(nslot, -1);
    NSLX_HIP(hipMemcpyAsync(w.int_h.p, w.done.p, nslot * sizeof(int), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < nslot; ++i) done_h[i] = w.int_h.p[i];
    return it;
}

void copy_dev(void* dst, const void* src, size_t bytes, hipStream_t s) {
    NSLX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
}

double host_rsqr(SolverWork& w, int64_t nslot, hipStream_t s, double* maxv = nullptr) {
    NSLX_HIP(hipMemcpyAsync(w.dbl_h.p, w.rr.p, nslot * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    double t = 0.0, m = 0.0;
    for (int64_t i = 0; i < nslot; ++i) { t += w.dbl_h.p[i]; m = std::max(m, w.dbl_h.p[i]); }
    if (maxv) *maxv = m;
    return t;
}

// max_k |b_k - A x_k| / |b_k| in FP64
double true_rel_residual(FermionMatrix& fm, int op, const void* b, const void* x, int64_t nb, hipStream_t s) {
    SolverWork& w = fm.sw;
    const size_t bytes = (size_t)nb * fm.len() * 16;
    w.t64.ensure(bytes);
    w.r64.ensure(bytes);
    w.ensure_scalars(nb, 1, s);
    fm.apply(op, x, w.t64.p, nb, NSLX_F64, s);
    resid64(b, w.t64.p, w.r64.p, nb, fm.len(), w.rn2.p, fm.vsc, s);
    norm2_perb(1, b, nb, fm.len(), w.bn2.p, fm.vsc, s);
    NSLX_HIP(hipMemcpyAsync(w.dbl_h.p, w.rn2.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipMemcpyAsync(w.dbl_h.p + nb, w.bn2.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    double m = 0.0;
    for (int64_t i = 0; i < nb; ++i) {
        const double bn = w.dbl_h.p[nb + i];
        const double rel = bn > 0 ? std::sqrt(w.dbl_h.p[i] / bn) : std::sqrt(w.dbl_h.p[i]);
        m = std::max(m, rel);
    }
    return m;
}

// --------------------------------------------------------------- FP64 CG
int solve_cg_f64(FermionMatrix& fm, int op, const void* b, void* x, const void* x0, int64_t nb, double tol,
                 int tol_kind, int64_t maxit, nslx_solve_report* rep, hipStream_t s) {
    SolverWork& w = fm.sw;
    const int64_t len = fm.len();
    const size_t bytes = (size_t)nb * len * 16;
    w.r.ensure(bytes); w.p0.ensure(bytes); w.p1.ensure(bytes); w.u.ensure(bytes);
    const int global = tol_kind == NSLX_TOL_ABS_BATCH;
    w.ensure_scalars(nb, 1, s);
    // x0 = b (CG.tpp:97-103) unless given
    copy_dev(x, x0 ? x0 : b, bytes, s);
    if (global) {
        set_tol2(nullptr, w.tol2.p, nb, tol, 0, 1, s);
    } else {
        norm2_perb(1, b, nb, len, w.bn2.p, fm.vsc, s);
        set_tol2(w.bn2.p, w.tol2.p, nb, tol, 1, 0, s);
    }
    const int inner = op == NSLX_OP_MMDAGGER ? HOP_MD : HOP_M;
    const int outer = op == NSLX_OP_MMDAGGER ? HOP_M : HOP_MD;
    std::vector<int> done;
    const int64_t it = cg_fused(fm, NSLX_F64, inner, outer, b, x, false, nb, global, maxit, w.r.p, w.p0.p, w.p1.p,
                                w.u.p, s, done);
    int64_t conv = 0, maxi = 0;
    for (int d : done) {
        if (d >= 0) { ++conv; maxi = std::max<int64_t>(maxi, d); }
    }
    const bool all = conv == (int64_t)done.size();
    if (global && all) conv = nb;
    if (rep) {
        rep->iterations = all ? maxi : it;
        rep->iterations_total = rep->iterations;
        rep->outer_steps = 0;
        rep->converged_rhs = conv;
        rep->rsqr = host_rsqr(w, global ? 1 : nb, s);
    }
    return all ? NSLX_OK : NSLX_NOT_CONVERGED;
}

// ------------------------------------------------- defect correction (mixed)
int solve_cg_mixed(FermionMatrix& fm, int op, int prec, const void* b, void* x, const void* x0, int64_t nb,
                   double tol, int64_t maxit, nslx_solve_report* rep, hipStream_t s) {
    SolverWork& w = fm.sw;
    const int64_t len = fm.len();
    const size_t b64 = (size_t)nb * len * 16, b32 = (size_t)nb * len * 8;
    w.r64.ensure(b64); w.t64.ensure(b64);
    w.r32.ensure(b32); w.x32.ensure(b32);
    w.r.ensure(b32); w.p0.ensure(b32); w.p1.ensure(b32); w.u.ensure(b32);
    w.ensure_scalars(nb, 1, s);
    const double tau = env_double(prec == NSLX_F32 ? "NSLX_INNER_TOL_F32" : "NSLX_INNER_TOL_BF16", prec == NSLX_F32 ? 3e-6 : 3e-5);
    const int max_outer = (int)env_double("NSLX_MAX_OUTER", 30.0);
    const int inner = op == NSLX_OP_MMDAGGER ? HOP_MD : HOP_M;
    const int outer = op == NSLX_OP_MMDAGGER ? HOP_M : HOP_MD;

    norm2_perb(1, b, nb, len, w.bn2.p, fm.vsc, s);
    set_tol2(w.bn2.p, w.otol2.p, nb, tol, 1, 0, s);
    bool xzero = x0 == nullptr;
    if (xzero) NSLX_HIP(hipMemsetAsync(x, 0, b64, s));
    else copy_dev(x, x0, b64, s);

    int64_t total_inner = 0, outer_steps = 0, conv = 0;
    double maxrel = 0.0;
    std::vector<int> done;
    std::vector<double> bn2h(nb), rn2h(nb);
    NSLX_HIP(hipMemcpyAsync(w.dbl_h.p, w.bn2.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < nb; ++i) bn2h[i] = w.dbl_h.p[i];
    for (int step = 0;; ++step) {
        if (xzero) {
            copy_dev(w.r64.p, b, b64, s);
            copy_dev(w.rn2.p, w.bn2.p, nb * sizeof(double), s);
        } else {
            fm.apply(op, x, w.t64.p, nb, NSLX_F64, s);
            resid64(b, w.t64.p, w.r64.p, nb, len, w.rn2.p, fm.vsc, s);
        }
        NSLX_HIP(hipMemcpyAsync(w.dbl_h.p, w.rn2.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
        NSLX_HIP(hipStreamSynchronize(s));
        conv = 0;
        maxrel = 0.0;
        for (int64_t i = 0; i < nb; ++i) {
            rn2h[i] = w.dbl_h.p[i];
            const double rel = bn2h[i] > 0 ? std::sqrt(rn2h[i] / bn2h[i]) : std::sqrt(rn2h[i]);
            maxrel = std::max(maxrel, rel);
            if (rn2h[i] <= tol * tol * bn2h[i]) ++conv;
        }
        if (conv == nb || step >= max_outer) break;
        ++outer_steps;
        scale_to32(w.r64.p, w.r32.p, w.rn2.p, w.otol2.p, nb, len, s);
        set_tol2(nullptr, w.tol2.p, nb, tau, 0, 0, s);
        const int64_t it = cg_fused(fm, prec, inner, outer, w.r32.p, w.x32.p, true, nb, 0, maxit, w.r.p, w.p0.p,
                                    w.p1.p, w.u.p, s, done);
        int64_t mi = 0;
        for (int d : done) mi = std::max<int64_t>(mi, d >= 0 ? d : it);
        total_inner += mi;
        update64(x, w.x32.p, w.rn2.p, w.otol2.p, nb, len, s);
        xzero = false;
    }
    if (rep) {
        rep->iterations = total_inner;
        rep->iterations_total = total_inner;
        rep->outer_steps = outer_steps;
        rep->converged_rhs = conv;
        double t = 0.0;
        for (double v : rn2h) t += v;
        rep->rsqr = t;
        rep->max_rel_residual = maxrel;
    }
    return conv == nb ? NSLX_OK : NSLX_NOT_CONVERGED;
}

// ----------------------------------------- generic Krylov (host scalars)
struct HostDots {
    FermionMatrix& fm;
    int64_t nb;
    int global;
    hipStream_t s;
    // returns nslot complex values
    std::vector<std::pair<double, double>> dot(const void* a, const void* b) {
        const int64_t nslot = global ? 1 : nb;
        dot_perb(1, a, b, nb, fm.len(), fm.sw.scal.p, global, fm.vsc, s);
        NSLX_HIP(hipMemcpyAsync(fm.sw.dbl_h.p, fm.sw.scal.p, 2 * nslot * sizeof(double), hipMemcpyDeviceToHost, s));
        NSLX_HIP(hipStreamSynchronize(s));
        std::vector<std::pair<double, double>> out(nslot);
        for (int64_t i = 0; i < nslot; ++i) out[i] = {fm.sw.dbl_h.p[2 * i], fm.sw.dbl_h.p[2 * i + 1]};
        return out;
    }
    // z = x + sign * c_k * y with per-slot complex c (host), skipping done slots
    void axpy(const void* x, const void* y, void* z, const std::vector<std::pair<double, double>>& c, double sign,
              const int* done_dev) {
        const int64_t nslot = c.size();
        for (int64_t i = 0; i < nslot; ++i) { fm.sw.dbl_h.p[2 * i] = c[i].first; fm.sw.dbl_h.p[2 * i + 1] = c[i].second; }
        NSLX_HIP(hipMemcpyAsync(fm.sw.alpha.p, fm.sw.dbl_h.p, 2 * nslot * sizeof(double), hipMemcpyHostToDevice, s));
        axpy_perb(1, x, y, z, fm.sw.alpha.p, sign, nb, fm.len(), global, done_dev, s);
        NSLX_HIP(hipStreamSynchronize(s));   // dbl_h is reused immediately
    }
};

using cd = std::pair<double, double>;
inline cd cmul(cd a, cd b) { return {a.first * b.first - a.second * b.second, a.first * b.second + a.second * b.first}; }
inline cd cdiv(cd a, cd b) {
    const double d = b.first * b.first + b.second * b.second;
    return {(a.first * b.first + a.second * b.second) / d, (a.second * b.first - a.first * b.second) / d};
}

int solve_generic(FermionMatrix& fm, int op, int method, const void* b, void* x, const void* x0, int64_t nb, double tol,
                  int tol_kind, int64_t maxit, nslx_solve_report* rep, hipStream_t s) {
    SolverWork& w = fm.sw;
    const int64_t len = fm.len();
    const size_t bytes = (size_t)nb * len * 16;
    const int global = tol_kind == NSLX_TOL_ABS_BATCH;
    const int64_t nslot = global ? 1 : nb;
    w.ensure_scalars(nb, 1, s);
    w.r.ensure(bytes); w.p0.ensure(bytes); w.u.ensure(bytes);
    w.v1.ensure(bytes); w.v2.ensure(bytes); w.v3.ensure(bytes); w.v4.ensure(bytes);
    HostDots hd{fm, nb, global, s};
    std::vector<double> t2(nslot);
    if (global) t2[0] = tol * tol;
    else {
        auto bb = hd.dot(b, b);
        for (int64_t i = 0; i < nb; ++i) t2[i] = tol * tol * bb[i].first;
    }
    copy_dev(x, x0 ? x0 : b, bytes, s);
    // done flags on the device for axpy masking
    std::vector<int> done(nslot, -1);
    auto push_done = [&]() {
        for (int64_t i = 0; i < nslot; ++i) w.int_h.p[i] = done[i];
        NSLX_HIP(hipMemcpyAsync(w.done.p, w.int_h.p, nslot * sizeof(int), hipMemcpyHostToDevice, s));
        NSLX_HIP(hipStreamSynchronize(s));
    };
    const std::vector<cd> one(nslot, cd{1.0, 0.0});
    // r = b - A x
    fm.apply(op, x, w.u.p, nb, NSLX_F64, s);
    push_done();
    hd.axpy(b, w.u.p, w.r.p, one, -1.0, nullptr);
    auto rr = hd.dot(w.r.p, w.r.p);
    int64_t active = 0;
    for (int64_t i = 0; i < nslot; ++i) { if (rr[i].first <= t2[i]) done[i] = 0; else ++active; }
    push_done();
    int64_t it = 0;
    if (method == NSLX_CG) {
        void* p = w.p0.p;
        void* t = w.u.p;
        copy_dev(p, w.r.p, bytes, s);
        while (active > 0 && it < maxit) {
            ++it;
            fm.apply(op, p, t, nb, NSLX_F64, s);
            auto pt = hd.dot(p, t);
            std::vector<cd> al(nslot);
            for (int64_t i = 0; i < nslot; ++i) al[i] = cdiv(cd{rr[i].first, 0.0}, pt[i]);
            hd.axpy(x, p, x, al, 1.0, w.done.p);
            hd.axpy(w.r.p, t, w.r.p, al, -1.0, w.done.p);
            auto rn = hd.dot(w.r.p, w.r.p);
            std::vector<cd> be(nslot);
            for (int64_t i = 0; i < nslot; ++i) {
                if (done[i] >= 0) { be[i] = {0, 0}; continue; }
                be[i] = {rn[i].first / rr[i].first, 0.0};
                rr[i] = rn[i];
                if (rn[i].first <= t2[i]) { done[i] = (int)it; --active; }
            }
            // p = r + beta p for still-active slots (done slots are frozen)
            hd.axpy(w.r.p, p, p, be, 1.0, w.done.p);
            push_done();
        }
    } else {
        // BiCGStab.tpp:38-117: y = s = r; As = A s; delta = <y,r>; phi = <y,As>/delta
        void* y = w.v1.p; void* sv = w.v2.p; void* As = w.v3.p; void* wv = w.v4.p; void* Aw = w.u.p;
        copy_dev(y, w.r.p, bytes, s);
        copy_dev(sv, w.r.p, bytes, s);
        fm.apply(op, sv, As, nb, NSLX_F64, s);
        auto delta = hd.dot(y, w.r.p);
        auto yAs = hd.dot(y, As);
        std::vector<cd> phi(nslot);
        for (int64_t i = 0; i < nslot; ++i) phi[i] = cdiv(yAs[i], delta[i]);
        while (active > 0 && it < maxit) {
            ++it;
            std::vector<cd> om(nslot);
            for (int64_t i = 0; i < nslot; ++i) om[i] = cdiv(cd{1.0, 0.0}, phi[i]);
            hd.axpy(w.r.p, As, wv, om, -1.0, w.done.p);             // w = r - omega As
            fm.apply(op, wv, Aw, nb, NSLX_F64, s);
            auto aww = hd.dot(Aw, wv);
            auto awaw = hd.dot(Aw, Aw);
            std::vector<cd> chi(nslot);
            for (int64_t i = 0; i < nslot; ++i) chi[i] = cdiv(aww[i], cd{awaw[i].first, 0.0});
            hd.axpy(wv, Aw, w.r.p, chi, -1.0, w.done.p);            // r = w - chi Aw
            hd.axpy(x, sv, x, om, 1.0, w.done.p);                   // x += omega s
            hd.axpy(x, wv, x, chi, 1.0, w.done.p);                  // x += chi w
            auto dp1 = hd.dot(y, w.r.p);
            std::vector<cd> psi(nslot), mpsi(nslot), pc(nslot);
            for (int64_t i = 0; i < nslot; ++i) {
                psi[i] = cdiv(cmul(cd{-1.0, 0.0}, cmul(om[i], dp1[i])), cmul(delta[i], chi[i]));
                mpsi[i] = cd{-psi[i].first, -psi[i].second};
                pc[i] = cmul(psi[i], chi[i]);
            }
            // s = r - psi s + psi chi As  (t = s; s = r; s -= psi t; s += psi chi As)
            hd.axpy(w.r.p, sv, sv, mpsi, 1.0, w.done.p);
            hd.axpy(sv, As, sv, pc, 1.0, w.done.p);
            fm.apply(op, sv, As, nb, NSLX_F64, s);
            auto yAs2 = hd.dot(y, As);
            auto rn = hd.dot(w.r.p, w.r.p);
            for (int64_t i = 0; i < nslot; ++i) {
                if (done[i] >= 0) continue;
                phi[i] = cdiv(yAs2[i], dp1[i]);
                delta[i] = dp1[i];
                rr[i] = rn[i];
                if (rn[i].first <= t2[i]) { done[i] = (int)it; --active; }
            }
            push_done();
        }
    }
    int64_t conv = 0, maxi = 0;
    for (int d : done) if (d >= 0) { ++conv; maxi = std::max<int64_t>(maxi, d); }
    const bool all = conv == nslot;
    if (rep) {
        rep->iterations = all ? maxi : it;
        rep->iterations_total = rep->iterations;
        rep->outer_steps = 0;
        rep->converged_rhs = global ? (all ? nb : 0) : conv;
        double t = 0.0;
        for (auto& v : rr) t += v.first;
        rep->rsqr = t;
    }
    return all ? NSLX_OK : NSLX_NOT_CONVERGED;
}

}  // namespace

int solve(FermionMatrix& fm, int op, int method, const void* b, void* x, const void* x0, int64_t nb, double tol,
          int tol_kind, int64_t maxit, int prec, int check, nslx_solve_report* rep, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    NSLX_REQUIRE(b && x, "solve: null buffer");
    NSLX_REQUIRE(b != x && x0 != x, "solve: x must not alias b or x0");
    NSLX_REQUIRE(op >= NSLX_OP_M && op <= NSLX_OP_MDAGGERM, "solve: bad op");
    NSLX_REQUIRE(method == NSLX_CG || method == NSLX_BICGSTAB, "solve: bad method");
    NSLX_REQUIRE(tol_kind == NSLX_TOL_ABS_BATCH || tol_kind == NSLX_TOL_REL_PER_RHS, "solve: bad tol_kind");
    NSLX_REQUIRE(tol >= 0.0 && maxit >= 0, "solve: bad tol/maxit");
    NSLX_REQUIRE(prec == NSLX_F64 || prec == NSLX_F32 || prec == NSLX_BF16X3, "solve: bad precision");
    fm.require_populated();
    NSLX_REQUIRE(nb >= 1 && nb % fm.nfield == 0, "solve: nbatch must be a positive multiple of nfield");
    nslx_solve_report local{};
    if (!rep) rep = &local;
    *rep = nslx_solve_report{};
    const bool herm = op == NSLX_OP_MMDAGGER || op == NSLX_OP_MDAGGERM;
    int st;
    if (prec != NSLX_F64) {
        NSLX_REQUIRE(method == NSLX_CG && herm, "solve: low-precision inner solves need CG on MMdagger/MdaggerM");
        NSLX_REQUIRE(tol_kind == NSLX_TOL_REL_PER_RHS, "solve: low-precision inner solves need NSLX_TOL_REL_PER_RHS");
        st = solve_cg_mixed(fm, op, prec, b, x, x0, nb, tol, maxit, rep, s);
    } else if (method == NSLX_CG && herm) {
        st = solve_cg_f64(fm, op, b, x, x0, nb, tol, tol_kind, maxit, rep, s);
    } else {
        st = solve_generic(fm, op, method, b, x, x0, nb, tol, tol_kind, maxit, rep, s);
    }
    if (check && prec == NSLX_F64) rep->max_rel_residual = true_rel_residual(fm, op, b, x, nb, s);
    NSLX_HIP(hipStreamSynchronize(s));
    rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

}  // namespace nslx
