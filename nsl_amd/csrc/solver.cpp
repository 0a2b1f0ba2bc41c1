// Krylov solvers on the device (reference: src/NSL/LinAlg/Solver/Impl/CG.tpp,
// BiCGStab.tpp, CGpreconditioned.tpp).
//
// * Fused CG on MM^dagger / M^dagger M: two hop launches per iteration carry
//   the whole recurrence (hop.hpp MODE_CGA/MODE_CGB); alpha, beta, |r|^2 and
//   the per-RHS stop flags stay on the device; the host polls a status word
//   every few iterations.  Iterations after convergence are no-ops, so the
//   returned x is the iterate of the converging step, as in CG.tpp:70-73.
// * Per-RHS relative tolerance (every precision): CG with FP64 reliable
//   updates -- the recurrence runs in FP32 / BF16x3 / FP64 storage while x
//   and the true residual are periodically recomputed in FP64 (the
//   inner/outer correction pattern of CGpreconditioned.tpp:14-120, without
//   restarting the Krylov process).  The stop test is the FP64 TRUE residual.
// * BiCGStab and CG on the non-Hermitian single hops: device dots with the
//   scalars formed on the host (as the reference, one sync per inner product).
#include <algorithm>
#include <chrono>
#include <cmath>

#include "engine.hpp"

namespace nslx {

void SolverWork::ensure_scalars(int64_t nb, int64_t nparts, hipStream_t s) {
    (void)s;
    // complex scalars of the generic solvers need 2 doubles per slot
    const size_t n = (size_t)std::max<int64_t>(nb, 1);
    rr.ensure(2 * n); alpha.ensure(2 * n); beta.ensure(2 * n); tol2.ensure(n); bn2.ensure(n); rn2.ensure(n);
    otol2.ensure(n); scal.ensure(16 * n + 16);
    part.ensure(n * std::max<int64_t>(nparts, 1) + 16);
    upart.ensure(n * std::max<int64_t>(nparts, 1) + 16);
    rpart0.ensure(n * std::max<int64_t>(nparts, 1) + 16);
    rpart1.ensure(n * std::max<int64_t>(nparts, 1) + 16);
    done.ensure(n); status.ensure(4); want.ensure(n);
    status_h.ensure(4);
    dbl_h.ensure(8 * n + 16);
    int_h.ensure(n + 4);
}

namespace {

// set by solve() for NSLX_TOL_REL_PER_RHS_FLOOR: a stop at the FP64 floor counts as converged
thread_local bool g_floor_ok = false;

inline int poll_every() {
    static const int v = (int)env_double("NSLX_POLL", 16.0);
    return std::max(1, v);
}

CgDev make_cg(SolverWork& w, int global) {
    CgDev c;
    c.upart = w.upart.p;
    c.rpart[0] = w.rpart0.p;
    c.rpart[1] = w.rpart1.p;
    c.tol2 = w.tol2.p;
    c.done = w.done.p;
    c.want = w.want.p;
    c.global = global;
    return c;
}

// Number of RHS (or batch slots) still active: done < 0 on the device.
int64_t count_active(SolverWork& w, int64_t nslot, hipStream_t s) {
    NSLX_HIP(hipMemcpyAsync(w.int_h.p, w.done.p, nslot * sizeof(int), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    int64_t a = 0;
    for (int64_t i = 0; i < nslot; ++i) a += w.int_h.p[i] < 0 ? 1 : 0;
    return a;
}

// Fused CG on A = outer(inner(.)) with the reference's recurrence from x0
// (held in x).  tol2 must be set.  Buffers r, p0, p1, u: nb*len elements.
// Returns the number of launched iterations; per-slot iteration counts in done_h.
int64_t cg_fused(FermionMatrix& fm, int prec, int inner, int outer, const void* b, void* x, int64_t nb, int global,
                 int64_t maxit, void* r, void* p0, void* p1, void* u, hipStream_t s, std::vector<int>& done_h) {
    SolverWork& w = fm.sw;
    const int tile = tuned_tile(fm, prec, nb, s);
    const int nparts = hop_nparts(prec, tile, fm.nx, fm.nt);
    w.ensure_scalars(nb, nparts, s);
    HopArgs a = fm.hop_args(prec, nb, s);
    a.cg = make_cg(w, global);
    const int64_t nslot = global ? 1 : nb;
    NSLX_HIP(hipMemsetAsync(w.done.p, 0xFF, nslot * sizeof(int), s));   // all active (-1)
    {
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
        if (it % chk == 0 || it == maxit) finished = count_active(w, nslot, s) == 0;
    }
    done_h.assign(nslot, -1);
    NSLX_HIP(hipMemcpyAsync(w.int_h.p, w.done.p, nslot * sizeof(int), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < nslot; ++i) done_h[i] = w.int_h.p[i];
    // |r_j|^2 of the iteration j a slot stopped at, from the partials buffer
    // of j's parity (a stopped slot's later launches write no partials), or
    // of iteration `it` for a slot still running, whose stop test (made by
    // the next CGA) is made here
    const size_t np = (size_t)nb * nparts;
    std::vector<double> rp[2] = {std::vector<double>(np), std::vector<double>(np)};
    std::vector<double> t2(nslot);
    NSLX_HIP(hipMemcpyAsync(rp[0].data(), w.rpart0.p, np * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipMemcpyAsync(rp[1].data(), w.rpart1.p, np * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipMemcpyAsync(t2.data(), w.tol2.p, nslot * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    w.last_rr.assign(nslot, 0.0);
    for (int64_t b2 = 0; b2 < nb; ++b2) {
        const int64_t slot = global ? 0 : b2;
        const int64_t j = done_h[slot] >= 0 ? done_h[slot] : it;
        const std::vector<double>& part = rp[j & 1];
        for (int k = 0; k < nparts; ++k) w.last_rr[slot] += part[(size_t)b2 * nparts + k];
    }
    for (int64_t i = 0; i < nslot; ++i)
        if (done_h[i] < 0 && w.last_rr[i] <= t2[i]) done_h[i] = (int)it;
    return it;
}

void copy_dev(void* dst, const void* src, size_t bytes, hipStream_t s) {
    NSLX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
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
    const int64_t it = cg_fused(fm, NSLX_F64, inner, outer, b, x, nb, global, maxit, w.r.p, w.p0.p, w.p1.p, w.u.p, s,
                                done);
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
        double t = 0.0;
        for (double v : w.last_rr) t += v;
        rep->rsqr = t;
    }
    return all ? NSLX_OK : NSLX_NOT_CONVERGED;
}

// ------------------------------- CG with FP64 reliable updates (per RHS)
// The CG recurrence runs in the "low" storage (c64 for F32/BF16X3, c128 for
// F64) on the fused hop kernels; x and the true residual live in FP64.  When
// a fraction frac of the active RHS has reduced its recursive |r|^2 by
// delta^2 since the last replacement (or reached tol^2), or maxgap iterations
// passed, the low-storage increment is folded into x, r = b - A x is
// recomputed in FP64 and replaces the recursive residual, while the search
// direction p is kept (residual replacement: Sleijpen & van der Vorst 1996;
// van der Vorst & Ye 2000), so the Krylov process is not restarted.  A RHS
// also stops at the attainable accuracy of an FP64 residual,
// |r| <= c eps ||A|| |x| (no FP64 vector x has a smaller true residual; the
// near-singular MM^dagger of random fields at large beta puts this above
// 1e-12 |b|).  Otherwise a RHS stops once its FP64 TRUE relative
// residual is <= tol.
struct RuPhase {
    FermionMatrix& fm;
    int op;
    const void* b;
    void* x;
    int64_t nb;
    double floor2;
    double delta, frac;
    int64_t maxgap;
    int chk;
    hipStream_t s;
    // double-double mode: x + xlo is the solution, accumulated with TwoSum;
    // once the FP64 true residual of some RHS comes within 1e4 of the FP64
    // floor (floor2_64 |x|^2) the replacements evaluate it in double-double
    void* xlo = nullptr;
    double floor2_64 = 0.0;
    bool dd_on = false;
    int64_t it = 0;      // iterations so far (all phases)
    int64_t nupd = 0;    // FP64 residual replacements so far
    int64_t ndd = 0;     // of which in double-double
    int64_t active = 0, waiting = 0;
    int64_t rhs_iters[2] = {0, 0};   // active RHS x iterations: [0] low precision, [1] FP64

    // host copy of done[] / want[]: active RHS and those not yet asking for
    // a residual replacement
    void read_flags() {
        SolverWork& w = fm.sw;
        NSLX_HIP(hipMemcpyAsync(w.int_h.p, w.done.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        NSLX_HIP(hipMemcpyAsync(w.status_h2.p, w.want.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        NSLX_HIP(hipStreamSynchronize(s));
        active = waiting = 0;
        for (int64_t i = 0; i < nb; ++i)
            if (w.int_h.p[i] < 0) { ++active; if (!w.status_h2.p[i]) ++waiting; }
    }

    // FP64 (or, in double-double mode near the floor, double-double) true
    // residual of the current x for every RHS, and |x|^2.
    void true_residual(bool x_is_zero) {
        SolverWork& w = fm.sw;
        const int64_t len = fm.len();
        if (x_is_zero) {
            NSLX_HIP(hipMemcpyAsync(w.r64.p, b, (size_t)nb * len * 16, hipMemcpyDeviceToDevice, s));
            NSLX_HIP(hipMemcpyAsync(w.rn2.p, w.bn2.p, nb * sizeof(double), hipMemcpyDeviceToDevice, s));
        } else if (dd_on) {
            dd_residual(fm, op, b, x, xlo, w.r64.p, w.rn2.p, nb, fm.ddw, s);
            ++ndd;
        } else {
            fm.apply(op, x, w.t64.p, nb, NSLX_F64, s);
            resid64(b, w.t64.p, w.r64.p, nb, len, w.rn2.p, fm.vsc, s);
        }
        norm2_perb(1, x, nb, len, w.xn2.p, fm.vsc, s);
        if (xlo && !dd_on && !x_is_zero) {
            // switch to double-double residuals within 1e4 of the FP64 floor
            std::vector<double> rn(nb), xn(nb);
            NSLX_HIP(hipMemcpyAsync(rn.data(), w.rn2.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
            NSLX_HIP(hipMemcpyAsync(xn.data(), w.xn2.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
            NSLX_HIP(hipStreamSynchronize(s));
            for (int64_t i = 0; i < nb; ++i)
                if (rn[i] <= 1e8 * floor2_64 * xn[i]) { dd_on = true; break; }
            if (dd_on) {   // this replacement already in double-double
                dd_residual(fm, op, b, x, xlo, w.r64.p, w.rn2.p, nb, fm.ddw, s);
                ++ndd;
            }
        }
    }

    // x (+ xlo) += the recurrence's increment xs (then xs = 0)
    void accumulate(int low64, void* xs) {
        SolverWork& w = fm.sw;
        const int64_t len = fm.len();
        if (!xlo) {
            ru_accumulate(low64, x, xs, nb, len, s);
            return;
        }
        const void* d = xs;
        if (!low64) {
            c64_to_c128(xs, w.t64.p, nb * len, s);
            d = w.t64.p;
        }
        dd_accumulate(x, xlo, d, nb, len, nullptr, s);
        NSLX_HIP(hipMemsetAsync(xs, 0, (size_t)nb * len * (low64 ? 16 : 8), s));
    }

    // One reliable-update CG phase in precision prec over the RHS with
    // done < 0 or done == RU_ESCALATE (a fresh recurrence from the current x).
    void run(int prec, bool x_is_zero, int stall_max, int64_t maxit) {
        SolverWork& w = fm.sw;
        // a low-precision phase escalates RHS whose projected iterations
        // exceed its budget (ru_scalars); FP64 phases run to maxit
        const int budget = prec == NSLX_F64 ? 0 : (int)std::min<int64_t>(maxit, 1 << 30);
        const int it0 = (int)it;
        const int64_t len = fm.len();
        const int low64 = prec == NSLX_F64;
        const size_t blo = (size_t)nb * len * (low64 ? 16 : 8);
        w.r.ensure(blo); w.p0.ensure(blo); w.p1.ensure(blo); w.u.ensure(blo); w.x32.ensure(blo);
        const int tile = tuned_tile(fm, prec, nb, s);
        const int nparts = hop_nparts(prec, tile, fm.nx, fm.nt);
        w.ensure_scalars(nb, nparts, s);
        w.status_h2.ensure(nb + 4);
        CgDev cg = make_cg(w, 0);
        cg.ru = 1;
        cg.delta2 = delta * delta;
        cg.rr_ref = w.rr_ref.p;
        const int inner = op == NSLX_OP_MMDAGGER ? HOP_MD : HOP_M;
        const int outer = op == NSLX_OP_MMDAGGER ? HOP_M : HOP_MD;

        true_residual(x_is_zero);
        ru_residual_to_low(low64, w.r64.p, w.r.p, w.p0.p, w.x32.p, 1, nb, len, s);
        ru_scalars(w.rn2.p, w.xn2.p, floor2, cg, nb, nparts, 0, (int)it, 1, w.best.p, w.stall.p, stall_max, w.rstart.p,
                   it0, budget, s);
        read_flags();
        bool finished = active == 0;
        HopArgs a = fm.hop_args(prec, nb, s);
        a.cg = cg;
        void* pb[2] = {w.p0.p, w.p1.p};
        int64_t k = 0, last_upd = 0, kprev = 0, act_prev = active;
        while (!finished && k < maxit) {
            ++k;
            ++it;
            HopArgs A = a;
            A.X = w.r.p;
            A.Pold = pb[(k - 1) & 1];
            A.Pnew = pb[k & 1];
            A.Y = w.u.p;
            A.iter = (int)k;
            launch_hop(prec, inner, MODE_CGA, A, tile, s);
            HopArgs B = a;
            B.X = w.u.p;
            B.Pold = pb[k & 1];
            B.Xs = w.x32.p;
            B.R = w.r.p;
            B.iter = (int)k;
            launch_hop(prec, outer, MODE_CGB, B, tile, s);
            if (k % chk == 0 || k == maxit) {
                read_flags();
                rhs_iters[low64] += act_prev * (k - kprev);   // active RHS x iterations since the last poll
                kprev = k;
                act_prev = active;
                if (waiting <= (1.0 - frac) * active || k - last_upd >= maxgap || k == maxit) {
                    accumulate(low64, w.x32.p);
                    true_residual(false);
                    ru_scalars(w.rn2.p, w.xn2.p, floor2, cg, nb, nparts, (int)(k & 1), (int)it, 0, w.best.p,
                               w.stall.p, stall_max, w.rstart.p, it0, budget, s);
                    ru_residual_to_low(low64, w.r64.p, w.r.p, nullptr, nullptr, 0, nb, len, s);
                    ++nupd;
                    last_upd = k;
                    read_flags();
                    act_prev = active;
                    finished = active == 0;
                }
            }
        }
    }
};

int solve_cg_ru(FermionMatrix& fm, int op, int prec, const void* b, void* x, const void* x0, int64_t nb, double tol,
                int64_t maxit, nslx_solve_report* rep, hipStream_t s, bool dd = false, void* xlo = nullptr) {
    SolverWork& w = fm.sw;
    const int64_t len = fm.len();
    const size_t b64 = (size_t)nb * len * 16;
    w.r64.ensure(b64); w.t64.ensure(b64);
    if (dd && !xlo) {
        w.xlo.ensure(b64);
        xlo = w.xlo.p;
    }
    if (!dd) xlo = nullptr;
    w.ensure_scalars(nb, 1, s);
    w.rr_ref.ensure(nb); w.rr_prev.ensure(nb); w.want.ensure(nb); w.xn2.ensure(nb);
    w.best.ensure(nb); w.stall.ensure(nb); w.rstart.ensure(nb);
    // ||MM^dag|| = ||M||^2 <= (1 + ||E|| max|P|)^2, ||E|| <= e^{|delta| ||K||},
    // |P| <= e^{|mu~| + max|Im phi|}
    const double nM = 1.0 + std::exp(std::fabs(fm.delta()) * fm.normK + fm.mut_abs() + fm.max_im_phi);
    const double fc = env_double("NSLX_FLOOR_C", 1.0) * 1.1102230246251565e-16 * nM * nM;
    // double-double mode: no attainable-accuracy stop (the residual is exact
    // enough to reach tol); floor2_64 only decides when to switch to it
    RuPhase ph{fm, op, b, x, nb, xlo ? 0.0 : fc * fc, env_double("NSLX_RU_DELTA", 0.1), env_double("NSLX_RU_FRAC", 0.5),
               (int64_t)env_double("NSLX_RU_MAXGAP", 200.0), poll_every(), s};
    ph.xlo = xlo;
    ph.floor2_64 = fc * fc;
    // A low-precision RHS is also finished in FP64 after its true residual
    // has made no 4x progress over this many replacements (the main test is
    // the true/recursive residual gap, see ru_scalars).
    const int stall_max = (int)env_double("NSLX_RU_STALL", 16.0);

    norm2_perb(1, b, nb, len, w.bn2.p, fm.vsc, s);
    set_tol2(w.bn2.p, w.tol2.p, nb, tol, 1, 0, s);
    if (x0) NSLX_HIP(hipMemcpyAsync(x, x0, b64, hipMemcpyDeviceToDevice, s));
    else NSLX_HIP(hipMemsetAsync(x, 0, b64, s));
    if (xlo) NSLX_HIP(hipMemsetAsync(xlo, 0, b64, s));
    NSLX_HIP(hipMemsetAsync(w.done.p, 0xFF, nb * sizeof(int), s));   // all active (-1)
    // Low-precision iteration budget: a system that low precision has not
    // solved by then is ill-conditioned beyond what its rounding allows
    // (eps_low * cond >> 1), and is solved in FP64 instead.
    const int64_t lowmax = (int64_t)env_double("NSLX_RU_LOWMAX", 2000.0);
    ph.run(prec, x0 == nullptr, prec == NSLX_F64 ? 0 : stall_max, prec == NSLX_F64 ? maxit : std::min(maxit, lowmax));
    int64_t escalated = 0;
    if (prec != NSLX_F64) {
        // RHS whose low-precision recurrence lost track, stagnated or ran out
        // of budget: solve again in FP64 from x0 (their low-precision iterate
        // carries rounding errors of order eps_low * cond and is discarded)
        std::vector<int> d(nb);
        NSLX_HIP(hipMemcpyAsync(d.data(), w.done.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        NSLX_HIP(hipStreamSynchronize(s));
        for (auto& v : d)
            if (v == RU_ESCALATE || v < 0) { v = RU_ESCALATE; ++escalated; }
        // NSLX_RU_ESC_KEEP=1 continues from the FP64-accumulated x instead
        const bool keep = env_double("NSLX_RU_ESC_KEEP", 0.0) != 0.0;
        if (escalated > 0 && ph.it < maxit) {
            const size_t rb = (size_t)len * 16;
            for (int64_t i = 0; i < nb && !keep; ++i) {
                if (d[i] != RU_ESCALATE) continue;
                char* xi = static_cast<char*>(x) + i * rb;
                if (x0) NSLX_HIP(hipMemcpyAsync(xi, static_cast<const char*>(x0) + i * rb, rb, hipMemcpyDeviceToDevice, s));
                else NSLX_HIP(hipMemsetAsync(xi, 0, rb, s));
                if (xlo) NSLX_HIP(hipMemsetAsync(static_cast<char*>(xlo) + i * rb, 0, rb, s));
            }
            NSLX_HIP(hipMemcpyAsync(w.done.p, d.data(), nb * sizeof(int), hipMemcpyHostToDevice, s));
            ph.run(NSLX_F64, false, 0, maxit - ph.it);
        }
    }
    const double floor2 = ph.floor2;
    const int64_t it = ph.it, nupd = ph.nupd;
    // final bookkeeping
    std::vector<double> rn2(nb), bn2(nb), xn2(nb, 0.0);
    std::vector<int> done(nb);
    NSLX_HIP(hipMemcpyAsync(rn2.data(), w.rn2.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipMemcpyAsync(bn2.data(), w.bn2.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipMemcpyAsync(done.data(), w.done.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    if (nupd > 0) NSLX_HIP(hipMemcpyAsync(xn2.data(), w.xn2.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    int64_t conv = 0, maxi = 0, floor_stops = 0;
    double maxrel = 0.0, tot = 0.0, maxfloor = 0.0;
    for (int64_t i = 0; i < nb; ++i) {
        const double fl2 = floor2 * xn2[i];
        if (done[i] >= 0 && done[i] != RU_ESCALATE) {
            ++conv;
            maxi = std::max<int64_t>(maxi, done[i]);
            if (rn2[i] > tol * tol * bn2[i]) ++floor_stops;
        }
        const double rel = bn2[i] > 0 ? std::sqrt(rn2[i] / bn2[i]) : std::sqrt(rn2[i]);
        maxrel = std::max(maxrel, rel);
        if (bn2[i] > 0) maxfloor = std::max(maxfloor, std::sqrt(fl2 / bn2[i]));
        tot += rn2[i];
    }
    if (rep) {
        rep->iterations = conv == nb ? maxi : it;
        rep->iterations_total = it;
        rep->outer_steps = nupd;
        rep->converged_rhs = conv;
        rep->rsqr = tot;
        rep->max_rel_residual = maxrel;
        rep->floor_stops = floor_stops;
        rep->max_floor_rel = maxfloor;
        rep->rhs_iters_low = ph.rhs_iters[0];
        rep->rhs_iters_f64 = ph.rhs_iters[1];
    }
    // a RHS that stopped at the FP64 floor above tol has not met the criterion
    // (report.floor_stops says so); NSLX_TOL_REL_PER_RHS_FLOOR accepts it
    const bool floor_ok = env_double("NSLX_FLOOR_OK", 0.0) != 0.0 || g_floor_ok;
    return (conv == nb && (floor_stops == 0 || floor_ok)) ? NSLX_OK : NSLX_NOT_CONVERGED;
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

}  // namespace

int solve(FermionMatrix& fm, int op, int method, const void* b, void* x, const void* x0, int64_t nb, double tol,
          int tol_kind, int64_t maxit, int prec, int check, nslx_solve_report* rep, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    NSLX_REQUIRE(b && x, "solve: null buffer");
    NSLX_REQUIRE(b != x && x0 != x, "solve: x must not alias b or x0");
    NSLX_REQUIRE(op >= NSLX_OP_M && op <= NSLX_OP_MDAGGERM, "solve: bad op");
    NSLX_REQUIRE(method == NSLX_CG || method == NSLX_BICGSTAB, "solve: bad method");
    NSLX_REQUIRE(tol_kind == NSLX_TOL_ABS_BATCH || tol_kind == NSLX_TOL_REL_PER_RHS ||
                     tol_kind == NSLX_TOL_REL_PER_RHS_FLOOR,
                 "solve: bad tol_kind");
    g_floor_ok = tol_kind == NSLX_TOL_REL_PER_RHS_FLOOR;
    if (tol_kind == NSLX_TOL_REL_PER_RHS_FLOOR) tol_kind = NSLX_TOL_REL_PER_RHS;
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
        st = solve_cg_ru(fm, op, prec, b, x, x0, nb, tol, maxit, rep, s);
    } else if (method == NSLX_CG && herm && tol_kind == NSLX_TOL_REL_PER_RHS) {
        st = solve_cg_ru(fm, op, NSLX_F64, b, x, x0, nb, tol, maxit, rep, s);
    } else if (method == NSLX_CG && herm) {
        st = solve_cg_f64(fm, op, b, x, x0, nb, tol, tol_kind, maxit, rep, s);
    } else {
        st = solve_generic_dev(fm, op, method, b, x, x0, nb, tol, tol_kind, maxit, rep, s);
    }
    if (check && prec == NSLX_F64) rep->max_rel_residual = true_rel_residual(fm, op, b, x, nb, s);
    NSLX_HIP(hipStreamSynchronize(s));
    rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

// ------------------------------------ CG with an inner-CG preconditioner
// The outer recurrence of CGpreconditioned.tpp:14-100 on fm's operator: x0 = b
// (or the given x0, :103-118), r = b - A x, z = P^{-1} r, p = z; per step
// t = A p, alpha = <r,z>/<p,t>, x += alpha p, r -= alpha t, stop on |r|^2,
// z = P^{-1} r, beta = <r,z>_new/<r,z>_old, p = z + beta p.  P^{-1} r is an
// inner CG on fp's operator with the same eps and maxIter started from r
// (the reference's innerCG_, CGpreconditioned.hpp:34-41,83-91); it may run in
// a low precision (reliable updates), which makes the preconditioner inexact.
// The outer vectors have their own buffers, so fp may be fm itself.  The
// reference computes z before its first stop test; here a RHS that is already
// converged skips that inner solve (x is the same).
int solve_dd(FermionMatrix& fm, int op, const void* b, void* x, void* xlo, const void* x0, int64_t nb, double tol,
             int64_t maxit, int prec, nslx_solve_report* rep, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    NSLX_REQUIRE(b && x, "solve_dd: null buffer");
    NSLX_REQUIRE(b != x && x0 != x && (!xlo || (xlo != x && xlo != b)), "solve_dd: x must not alias b or x0");
    NSLX_REQUIRE(op == NSLX_OP_MMDAGGER || op == NSLX_OP_MDAGGERM, "solve_dd: MMdagger or MdaggerM");
    NSLX_REQUIRE(prec == NSLX_F64 || prec == NSLX_F32 || prec == NSLX_BF16X3, "solve_dd: bad precision");
    NSLX_REQUIRE(tol > 0.0 && maxit >= 0, "solve_dd: bad tol/maxit");
    NSLX_REQUIRE(fm.disc == DISC_EXP, "solve_dd: HubbardExp only");
    fm.require_populated();
    NSLX_REQUIRE(nb >= 1 && nb % fm.nfield == 0, "solve_dd: nbatch must be a positive multiple of nfield");
    nslx_solve_report local{};
    if (!rep) rep = &local;
    *rep = nslx_solve_report{};
    g_floor_ok = false;
    const int st = solve_cg_ru(fm, op, prec, b, x, x0, nb, tol, maxit, rep, s, true, xlo);
    NSLX_HIP(hipStreamSynchronize(s));
    rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

int solve_precond(FermionMatrix& fm, FermionMatrix& fp, int op, const void* b, void* x, const void* x0, int64_t nb,
                  double tol, int tol_kind, int64_t maxit, int inner_prec, nslx_solve_report* rep, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    NSLX_REQUIRE(b && x, "solve_precond: null buffer");
    NSLX_REQUIRE(b != x && x0 != x, "solve_precond: x must not alias b or x0");
    NSLX_REQUIRE(op >= NSLX_OP_M && op <= NSLX_OP_MDAGGERM, "solve_precond: bad op");
    NSLX_REQUIRE(tol_kind == NSLX_TOL_ABS_BATCH || tol_kind == NSLX_TOL_REL_PER_RHS, "solve_precond: bad tol_kind");
    NSLX_REQUIRE(tol >= 0.0 && maxit >= 0, "solve_precond: bad tol/maxit");
    NSLX_REQUIRE(inner_prec == NSLX_F64 || inner_prec == NSLX_F32 || inner_prec == NSLX_BF16X3,
                 "solve_precond: bad inner precision");
    fm.require_populated();
    fp.require_populated();
    NSLX_REQUIRE(fp.nx == fm.nx && fp.nt == fm.nt, "solve_precond: the preconditioner's Nx and Nt must match");
    NSLX_REQUIRE(fp.ctx->device == fm.ctx->device, "solve_precond: both fermion matrices must be on one device");
    NSLX_REQUIRE(nb >= 1 && nb % fm.nfield == 0 && nb % fp.nfield == 0,
                 "solve_precond: nbatch must be a positive multiple of both nfield");
    nslx_solve_report local{};
    if (!rep) rep = &local;
    *rep = nslx_solve_report{};
    const int64_t len = fm.len();
    const size_t bytes = (size_t)nb * len * 16;
    const int global = tol_kind == NSLX_TOL_ABS_BATCH;
    const int64_t nslot = global ? 1 : nb;
    fm.sw.ensure_scalars(nb, 1, s);
    DevBuf<char> r(bytes), t(bytes), p(bytes), z(bytes);
    DevBuf<int> done_d(nslot);
    HostDots hd{fm, nb, global, s};
    std::vector<double> t2(nslot);
    if (global) t2[0] = tol * tol;
    else {
        auto bb = hd.dot(b, b);
        for (int64_t i = 0; i < nslot; ++i) t2[i] = tol * tol * bb[i].first;
    }
    std::vector<int> done(nslot, -1);
    std::vector<int> done_h(nslot);
    auto push_done = [&]() {
        for (int64_t i = 0; i < nslot; ++i) done_h[i] = done[i];
        NSLX_HIP(hipMemcpyAsync(done_d.p, done_h.data(), nslot * sizeof(int), hipMemcpyHostToDevice, s));
        NSLX_HIP(hipStreamSynchronize(s));
    };
    int64_t inner_its = 0, nprec = 0;
    auto precond = [&]() {
        nslx_solve_report ir{};
        // a non-converged inner solve still returns its current iterate (CG.tpp:90-93)
        (void)solve(fp, op, NSLX_CG, r.p, z.p, r.p, nb, tol,
                    tol_kind == NSLX_TOL_REL_PER_RHS ? NSLX_TOL_REL_PER_RHS_FLOOR : tol_kind, maxit, inner_prec, 0,
                    &ir, s);
        inner_its += ir.iterations_total;
        ++nprec;
    };
    const std::vector<cd> one(nslot, cd{1.0, 0.0});
    copy_dev(x, x0 ? x0 : b, bytes, s);
    fm.apply(op, x, t.p, nb, NSLX_F64, s);
    push_done();
    hd.axpy(b, t.p, r.p, one, -1.0, nullptr);   // r = b - A x
    auto rr = hd.dot(r.p, r.p);
    int64_t active = 0;
    for (int64_t i = 0; i < nslot; ++i) { if (rr[i].first <= t2[i]) done[i] = 0; else ++active; }
    push_done();
    std::vector<double> rz(nslot, 0.0);
    int64_t it = 0;
    if (active > 0) {
        precond();
        copy_dev(p.p, z.p, bytes, s);            // p = z
        auto d = hd.dot(r.p, z.p);
        for (int64_t i = 0; i < nslot; ++i) rz[i] = d[i].first;
    }
    while (active > 0 && it < maxit) {
        ++it;
        fm.apply(op, p.p, t.p, nb, NSLX_F64, s);
        auto pt = hd.dot(p.p, t.p);
        std::vector<cd> al(nslot);
        for (int64_t i = 0; i < nslot; ++i) al[i] = cdiv(cd{rz[i], 0.0}, pt[i]);
        hd.axpy(x, p.p, x, al, 1.0, done_d.p);
        hd.axpy(r.p, t.p, r.p, al, -1.0, done_d.p);
        auto rn = hd.dot(r.p, r.p);
        for (int64_t i = 0; i < nslot; ++i) {
            if (done[i] >= 0) continue;
            rr[i] = rn[i];
            if (rn[i].first <= t2[i]) { done[i] = (int)it; --active; }
        }
        push_done();
        if (active == 0) break;
        precond();
        auto rzn = hd.dot(r.p, z.p);
        std::vector<cd> be(nslot, cd{0.0, 0.0});
        for (int64_t i = 0; i < nslot; ++i) {
            if (done[i] >= 0) continue;
            be[i] = {rzn[i].first / rz[i], 0.0};
            rz[i] = rzn[i].first;
        }
        hd.axpy(z.p, p.p, p.p, be, 1.0, done_d.p);   // p = z + beta p
    }
    int64_t conv = 0, maxi = 0;
    for (int d : done) if (d >= 0) { ++conv; maxi = std::max<int64_t>(maxi, d); }
    const bool all = conv == nslot;
    rep->iterations = all ? maxi : it;
    rep->iterations_total = inner_its;
    rep->outer_steps = nprec;
    rep->converged_rhs = global ? (all ? nb : 0) : conv;
    double tot = 0.0;
    for (auto& v : rr) tot += v.first;
    rep->rsqr = tot;
    rep->max_rel_residual = true_rel_residual(fm, op, b, x, nb, s);
    NSLX_HIP(hipStreamSynchronize(s));
    rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return all ? NSLX_OK : NSLX_NOT_CONVERGED;
}

}  // namespace nslx
