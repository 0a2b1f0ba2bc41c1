// Hop tile autotuner: times the tile variants that fit the lattice on the
// fermion matrix's populated fields and records the fastest per
// (precision, nx, nt, nbatch class) in the context's tile table.
//
// Two workloads can be timed: plain M + M^dag applies (every tile, the
// E-slice tiles included), or one iteration of the fused CG on MM^dag -- the
// CGA launch (M^dag with p = r + beta p and the |u|^2 partials) followed by
// the CGB launch (M with the x, r updates and the |r|^2 partials) -- which is
// what the solvers and the batching engine run; that mode times the classic
// tiles 0-3, the ones both the single-problem solver and the grouped
// launches take.  The CG launches run on scratch vectors (all zero: the
// recurrence scalars come out as NaN, which changes no timing) with their
// own partial-sum and flag arrays, so no solver state is touched.
//
// tuned_tile() is what the solvers call: the table's entry, or, with
// NSLX_AUTOTUNE (default 1) and a block big enough for the choice to matter
// (>= 512 workgroups of the 32 x 96 tile), a CG-mode tuning run on first use.
#include <algorithm>
#include <limits>

#include "engine.hpp"

namespace nslx {

int fm_tune(FermionMatrix& fm, int prec, int64_t nb, int reps, int mode, int* best_tile, double* best_us,
            hipStream_t s) {
    NSLX_REQUIRE(prec == NSLX_F64 || prec == NSLX_F32 || prec == NSLX_BF16X3, "tune: bad precision");
    NSLX_REQUIRE(reps >= 1, "tune: reps must be >= 1");
    NSLX_REQUIRE(mode == NSLX_TUNE_APPLY || mode == NSLX_TUNE_CG, "tune: bad mode");
    HopArgs a = fm.hop_args(prec, nb, s);
    const size_t esz = prec == NSLX_F64 ? 16 : 8;
    const size_t bytes = (size_t)nb * fm.len() * esz;
    const bool cg = mode == NSLX_TUNE_CG;
    DevBuf<unsigned char> in(bytes), out(bytes);
    DevBuf<unsigned char> r, p0, p1, x;
    DevBuf<double> upart, rp0, rp1, tol2, rref;
    DevBuf<int> done, want;
    NSLX_HIP(hipMemsetAsync(in.p, 0, bytes, s));
    NSLX_HIP(hipMemsetAsync(out.p, 0, bytes, s));
    const int n = cg ? 4 : (int)tile_menu(prec).size();
    if (cg) {
        int maxparts = 1;
        for (int t = 0; t < n; ++t) maxparts = std::max(maxparts, hop_nparts(prec, t, fm.nx, fm.nt));
        const size_t np = (size_t)nb * maxparts;
        r.alloc(bytes); p0.alloc(bytes); p1.alloc(bytes); x.alloc(bytes);
        upart.alloc(np); rp0.alloc(np); rp1.alloc(np);
        tol2.alloc(nb); rref.alloc(nb); done.alloc(nb); want.alloc(nb);
        for (unsigned char* v : {r.p, p0.p, p1.p, x.p}) NSLX_HIP(hipMemsetAsync(v, 0, bytes, s));
        for (double* v : {upart.p, rp0.p, rp1.p}) NSLX_HIP(hipMemsetAsync(v, 0, np * sizeof(double), s));
        NSLX_HIP(hipMemsetAsync(tol2.p, 0, nb * sizeof(double), s));
        NSLX_HIP(hipMemsetAsync(rref.p, 0, nb * sizeof(double), s));
        NSLX_HIP(hipMemsetAsync(done.p, 0xFF, nb * sizeof(int), s));   // every RHS active
        NSLX_HIP(hipMemsetAsync(want.p, 0, nb * sizeof(int), s));
        a.cg.upart = upart.p;
        a.cg.rpart[0] = rp0.p;
        a.cg.rpart[1] = rp1.p;
        a.cg.tol2 = tol2.p;
        a.cg.done = done.p;
        a.cg.want = want.p;
        a.cg.rr_ref = rref.p;
        a.cg.global = 0;
        a.cg.ru = prec != NSLX_F64;
    }
    auto run = [&](int t, int it) {
        if (!cg) {
            HopArgs h = a;
            h.X = in.p;
            h.Y = out.p;
            launch_hop(prec, HOP_M, MODE_PLAIN, h, t, s);
            launch_hop(prec, HOP_MD, MODE_PLAIN, h, t, s);
            return;
        }
        void* pb[2] = {p0.p, p1.p};
        HopArgs A = a;
        A.X = r.p;
        A.Pold = pb[(it - 1) & 1];
        A.Pnew = pb[it & 1];
        A.Y = out.p;
        A.iter = it;
        launch_hop(prec, HOP_MD, MODE_CGA, A, t, s);
        HopArgs B = a;
        B.X = out.p;
        B.Pold = pb[it & 1];
        B.Xs = x.p;
        B.R = r.p;
        B.iter = it;
        launch_hop(prec, HOP_M, MODE_CGB, B, t, s);
    };
    hipEvent_t e0, e1;
    NSLX_HIP(hipEventCreate(&e0));
    NSLX_HIP(hipEventCreate(&e1));
    int best = -1;
    double best_t = std::numeric_limits<double>::infinity();
    try {
        for (int t = 0; t < n; ++t) {
            if (!tile_fits(prec, t, fm.nx)) continue;
            if (fm.disc == DISC_DIAG && t >= 4) continue;   // HubbardDiag: classic tiles only
            run(t, 2);                                      // warm-up
            NSLX_HIP(hipEventRecord(e0, s));
            for (int k = 0; k < reps; ++k) run(t, 3 + k);
            NSLX_HIP(hipEventRecord(e1, s));
            NSLX_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            NSLX_HIP(hipEventElapsedTime(&ms, e0, e1));
            const double us = 1e3 * ms / reps;
            if (us < best_t) {
                best_t = us;
                best = t;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    NSLX_HIP(hipStreamSynchronize(s));   // scratch buffers are freed on return
    NSLX_REQUIRE(best >= 0, "tune: no tile fits this lattice");
    fm.ctx->tiles.set(prec, fm.nx, fm.nt, nb, best);
    if (best_tile) *best_tile = best;
    if (best_us) *best_us = best_t;
    return NSLX_OK;
}

int tuned_tile(FermionMatrix& fm, int prec, int64_t nb, hipStream_t s) {
    static const int autotune = (int)env_double("NSLX_AUTOTUNE", 1.0);
    int t;
    if (fm.ctx->tiles.lookup(prec, fm.nx, fm.nt, nb, &t)) return fm.tile(prec, nb);
    const int64_t wg_mid = ((nb * fm.nt + 31) / 32) * ((fm.nx + 95) / 96);
    if (autotune && wg_mid >= 512) fm_tune(fm, prec, nb, 3, NSLX_TUNE_CG, nullptr, nullptr, s);
    return fm.tile(prec, nb);
}

}  // namespace nslx
