// Ragged batched CG engine (see pool.hpp).
//
// Every job owns a contiguous range of RHS slots (per-RHS tol2 / done flags)
// and of tile-partial slots.  A global iteration counter k lives on the
// device: iteration k is one grouped CGA launch (M^dag hop with p = r + beta
// p, |u|^2 partials) and one grouped CGB launch (M hop with x += alpha p,
// r -= alpha t, |r|^2 partials) per (precision, tile) class.  The kernels read
// the base iteration and the tile count from device memory (iteration = base
// + the launch's offset in the sequence; one one-thread kernel advances the
// base by P after the P iterations), so the launch sequence of P iterations is built once
// as a HIP graph and replayed; jobs join and leave between replays by
// rewriting the device tile lists.  A job joining before iteration k puts its
// first search direction in p[(k-1)&1], its |r|^2 in rpart[(k-1)&1] and zeros
// in rpart[k&1], so that its first beta is 0; P is even, so every replay
// starts at the same parity and the graph's roles stay valid.
//
// FP64 jobs: when all RHS have met their recursive tolerance the job is
// finalised: the true residual b - A x and the per-RHS test |r| <= tol |b| or
// |r| <= c eps ||A|| |x| (the FP64 attainable accuracy); otherwise it restarts
// from x with a fresh recurrence.
// Low-precision jobs run CG with FP64 reliable updates, as the single-problem
// solver: the kernels never stop a RHS on its recursive residual but flag it
// (want) once it has fallen by delta since the last replacement; at a replay
// boundary where half the job's active RHS want it (or 200 iterations passed)
// the low-precision increment is folded into x, r = b - A x is recomputed in
// FP64 and replaces the recursive residual, and the search direction is kept
// (ru_scalars: converged RHS stop, the others continue; a RHS that lost track,
// stalled or projects past the low-precision budget escalates the job, which
// then restarts from x in FP64).
#include "pool.hpp"

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstring>

namespace nslx {
namespace {

__global__ void k_advance(int* it, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *it += n;
}

// Slot initialisation of a job joining before iteration k.  With keep, RHS
// that reliable updates found converged (done >= 0, not RU_ESCALATE) stay done.
__global__ void k_join(int64_t nb, int64_t rhs0, int64_t part0, int nparts, const double* __restrict__ rn2,
                       const double* __restrict__ bn2, double tol, int keep, double* tol2, int* done, int* want,
                       double* rp_cur, double* rp_prev) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb * nparts) return;
    const int64_t b = i / nparts;
    const int k = (int)(i - b * nparts);
    rp_cur[part0 + i] = k == 0 ? rn2[b] : 0.0;
    rp_prev[part0 + i] = 0.0;
    if (k == 0) {
        tol2[rhs0 + b] = tol * tol * bn2[b];
        want[rhs0 + b] = 0;
        const int d = done[rhs0 + b];
        if (!(keep && d >= 0 && d != RU_ESCALATE)) done[rhs0 + b] = -1;
    }
}

inline unsigned g1(int64_t n) { return (unsigned)((n + 255) / 256); }

// c eps ||MM^dag||: the attainable accuracy of an FP64 residual per |x|, with
// ||MM^dag|| = ||M||^2 <= (1 + e^{|delta| ||K||} e^{|mu~| + max|Im phi|})^2.
double floor_c(const FermionMatrix& fm) {
    const double nM = 1.0 + std::exp(std::fabs(fm.delta()) * fm.normK + fm.mut_abs() + fm.max_im_phi);
    return env_double("NSLX_FLOOR_C", 1.0) * 1.1102230246251565e-16 * nM * nM;
}

}  // namespace

size_t PoolJob::bytes() const {
    const size_t len = fm ? (size_t)fm->len() : 0;
    const size_t v = (size_t)nb * len;
    const size_t low = prec == NSLX_F64 ? 16 : 8;
    return v * (4 * low + (prec == NSLX_F64 ? 0 : low) + 2 * 16);
}

Pool::Pool(Ctx* ctx, hipStream_t s, int64_t cap_rhs, int use_graph, int iters_per_replay)
    : ctx_(ctx), s_(s), cap_rhs_(cap_rhs), use_graph_(use_graph), P_(std::max(2, iters_per_replay & ~1)) {
    NSLX_REQUIRE(cap_rhs >= 1, "pool: capacity must be >= 1");
    cap_parts_ = cap_rhs * 128;
    tol2_.alloc(cap_rhs);
    rr_ref_.alloc(cap_rhs);
    done_.alloc(cap_rhs);
    want_.alloc(cap_rhs);
    upart_.alloc(cap_parts_);
    rpart0_.alloc(cap_parts_);
    rpart1_.alloc(cap_parts_);
    it_dev_.alloc(1);
    const int one = 1;
    NSLX_HIP(hipMemcpyAsync(it_dev_.p, &one, sizeof(int), hipMemcpyHostToDevice, s_));
    rhs_used_.assign(cap_rhs, 0);
    part_used_.assign(cap_parts_, 0);
    done_h_.assign(cap_rhs, 0);
    want_h_.assign(cap_rhs, 0);
    const int ncu = ctx->num_cu > 0 ? ctx->num_cu : 256;
    for (int c = 0; c < NCLS; ++c) {
        const int tile = tile_of(c);
        const int per_cu = tile == 2 ? 8 : 4;
        cls_[c].grid = (ncu * per_cu + 7) / 8 * 8;
        cls_[c].ntiles.alloc(1);
        const int zero = 0;
        NSLX_HIP(hipMemcpyAsync(cls_[c].ntiles.p, &zero, sizeof(int), hipMemcpyHostToDevice, s_));
    }
    NSLX_HIP(hipStreamSynchronize(s_));
}

Pool::~Pool() {
    (void)hipStreamSynchronize(s_);
    drop_graph();
}

void Pool::drop_graph() {
    if (exec_) { (void)hipGraphExecDestroy(exec_); exec_ = nullptr; }
    if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
    cls_mask_ = 0;
}

int64_t Pool::free_rhs() const {
    return (int64_t)std::count(rhs_used_.begin(), rhs_used_.end(), (char)0);
}

int64_t Pool::reserve_rhs(int64_t nb) { return nb >= 1 ? alloc_range(rhs_used_, nb) : -1; }

void Pool::release_rhs(int64_t start, int64_t nb) { free_range(rhs_used_, start, nb); }

// Partial-sum slots per RHS depend on the tile a job gets (ceil(nt/BM) *
// ceil(nx/BN) of them: 144 for the 16 x 32 tile at Nx = 1152, Nt = 64), so the
// arrays grow on demand: the live partials of running jobs are copied at the
// same offsets and the graph, which holds the old addresses, is rebuilt.
void Pool::grow_parts(int64_t need) {
    const int64_t used = (int64_t)std::count(part_used_.begin(), part_used_.end(), (char)1);
    const int64_t cap = std::max<int64_t>(2 * cap_parts_, used + need + (int64_t)part_used_.size() / 4 + 1024);
    DevBuf<double> u2(cap), r0(cap), r1(cap);
    const size_t old = (size_t)cap_parts_ * sizeof(double);
    NSLX_HIP(hipMemsetAsync(u2.p, 0, cap * sizeof(double), s_));
    NSLX_HIP(hipMemsetAsync(r0.p, 0, cap * sizeof(double), s_));
    NSLX_HIP(hipMemsetAsync(r1.p, 0, cap * sizeof(double), s_));
    NSLX_HIP(hipMemcpyAsync(u2.p, upart_.p, old, hipMemcpyDeviceToDevice, s_));
    NSLX_HIP(hipMemcpyAsync(r0.p, rpart0_.p, old, hipMemcpyDeviceToDevice, s_));
    NSLX_HIP(hipMemcpyAsync(r1.p, rpart1_.p, old, hipMemcpyDeviceToDevice, s_));
    NSLX_HIP(hipStreamSynchronize(s_));   // the old arrays are freed below
    upart_ = std::move(u2);
    rpart0_ = std::move(r0);
    rpart1_ = std::move(r1);
    part_used_.resize(cap, 0);
    cap_parts_ = cap;
    drop_graph();
}

int64_t Pool::alloc_range(std::vector<char>& used, int64_t n) {
    int64_t run = 0;
    for (int64_t i = 0; i < (int64_t)used.size(); ++i) {
        run = used[i] ? 0 : run + 1;
        if (run == n) {
            const int64_t s = i - n + 1;
            std::fill(used.begin() + s, used.begin() + s + n, (char)1);
            return s;
        }
    }
    return -1;
}

void Pool::free_range(std::vector<char>& used, int64_t start, int64_t n) {
    if (start >= 0) std::fill(used.begin() + start, used.begin() + start + n, (char)0);
}

CgDev Pool::cgdev() const {
    CgDev c;
    c.upart = upart_.p;
    c.rpart[0] = rpart0_.p;
    c.rpart[1] = rpart1_.p;
    c.tol2 = tol2_.p;
    c.done = done_.p;
    c.global = 0;
    c.ru = 0;
    c.rr_ref = rr_ref_.p;
    c.want = want_.p;
    return c;
}

CgDev Pool::slice(const PoolJob& j) const {
    CgDev c = cgdev();
    c.upart += j.part0;
    c.rpart[0] += j.part0;
    c.rpart[1] += j.part0;
    c.tol2 += j.rhs0;
    c.done += j.rhs0;
    c.rr_ref += j.rhs0;
    c.want += j.rhs0;
    return c;
}

// (Re)start the recurrence of job j before global iteration iter_: fresh
// (x = 0, r = b) or from the current x (r = b - A x in FP64); keep_done:
// RHS converged in the previous (reliable-update) recurrence stay done.
void Pool::start_recurrence(PoolJob& j, bool fresh, bool keep_done) {
    FermionMatrix& fm = *j.fm;
    const int64_t len = fm.len();
    const size_t n = (size_t)j.nb * len;
    const size_t esz = j.prec == NSLX_F64 ? 16 : 8;
    j.r.ensure(n * esz); j.p0.ensure(n * esz); j.p1.ensure(n * esz); j.u.ensure(n * esz);
    if (j.prec != NSLX_F64) j.xs.ensure(n * esz);
    j.r64.ensure(n); j.t64.ensure(n);
    j.rn2.ensure(j.nb); j.xn2.ensure(j.nb);
    if (fresh) {
        NSLX_HIP(hipMemsetAsync(j.x, 0, n * 16, s_));
        NSLX_HIP(hipMemcpyAsync(j.r64.p, j.b, n * 16, hipMemcpyDeviceToDevice, s_));
        NSLX_HIP(hipMemcpyAsync(j.rn2.p, j.bn2.p, j.nb * sizeof(double), hipMemcpyDeviceToDevice, s_));
    } else {
        fm.apply(j.op, j.x, j.t64.p, j.nb, NSLX_F64, s_);
        resid64(j.b, j.t64.p, j.r64.p, j.nb, len, j.rn2.p, fm.vsc, s_);
    }
    if (j.prec == NSLX_F64) NSLX_HIP(hipMemcpyAsync(j.r.p, j.r64.p, n * 16, hipMemcpyDeviceToDevice, s_));
    else {
        c128_to_c64(j.r64.p, j.r.p, (int64_t)n, s_);
        NSLX_HIP(hipMemsetAsync(j.xs.p, 0, n * esz, s_));
    }
    void* pprev = ((iter_ - 1) & 1) ? j.p1.p : j.p0.p;
    NSLX_HIP(hipMemcpyAsync(pprev, j.r.p, n * esz, hipMemcpyDeviceToDevice, s_));

    // class, tile and partial slots.  A job too small to fill the GPU with the
    // 32 x 96 tile (fewer than 512 of its workgroups) takes the 16 x 32 one
    // (classic tile 2): 6x the workgroups per iteration, and all such jobs of
    // a precision share one class, so one grouped launch per phase.  (Batching
    // leg 1340 -> 1680 solves/s, serving FP64 p50 31 -> 19 ms.)  Larger jobs
    // keep the tile the tile table or the default gives.
    const int64_t wg_mid = (((int64_t)j.nb * fm.nt + 31) / 32) * ((fm.nx + 95) / 96);
    int tile = 2;
    if (wg_mid >= 512) {
        tile = tuned_tile(fm, j.prec, j.nb, s_);
        if (tile >= 4 || !tile_fits(j.prec, tile, fm.nx)) tile = default_tile(j.prec, fm.nx, fm.nt, j.nb);
    }
    j.tile = tile;
    j.cls = cls_of(j.prec, tile, fm.disc, j.op);
    j.nparts = hop_nparts(j.prec, tile, fm.nx, fm.nt);
    j.part0 = alloc_range(part_used_, j.nb * j.nparts);
    if (j.part0 < 0) {
        grow_parts(j.nb * j.nparts);
        j.part0 = alloc_range(part_used_, j.nb * j.nparts);
    }
    NSLX_REQUIRE(j.part0 >= 0, "pool: out of partial-sum slots");
    double* cur = ((iter_ - 1) & 1) ? rpart1_.p : rpart0_.p;
    double* prev = (iter_ & 1) ? rpart1_.p : rpart0_.p;
    const int64_t m = j.nb * j.nparts;
    hipLaunchKernelGGL(k_join, dim3(g1(m)), dim3(256), 0, s_, j.nb, j.rhs0, j.part0, j.nparts, j.rn2.p, j.bn2.p, j.tol,
                       keep_done ? 1 : 0, tol2_.p, done_.p, want_.p, cur, prev);
    NSLX_LAUNCH_CHECK();
    if (j.prec != NSLX_F64) {
        // reliable-update state of the job's RHS (rr_ref, best, stall, rstart)
        j.best.ensure(j.nb); j.stall.ensure(j.nb); j.rstart.ensure(j.nb);
        norm2_perb(1, j.x, j.nb, len, j.xn2.p, fm.vsc, s_);
        const double fc = floor_c(fm);
        ru_scalars(j.rn2.p, j.xn2.p, fc * fc, slice(j), j.nb, j.nparts, (int)((iter_ - 1) & 1), (int)iter_, 1,
                   j.best.p, j.stall.p, 0, j.rstart.p, (int)iter_, 0, s_);
        j.ru_it0 = iter_;
        j.last_upd = iter_;
    }
    j.join_iter = iter_;
}

void Pool::admit(std::unique_ptr<PoolJob> jp) {
    PoolJob& j = *jp;
    NSLX_REQUIRE(j.fm && j.b && j.x && j.nb >= 1, "pool: bad job");
    j.fm->require_populated();
    NSLX_REQUIRE(j.nb % j.fm->nfield == 0, "pool: nbatch must be a multiple of the populated fields");
    NSLX_REQUIRE(j.op == NSLX_OP_MMDAGGER || j.op == NSLX_OP_MDAGGERM, "pool: the batched systems are MMdagger or MdaggerM");
    NSLX_REQUIRE(j.op == NSLX_OP_MMDAGGER || j.fm->disc == DISC_EXP, "pool: HubbardDiag jobs are MMdagger systems");
    NSLX_REQUIRE(j.prec == NSLX_F64 || j.prec == NSLX_F32 || j.prec == NSLX_BF16X3, "pool: bad precision");
    if (j.rhs0 < 0) j.rhs0 = alloc_range(rhs_used_, j.nb);   // else reserved by reserve_rhs
    NSLX_REQUIRE(j.rhs0 >= 0, "pool: out of RHS slots");
    try {
        j.bn2.ensure(j.nb);
        norm2_perb(1, j.b, j.nb, j.fm->len(), j.bn2.p, j.fm->vsc, s_);
        start_recurrence(j, true, false);
    } catch (...) {
        free_range(rhs_used_, j.rhs0, j.nb);
        if (j.part0 >= 0) free_range(part_used_, j.part0, j.nb * j.nparts);
        throw;
    }
    const int c = j.cls;
    jobs_.push_back(std::move(jp));
    rebuild_class(c);
}

void Pool::rebuild_class(int c) {
    Cls& K = cls_[c];
    std::vector<HopProblem> probs;
    std::vector<PoolJob*> owners;
    for (auto& jp : jobs_) {
        PoolJob& j = *jp;
        if (j.cls != c || j.finished) continue;
        FermionMatrix& fm = *j.fm;
        HopProblem P;
        P.nx = fm.nx;
        P.nt = fm.nt;
        P.nb = (int)j.nb;
        P.rpf = (int)(j.nb / fm.nfield);
        P.rhs0 = (int)j.rhs0;
        P.part0 = j.part0;
        P.E = fm.Eptr(j.prec, s_);
        P.P = fm.Pptr(j.prec);
        P.v[VS_R] = j.r.p;
        P.v[VS_P0] = j.p0.p;
        P.v[VS_P1] = j.p1.p;
        P.v[VS_U] = j.u.p;
        P.v[VS_X] = j.prec == NSLX_F64 ? j.x : (void*)j.xs.p;
        P.v[VS_B] = const_cast<void*>(j.b);
        probs.push_back(P);
        owners.push_back(&j);
    }
    K.njobs = (int)probs.size();
    std::vector<int4> tiles;
    if (!probs.empty()) {
        const int tile = tile_of(c);
        const int prec = prec_of(c);
        tiles = build_tile_list(probs, prec, tile);
    }
    if (probs.size() > K.cap_probs) {
        K.cap_probs = std::max<size_t>(probs.size(), 2 * K.cap_probs + 4);
        K.probs.alloc(K.cap_probs);
        drop_graph();
    }
    if (tiles.size() > K.cap_tiles) {
        K.cap_tiles = std::max<size_t>(tiles.size(), 2 * K.cap_tiles + 1024);
        K.tiles.alloc(K.cap_tiles);
        drop_graph();
    }
    if (!probs.empty()) {
        NSLX_HIP(hipMemcpyAsync(K.probs.p, probs.data(), probs.size() * sizeof(HopProblem), hipMemcpyHostToDevice, s_));
        NSLX_HIP(hipMemcpyAsync(K.tiles.p, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice, s_));
    }
    const int nt = (int)tiles.size();
    NSLX_HIP(hipMemcpyAsync(K.ntiles.p, &nt, sizeof(int), hipMemcpyHostToDevice, s_));
    NSLX_HIP(hipStreamSynchronize(s_));   // host vectors are freed on return
}

// P iterations over the active classes, launched directly or replayed from a
// graph.  The graph is built node by node (kernel nodes in launch order), not
// by stream capture, so that work other threads put on the legacy stream
// meanwhile never meets a capturing stream.
void Pool::run_iterations() {
    uint64_t mask = 0;
    for (int c = 0; c < NCLS; ++c)
        if (cls_[c].njobs > 0) mask |= uint64_t(1) << c;
    if (!mask) return;
    // the launch sequence: hop(c, pass, args) per grouped launch, then advance()
    auto sequence = [&](auto&& hop, auto&& advance) {
        const CgDev cg = cgdev();
        const double delta = env_double("NSLX_RU_DELTA", 0.1);
        for (int k = 0; k < P_; ++k) {
            const int64_t K = iter_ + k;   // parity only (the kernels read the counter)
            const int pold = ((K - 1) & 1) ? VS_P1 : VS_P0;
            const int pnew = ((K - 1) & 1) ? VS_P0 : VS_P1;
            for (int pass = 0; pass < 2; ++pass) {
                for (int c = 0; c < NCLS; ++c) {
                    if (!(mask & (uint64_t(1) << c))) continue;
                    GroupArgs g;
                    g.probs = cls_[c].probs.p;
                    g.nprob = cls_[c].njobs;
                    g.tiles = cls_[c].tiles.p;
                    g.ntiles = (int)std::max<size_t>(1, cls_[c].cap_tiles);
                    g.ntiles_dev = cls_[c].ntiles.p;
                    g.iter_dev = it_dev_.p;
                    g.iter = k;
                    g.cg = cg;
                    if (prec_of(c) != NSLX_F64) {   // low precision: reliable-update flags
                        g.cg.ru = 1;
                        g.cg.delta2 = delta * delta;
                    }
                    if (pass == 0) {   // CGA: Y = M^dag(p_new) (M(p_new) for MdaggerM), p_new = r + beta p_old
                        g.role[0] = VS_R; g.role[1] = VS_U; g.role[2] = (signed char)pold; g.role[3] = (signed char)pnew;
                    } else {           // CGB: t = M u (M^dag u), x += alpha p_new, r -= alpha t
                        g.role[0] = VS_U; g.role[2] = (signed char)pnew; g.role[4] = VS_X; g.role[5] = VS_R;
                    }
                    hop(c, pass, g);
                }
            }
        }
        advance();
    };
    // MMdagger = M (M^dag .): CGA applies M^dag, CGB M; MdaggerM the other way round
    auto op_of = [](int c, int pass) { return (pass == 0) != mdagm_of(c) ? HOP_MD : HOP_M; };
    auto mode_of = [](int pass) { return pass == 0 ? MODE_CGA : MODE_CGB; };
    if (use_graph_) {
        if (!exec_ || mask != cls_mask_) {
            drop_graph();
            NSLX_HIP(hipGraphCreate(&graph_, 0));
            hipGraphNode_t last = nullptr;
            // params: pointers to the argument values, which are copied into the node
            auto add = [&](const void* fn, unsigned grid, unsigned block, void** params) {
                hipKernelNodeParams kp{};
                kp.func = const_cast<void*>(fn);
                kp.gridDim = dim3(grid);
                kp.blockDim = dim3(block);
                kp.sharedMemBytes = 0;
                kp.kernelParams = params;
                kp.extra = nullptr;
                hipGraphNode_t n = nullptr;
                NSLX_HIP(hipGraphAddKernelNode(&n, graph_, last ? &last : nullptr, last ? 1 : 0, &kp));
                last = n;
            };
            int* itp = it_dev_.p;
            int pn = P_;
            sequence(
                [&](int c, int pass, GroupArgs& g) {
                    NSLX_REQUIRE(g.nprob > 0 && cls_[c].grid > 0, "pool: empty grouped launch");
                    const GroupedKernel gk = hop_grouped_kernel(prec_of(c), op_of(c, pass), mode_of(pass), tile_of(c), disc_of(c));
                    void* params[1] = {&g};
                    add(gk.fn, (unsigned)cls_[c].grid, (unsigned)gk.nthreads, params);
                },
                [&]() {
                    void* params[2] = {&itp, &pn};
                    add(reinterpret_cast<const void*>(&k_advance), 1, 64, params);
                });
            NSLX_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
            cls_mask_ = mask;
        }
        NSLX_HIP(hipGraphLaunch(exec_, s_));
    } else {
        sequence(
            [&](int c, int pass, GroupArgs& g) {
                launch_hop_grouped(prec_of(c), op_of(c, pass), mode_of(pass), tile_of(c), disc_of(c), g, cls_[c].grid, s_);
            },
            [&]() {
                hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s_, it_dev_.p, P_);
                NSLX_LAUNCH_CHECK();
            });
    }
    iter_ += P_;
    st_.iterations += P_;
    st_.replays += 1;
}

bool Pool::finalize(PoolJob& j) {
    FermionMatrix& fm = *j.fm;
    const int64_t len = fm.len();
    j.iters += iter_ - j.join_iter;
    if (j.prec != NSLX_F64) ru_accumulate(0, j.x, j.xs.p, j.nb, len, s_);
    fm.apply(j.op, j.x, j.t64.p, j.nb, NSLX_F64, s_);
    resid64(j.b, j.t64.p, j.r64.p, j.nb, len, j.rn2.p, fm.vsc, s_);
    norm2_perb(1, j.x, j.nb, len, j.xn2.p, fm.vsc, s_);
    std::vector<double> rn(j.nb), bn(j.nb), xn(j.nb);
    NSLX_HIP(hipMemcpyAsync(rn.data(), j.rn2.p, j.nb * sizeof(double), hipMemcpyDeviceToHost, s_));
    NSLX_HIP(hipMemcpyAsync(bn.data(), j.bn2.p, j.nb * sizeof(double), hipMemcpyDeviceToHost, s_));
    NSLX_HIP(hipMemcpyAsync(xn.data(), j.xn2.p, j.nb * sizeof(double), hipMemcpyDeviceToHost, s_));
    NSLX_HIP(hipStreamSynchronize(s_));
    const double fc = floor_c(fm);
    bool all = true;
    double maxrel = 0.0, maxfloor = 0.0;
    int64_t fstops = 0;
    for (int64_t i = 0; i < j.nb; ++i) {
        const double fl2 = fc * fc * xn[i];
        const bool ok = rn[i] <= j.tol * j.tol * bn[i] || rn[i] <= fl2;
        if (ok && rn[i] > j.tol * j.tol * bn[i]) ++fstops;
        const double r = bn[i] > 0 ? std::sqrt(rn[i] / bn[i]) : std::sqrt(rn[i]);
        maxrel = std::max(maxrel, r);
        if (bn[i] > 0) maxfloor = std::max(maxfloor, std::sqrt(fl2 / bn[i]));
        if (!ok) all = false;
    }
    j.max_rel = maxrel;
    j.max_floor_rel = maxfloor;
    j.floor_stops = fstops;
    if (all) { j.status = NSLX_OK; return true; }
    if (j.iters >= j.maxit || j.restarts >= 200) { j.status = NSLX_NOT_CONVERGED; return true; }
    // restart from x: FP64 jobs with a fresh recurrence; a low-precision job
    // ends here with RHS its reliable updates escalated (or out of its
    // low-precision budget) and continues in FP64
    ++j.restarts;
    ++st_.restarts;
    if (j.prec != NSLX_F64) {
        j.prec = NSLX_F64;
        j.escalated = true;
        ++st_.escalations;
        j.xs.reset();
    }
    return false;
}

// Reliable update of a low-precision job (in place, the search direction
// kept): x += xs, r = b - A x in FP64, ru_scalars on the job's slots, and the
// recurrence residual replaced by low(r).
void Pool::replace_residual(PoolJob& j) {
    FermionMatrix& fm = *j.fm;
    const int64_t len = fm.len();
    const double fc = floor_c(fm);
    const int64_t lowmax = (int64_t)env_double("NSLX_RU_LOWMAX", 2000.0);
    const int budget = (int)std::min<int64_t>(lowmax, std::max<int64_t>(1, j.maxit - j.iters));
    ru_accumulate(0, j.x, j.xs.p, j.nb, len, s_);
    fm.apply(j.op, j.x, j.t64.p, j.nb, NSLX_F64, s_);
    resid64(j.b, j.t64.p, j.r64.p, j.nb, len, j.rn2.p, fm.vsc, s_);
    norm2_perb(1, j.x, j.nb, len, j.xn2.p, fm.vsc, s_);
    ru_scalars(j.rn2.p, j.xn2.p, fc * fc, slice(j), j.nb, j.nparts, (int)((iter_ - 1) & 1), (int)iter_, 0, j.best.p,
               j.stall.p, (int)env_double("NSLX_RU_STALL", 16.0), j.rstart.p, (int)j.ru_it0, budget, s_);
    ru_residual_to_low(0, j.r64.p, j.r.p, nullptr, nullptr, 0, j.nb, len, s_);
    j.last_upd = iter_;
}

std::vector<std::unique_ptr<PoolJob>> Pool::step() {
    std::vector<std::unique_ptr<PoolJob>> out;
    if (jobs_.empty()) return out;
    run_iterations();
    int64_t hi = 0;
    for (auto& jp : jobs_) hi = std::max<int64_t>(hi, jp->rhs0 + jp->nb);
    auto read_flags = [&]() {
        NSLX_HIP(hipMemcpyAsync(done_h_.data(), done_.p, hi * sizeof(int), hipMemcpyDeviceToHost, s_));
        NSLX_HIP(hipMemcpyAsync(want_h_.data(), want_.p, hi * sizeof(int), hipMemcpyDeviceToHost, s_));
        NSLX_HIP(hipStreamSynchronize(s_));
    };
    read_flags();
    const int64_t lowmax = (int64_t)env_double("NSLX_RU_LOWMAX", 2000.0);
    const double frac = env_double("NSLX_RU_FRAC", 0.5);
    const int64_t maxgap = (int64_t)env_double("NSLX_RU_MAXGAP", 200.0);
    auto low_over = [&](const PoolJob& j) {
        return iter_ - j.ru_it0 >= std::min(lowmax, j.maxit - j.iters);
    };
    // reliable updates of the low-precision jobs that ask for one
    bool replaced = false;
    for (auto& jp : jobs_) {
        PoolJob& j = *jp;
        if (j.prec == NSLX_F64) continue;
        int64_t active = 0, waiting = 0;
        for (int64_t i = 0; i < j.nb; ++i)
            if (done_h_[j.rhs0 + i] < 0) { ++active; if (!want_h_[j.rhs0 + i]) ++waiting; }
        if (active > 0 && (waiting <= (1.0 - frac) * active || iter_ - j.last_upd >= maxgap || low_over(j))) {
            replace_residual(j);
            replaced = true;
        }
    }
    if (replaced) read_flags();
    uint64_t dirty = 0;
    std::vector<std::unique_ptr<PoolJob>> keep;
    for (auto& jp : jobs_) {
        PoolJob& j = *jp;
        bool all_done = true;
        for (int64_t i = 0; i < j.nb && all_done; ++i) all_done = done_h_[j.rhs0 + i] >= 0;
        const bool budget = j.prec == NSLX_F64 ? iter_ - j.join_iter >= j.maxit - j.iters : low_over(j);
        if (!all_done && !budget) { keep.push_back(std::move(jp)); continue; }
        const int old_cls = j.cls;
        dirty |= uint64_t(1) << old_cls;
        free_range(part_used_, j.part0, j.nb * j.nparts);
        j.part0 = -1;
        const bool was_low = j.prec != NSLX_F64;
        if (finalize(j)) {
            j.finished = true;
            free_range(rhs_used_, j.rhs0, j.nb);
            j.rhs0 = -1;
            out.push_back(std::move(jp));
        } else {
            start_recurrence(j, false, was_low);
            dirty |= uint64_t(1) << j.cls;
            keep.push_back(std::move(jp));
        }
    }
    jobs_ = std::move(keep);
    for (int c = 0; c < NCLS; ++c)
        if (dirty & (uint64_t(1) << c)) rebuild_class(c);
    return out;
}

}  // namespace nslx

namespace nslx {

int batch_solve(Ctx& ctx, int64_t nitems, const nslx_batch_item* items, double tol, int prec, int64_t maxit,
                int use_graph, nslx_batch_report* rep, hipStream_t user) {
    NSLX_REQUIRE(nitems >= 1 && items, "batch_solve: no items");
    NSLX_REQUIRE(tol > 0.0 && maxit >= 1, "batch_solve: bad tolerance or maxit");
    int64_t cap = 0;
    for (int64_t i = 0; i < nitems; ++i) {
        NSLX_REQUIRE(items[i].fm && items[i].b && items[i].x && items[i].nbatch >= 1, "batch_solve: bad item");
        NSLX_REQUIRE(&items[i].fm->m.ctx->tiles == &ctx.tiles, "batch_solve: every fm must belong to ctx");
        cap += items[i].nbatch;
    }
    const auto t0 = std::chrono::steady_clock::now();
    NSLX_HIP(hipStreamSynchronize(user));
    hipStream_t s;
    NSLX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    nslx_batch_report agg{};
    try {
        Pool pool(&ctx, s, cap, use_graph, (int)env_double("NSLX_POOL_P", 32.0));
        for (int64_t i = 0; i < nitems; ++i) {
            auto j = std::make_unique<PoolJob>();
            j->id = i;
            j->fm = &items[i].fm->m;
            j->b = items[i].b;
            j->x = items[i].x;
            j->nb = items[i].nbatch;
            j->op = items[i].op;
            j->tol = tol;
            j->prec = prec;
            j->maxit = maxit;
            pool.admit(std::move(j));
        }
        while (!pool.idle()) {
            for (auto& j : pool.step()) {
                if (j->status == NSLX_OK) ++agg.converged_items;
                agg.floor_stops += j->floor_stops;
                agg.max_rel_residual = std::max(agg.max_rel_residual, j->max_rel);
                agg.max_floor_rel = std::max(agg.max_floor_rel, j->max_floor_rel);
            }
        }
        agg.iterations = pool.stats().iterations;
        agg.replays = pool.stats().replays;
        agg.restarts = pool.stats().restarts;
        agg.escalations = pool.stats().escalations;
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
        throw;
    }
    NSLX_HIP(hipStreamSynchronize(s));
    NSLX_HIP(hipStreamDestroy(s));
    agg.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rep) *rep = agg;
    return agg.converged_items == nitems ? NSLX_OK : NSLX_NOT_CONVERGED;
}

}  // namespace nslx
