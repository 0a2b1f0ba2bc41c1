// Ragged batched CG over many fermion matrices (different Nx, Nt, fields) in
// one grouped hop launch per phase per Krylov iteration; the iteration loop
// is built once as a HIP graph and replayed.  Jobs join and leave between
// replays (continuous batching).  Used by nslx_batch_solve and by the
// correlator server.
#pragma once

#include <memory>
#include <vector>

#include "engine.hpp"

namespace nslx {

struct PoolJob {
    int64_t id = 0;
    FermionMatrix* fm = nullptr;
    const void* b = nullptr;     // device c128 [nb][len]
    void* x = nullptr;           // device c128 [nb][len], the solution
    int64_t nb = 0;
    double tol = 1e-12;          // per RHS: |b - A x| <= tol |b| (or the FP64 attainable accuracy)
    int prec = NSLX_F64;         // precision of the current recurrence
    int op = NSLX_OP_MMDAGGER;   // MMdagger or MdaggerM
    int64_t maxit = 100000;
    // -- pool bookkeeping --
    int cls = -1, tile = -1, nparts = 0;
    int64_t rhs0 = -1, part0 = -1;
    int64_t join_iter = 0, iters = 0;
    int restarts = 0;
    bool escalated = false;
    bool finished = false;
    int status = NSLX_OK;
    double max_rel = 0.0;         // final max_b |b - A x|_b / |b_b| (FP64)
    double max_floor_rel = 0.0;
    int64_t floor_stops = 0;
    DevBuf<unsigned char> r, p0, p1, u, xs;   // recurrence vectors (job precision)
    DevBuf<double2> r64, t64;                 // FP64 residual work
    DevBuf<double> bn2, rn2, xn2;             // per-RHS norms
    // reliable updates of a low-precision recurrence (ru_scalars state)
    DevBuf<double> best, rstart;
    DevBuf<int> stall;
    int64_t ru_it0 = 0, last_upd = 0;
    size_t bytes() const;
};

struct PoolStats {
    int64_t iterations = 0;       // global iterations run
    int64_t replays = 0;
    int64_t restarts = 0;
    int64_t escalations = 0;
};

class Pool {
   public:
    Pool(Ctx* ctx, hipStream_t s, int64_t cap_rhs, int use_graph, int iters_per_replay);
    ~Pool();
    // admit a job (device work on the pool's stream); the pool takes ownership
    void admit(std::unique_ptr<PoolJob> job);
    // run iters_per_replay iterations, then retire / restart finished jobs;
    // returns the jobs that completed in this step
    std::vector<std::unique_ptr<PoolJob>> step();
    bool idle() const { return jobs_.empty(); }
    size_t active() const { return jobs_.size(); }
    int64_t free_rhs() const;
    // reserve a contiguous run of nb RHS slots for a job about to be admitted
    // (its PoolJob::rhs0; -1 if no run is free: fragmented slots that count
    // enough by number do not suffice), or give such a run back unused.  The
    // partial-sum slots grow on demand.
    int64_t reserve_rhs(int64_t nb);
    void release_rhs(int64_t start, int64_t nb);
    hipStream_t stream() const { return s_; }
    const PoolStats& stats() const { return st_; }

   private:
    // class = (operator, discretisation, precision, classic tile 0-3):
    // ((opb * 2 + disc) * 3 + prec) * 4 + tile, opb = 1 for MdaggerM systems
    static constexpr int NCLS = 48;
    struct Cls {
        DevBuf<HopProblem> probs;
        DevBuf<int4> tiles;
        DevBuf<int> ntiles;
        size_t cap_probs = 0, cap_tiles = 0;
        int grid = 0;
        int njobs = 0;
    };
    Ctx* ctx_;
    hipStream_t s_;
    int64_t cap_rhs_;
    int use_graph_;
    int P_;
    DevBuf<double> tol2_, rr_ref_, upart_, rpart0_, rpart1_;
    DevBuf<int> done_, want_, it_dev_;
    std::vector<char> rhs_used_;
    std::vector<char> part_used_;
    int64_t cap_parts_ = 0;
    Cls cls_[NCLS];
    std::vector<std::unique_ptr<PoolJob>> jobs_;
    std::vector<int> done_h_, want_h_;
    int64_t iter_ = 1;                 // next global iteration (device counter mirror)
    uint64_t cls_mask_ = 0;            // classes in the graph
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    PoolStats st_;

    static int cls_of(int prec, int tile, int disc, int op) {
        return (((op == NSLX_OP_MDAGGERM ? 1 : 0) * 2 + (disc & 1)) * 3 + prec % 3) * 4 + (tile & 3);
    }
    static int prec_of(int c) { return (c / 4) % 3; }
    static int tile_of(int c) { return c % 4; }
    static int disc_of(int c) { return (c / 12) % 2; }
    static bool mdagm_of(int c) { return c >= 24; }
    int64_t alloc_range(std::vector<char>& used, int64_t n);
    void free_range(std::vector<char>& used, int64_t start, int64_t n);
    void start_recurrence(PoolJob& j, bool fresh, bool keep_done);
    void grow_parts(int64_t need);
    void rebuild_class(int c);
    void run_iterations();
    void drop_graph();
    bool finalize(PoolJob& j);   // true if the job is complete
    void replace_residual(PoolJob& j);
    CgDev cgdev() const;
    CgDev slice(const PoolJob& j) const;   // cgdev() offset to the job's RHS and partial slots
};

}  // namespace nslx

namespace nslx {
// nslx_batch_solve (see include/nslx.h)
int batch_solve(Ctx& ctx, int64_t nitems, const nslx_batch_item* items, double tol, int prec, int64_t maxit,
                int use_graph, nslx_batch_report* rep, hipStream_t user);
}  // namespace nslx
