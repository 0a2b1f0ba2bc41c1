// Time-slice direct solve of the HubbardExp fermion matrix (M^{-1}, M^{-dag})
// and the batched complex GEMM it runs on.  No counterpart in the reference,
// which only applies M (hubbardExp.tpp:12-125) and solves with Krylov methods
// (CG.tpp:14-113); here the solve serves as the preconditioner of CG on
// MM^dag (tss_pcg.cpp).
//
// Algebra.  (M x)_t = x_t - A_t x_{t-1} with the transfer matrices
// A_t = b_t E diag(P_{t-1}) (b_0 = -1), so M x = y is the recurrence
// x_t = y_t + A_t x_{t-1} around the time circle.  The Nt slices are cut into
// nc chunks of c slices, [t_j, t_j + c); eliminating the inner slices leaves
// the block-cyclic system R z = w of size nc*Nx for the last slice of each
// chunk, z_j = x_{t_j + c - 1}:
//     R = I - S,   S_{j, j-1} = T_j = A_{t_j + c - 1} ... A_{t_j}   (j-1 mod nc),
// w_j = the chunk's forward recurrence from y with zero inflow.  Chunks are
// short enough (growth of T_j ~ e^{c delta ||K||} <= e^3) that T_j is well
// conditioned; R is factorised once per field by Gauss-Jordan elimination
// with partial pivoting, block column by block column: step k eliminates
// block column k from block rows k, k+1 (the only rows where it is nonzero;
// fill appears only in the last block column), which leaves
//     top rows:    [I | V_k | W_k | G_k]      bottom rows: [0 | D_{k+1} | S_{k+1} | F_k]
// where [G_k; F_k] (2Nx x 2Nx) is the whole elimination (row swaps included)
// applied to the identity.  The solve is then a forward sweep
// [a_k; v_{k+1}] = [G_k; F_k] [v_k; w_{k+1}], z_{nc-1} = D^{-1} v_{nc-1}, and a
// back sweep z_k = a_k - V_k z_{k+1} - W_k z_{nc-1}; every step is a GEMM
// with a stored matrix.  M^dag = (A'_s = A_{s+1}^H, backward recurrence) uses
// the same factorisation through R^H (chunks shifted by one slice).
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <vector>

#include "common.hpp"

namespace nslx {

struct FermionMatrix;

// One job of a batched NT GEMM launch:
//   Y[r][c] = Add[r][c] + sgn * sum_k A[r][k] B[c][k]            (PH = 0)
//   A[r][k] <- A[r][k] * ph_r[k]                                  (PH = 1)
//   Y[r][c] = Add[r][c] + sgn * conj(ph_r[c]) * sum_k A[r][k] B[c][k]   (PH = 2)
// for r < rows, c < cols, k < K.  A's row r is A0 + r*lda0 for k < kseg and
// A1 + r*lda1 (shifted by kseg) beyond; ph_r = ph + (r / phdiv) * ldph.  All
// strides count complex elements (B: real elements for a real B).  Add may
// be null, and may equal Y (each element is read and written by one thread).
// zf (jobs with Add == Y only): per 32 output columns, counted from zfoff
// columns before the job's first, whether any of those rows of B is nonzero;
// a tile whose groups are all zero is left as it is (Y = Add exactly)
// without reading or writing it.
struct BJob {
    const void* A0 = nullptr;
    const void* A1 = nullptr;
    long long lda0 = 0, lda1 = 0;
    int kseg = 1 << 30;
    int rows = 0, cols = 0, K = 0;
    const void* B = nullptr;
    long long ldb = 0;
    void* Y = nullptr;
    long long ldy = 0;
    const void* Add = nullptr;
    long long ldadd = 0;
    const void* ph = nullptr;
    long long ldph = 0;
    int phdiv = 1 << 30;
    double sgn = 1.0;
    const int* zf = nullptr;
    int zfoff = 0;
};

// Kernel variant of a launch: element precision (NSLX_F64 / NSLX_F32),
// complex or real B, phase mode, tile (0: 64 x 64, 1: 32 x 32).
struct BVariant {
    int prec = NSLX_F64;
    bool cb = true;
    int ph = 0;
    int tile = 0;
};

// Tile list of a set of jobs for a variant.
void bgemm_tiles(const std::vector<BJob>& jobs, int tile, std::vector<int4>& out);
int bgemm_tile_rows(int tile);
// flops / bytes: the launch's algorithmic work, recorded by kernel timing
void launch_bgemm(const BVariant& v, const BJob* d_jobs, const int4* d_tiles, int ntiles, hipStream_t s,
                  double flops = 0.0, double bytes = 0.0);
// Algorithmic flops (8 per complex x complex, 4 per complex x real
// multiply-add) and bytes (A, B, Y, Add and phase rows read or written once)
// of a job list.
void bgemm_work(const std::vector<BJob>& jobs, const BVariant& v, double* flops, double* bytes);

// Persistent recurrence launch (chain.hip) of steps x nchain jobs, job
// s * nchain + i = step s of chain i, all with a real B (the same B, rows,
// cols and K), step s + 1 of a chain reading as A the rows step s wrote.
// chain_supported: whether the launch fits the current device (K within the
// LDS budget, one workgroup per CU for the whole grid); sync: the plan's
// chain_sync_words(nchain, rows) counters (zero between launches); err: set
// to 1 when a workgroup gave up waiting for its group (results invalid).
bool chain_supported(const BVariant& v, int steps, int nchain, int rows, int cols, int K);
size_t chain_sync_words(int nchain, int rows);
std::string chain_class(const BVariant& v);
void launch_chain(const BVariant& v, const BJob* d_jobs, int steps, int nchain, int rows, int cols, unsigned* sync,
                  int* err, hipStream_t s, double flops = 0.0, double bytes = 0.0);

// Chunk products T_j (row i of Tt_j = T_j e_i) for every field and chunk in
// one launch (chunkprod.hip): a workgroup owns 16 rows of one chain for all
// c steps, its rows resident in LDS.  chunkprod_supported: whether the rows
// fit (two workgroups per CU); else the per-step batched GEMM is used.
// E: exp(s delta K) in prec (ld ldE), P: phases [nf][nt][n], Tt: [nf][nc][n][n].
bool chunkprod_supported(int prec, int n);
void launch_chunkprod(int prec, int n, int nt, int c, int nc, int nf, const void* E, long long ldE, const void* P,
                      void* Tt, hipStream_t s);

// A recorded sequence of launches whose job and tile arrays live in one
// device buffer each (uploaded once); replayed by run().
class BPlan {
   public:
    // append a batched GEMM over jobs (tile shape picked from the tile count
    // when tile < 0)
    void gemm(const std::vector<BJob>& jobs, BVariant v, int tile = -1);
    // append a recurrence: steps[s][i] = step s of chain i, where step s + 1
    // reads as A the rows step s wrote as Y (A0 == the previous Y, same
    // stride).  One persistent launch when allowed, chain_supported() and
    // NSLX_CHAINS is not 0, else one batched GEMM launch per step.
    void chain(const std::vector<std::vector<BJob>>& steps, BVariant v, bool allow = true);
    // append a batched GEMM launched by launch (given the uploaded jobs and
    // tiles and the resolved variant) instead of launch_bgemm: a launch that
    // runs other work beside the tiles
    using GemmLauncher = std::function<void(hipStream_t, const BJob*, const int4*, int, const BVariant&)>;
    void gemm_with(const std::vector<BJob>& jobs, BVariant v, int tile, GemmLauncher launch);
    // append any other launch
    void custom(std::function<void(hipStream_t)> fn);
    void upload(hipStream_t s);
    void run(hipStream_t s) const;
    size_t launches() const { return ops_.size(); }
    // whether a chain launch of the plan gave up waiting (clears the flag
    // and the counters; synchronises s)
    bool chain_failed(hipStream_t s);
    // the plan's give-up flag on the device (null without chain launches),
    // and the reset of flag and counters after one was seen (synchronises s)
    const int* chain_err() const { return err_.p; }
    void chain_reset(hipStream_t s);

   private:
    struct Op {
        bool is_gemm = false, is_chain = false;
        BVariant v;
        size_t job0 = 0, njobs = 0, tile0 = 0;
        int ntiles = 0;
        int steps = 0, nchain = 0, rows = 0, cols = 0;
        double flops = 0.0, bytes = 0.0;
        std::function<void(hipStream_t)> fn;
        GemmLauncher fused;
    };
    std::vector<Op> ops_;
    std::vector<BJob> jobs_;
    std::vector<int4> tiles_;
    DevBuf<BJob> djobs_;
    DevBuf<int4> dtiles_;
    DevBuf<unsigned> sync_;
    DevBuf<int> err_;
    size_t sync_words_ = 0;
    bool uploaded_ = false;
};

// Chunk length of the time-slice solve for nt slices whose one-slice growth
// is g (delta ||K|| + |delta mu| + max |Im phi|): the largest divisor c of nt
// with c g <= bound (at least 1, at most cap).  A chunk product's condition
// number is at most e^{2 c g}: bound 3 keeps it below ~400, where an FP32
// product is accurate to ~3e-5 (the direct solve's rule in both precisions);
// bound 12 keeps it below ~2.6e10, ~3e-6 relative in FP64 (enough for a
// preconditioner, with a quarter of the block steps).
constexpr double TSS_BOUND_DIRECT = 3.0, TSS_BOUND_PRECOND_F64 = 12.0;
constexpr int TSS_CAP_DIRECT = 16, TSS_CAP_PRECOND = 64;
int tss_chunk(int nt, double g, double bound = TSS_BOUND_DIRECT, int cap = TSS_CAP_DIRECT);

// Factorisation of R per field and the apply plans, in one precision.
class TimeSliceSolver {
   public:
    // Factorises R for every populated field of fm (chunk products, blocked
    // Gauss-Jordan) in precision prec (NSLX_F64 or NSLX_F32).  nc <= 0 picks
    // the chunk count from tss_chunk.
    TimeSliceSolver(FermionMatrix& fm, int prec, int nc, hipStream_t s);
    // x = M^{-1} y (op NSLX_OP_M), M^{-dag} y (NSLX_OP_MDAGGER) or
    // (M M^dag)^{-1} y (NSLX_OP_MMDAGGER) for nb RHS (nb a multiple of the
    // fields).  y and x: distinct device buffers, c128 when io_c128 (converted
    // to and from an FP32 solver's precision), else in the solver's precision.
    void apply(int op, const void* y, void* x, int64_t nb, bool io_c128, hipStream_t s);
    int prec() const { return prec_; }
    int chunks() const { return nc_; }
    int chunk_len() const { return c_; }
    int64_t fields() const { return nf_; }
    // setup launches and device bytes held
    size_t setup_launches() const { return setup_launches_; }
    size_t bytes() const;
    // generation of the populate() the factorisation belongs to
    uint64_t generation() const { return gen_; }
    // factorise again for the fields now populated (same count and chunking)
    void refactor(hipStream_t s);
    // throws when a persistent recurrence launch of an apply since the last
    // check gave up waiting (synchronises s)
    void check_passes(hipStream_t s);
    // the same in two parts around a wait the caller does anyway: the flags'
    // copies queued on s, then (after s was synchronised) the test
    void queue_pass_checks(hipStream_t s);
    void check_passes_synced(hipStream_t s);
    // throws when the last factorisation found the reduced block system
    // singular.  The factorisation only queues the flag's copy to the host;
    // synced: s has been synchronised since (else this synchronises it).
    void check_singular(hipStream_t s, bool synced);

   private:
    // plans of a product solve bound to a caller's input and output
    // buffers (through mid): no copies in or out
    struct BoundPlans {
        int op = 0;
        const void* y = nullptr;
        void* x = nullptr;
        BPlan first, second;
    };
    struct ApplyBufs {
        DevBuf<unsigned char> mid, atil, v[2], z, q, cp[2], sb;
        BPlan Minv, Mdinv;     // M^{-1}: y -> x; M^{-dag}: y -> x (both via bound buffers)
        DevBuf<unsigned char> in, out;
        std::vector<std::unique_ptr<BoundPlans>> bound;
    };
    FermionMatrix& fm_;
    int prec_, n_, nt_, nc_, c_, nf_;
    size_t esz_;
    uint64_t gen_ = 0;
    size_t setup_launches_ = 0;
    DevBuf<unsigned char> Tt_, Q_, Qfin_, PhiH_, VH_, WHcat_, DinvH_, ident_;
    std::map<int64_t, std::unique_ptr<ApplyBufs>> plans_;
    bool plans_chains_ = true;    // the chains setting plans_ were built with
    BPlan setup_plan_;
    bool setup_built_ = false;
    const void* e_used_ = nullptr;
    const void* p_used_ = nullptr;
    DevBuf<int> gj_piv_, gj_sing_, gj_zf_;
    PinBuf<int> sing_h_;
    bool sing_pending_ = false;
    PinBuf<int> pass_h_;                 // queued give-up flags
    std::vector<BPlan*> pass_plans_;     // the plans they belong to
    DevBuf<unsigned char> gj_binv_, gj_C_, gj_YT_;
    void setup(hipStream_t s);
    void build_setup(hipStream_t s);
    void gj(BPlan& p, unsigned char* Q, long long ldq, long long strideQ, int rows, int width);
    ApplyBufs& bufs(int64_t nb, hipStream_t s);
    void build_minv(BPlan& p, const void* y, void* x, ApplyBufs& B, int64_t nb);
    void build_mdinv(BPlan& p, const void* y, void* x, ApplyBufs& B, int64_t nb);
    const void* E() const;
    const void* P() const;
};

}  // namespace nslx
