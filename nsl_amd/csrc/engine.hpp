// Engine objects behind the C ABI: device context, fermion matrix
// (HubbardExp) with its memoised exponential hopping matrices, solvers.
#pragma once

#include <cmath>
#include <map>
#include <memory>
#include <mutex>

#include "common.hpp"
#include "dd.hpp"
#include "dense.hpp"
#include "hop.hpp"
#include "tss.hpp"
#include "vec.hpp"

namespace nslx {

// Tuned tile choice per (precision, nx, nt, nb class); filled by the
// autotuner (tune.cpp) or by default_tile().
struct TileTable {
    std::mutex mu;
    std::map<std::tuple<int, int, int, int>, int> best;
    int get(int prec, int nx, int nt, int64_t nb);
    void set(int prec, int nx, int nt, int64_t nb, int tile);   // tile < 0 clears
    bool lookup(int prec, int nx, int nt, int64_t nb, int* tile);
    static int nb_class(int64_t nb);
};

struct Ctx {
    int device = 0;
    int num_cu = 0;
    TileTable tiles;
};

// E(a) = exp(a K) in the padded per-precision layouts (hop.hpp).
struct EEntry {
    DevBuf<double> e64;     // padded FP64
    DevBuf<float> e32;      // padded FP32
    DevBuf<uint16_t> ebf;   // padded BF16 hi + lo planes
    DevBuf<double> dense;   // nx x nx FP64, ld nx (for the dense kernels)
};

// Device workspace of the solvers, kept on the fermion matrix between calls.
struct SolverWork {
    DevBuf<unsigned char> r, p0, p1, u, x;          // Krylov vectors (storage precision)
    DevBuf<unsigned char> r64, t64, r32, x32;       // defect correction
    DevBuf<unsigned char> v1, v2, v3, v4, v5;       // BiCGStab extra vectors
    DevBuf<double> rr, alpha, beta, tol2, bn2, rn2, otol2, part, scal;
    DevBuf<int> done, status, want;
    DevBuf<double> upart, rpart0, rpart1;           // fused-CG tile partials
    std::vector<double> last_rr;                    // |r|^2 per slot after the last iteration
    DevBuf<double> rr_ref, rr_prev, xn2, best, rstart;
    DevBuf<unsigned char> xlo;                      // low part of a double-double x
    DevBuf<int> stall;
    PinBuf<int> status_h;
    PinBuf<double> dbl_h;
    PinBuf<int> int_h, status_h2;
    void ensure_scalars(int64_t nb, int64_t nparts, hipStream_t s);
};

struct FermionMatrix {
    Ctx* ctx = nullptr;
    int disc = DISC_EXP;             // HubbardExp or HubbardDiag (hop.hpp)
    int nx = 0, nt = 0;
    double beta = 1.0, mu = 0.0, mu_im = 0.0;   // mu = mu + i mu_im (complex chemical potential)
    int species = NSLX_PARTICLE;
    std::vector<double> K_host;
    double normK = 0.0;              // max row sum of |K|
    DevBuf<double> K;
    std::map<double, std::unique_ptr<EEntry>> memo;

    int64_t nfield = 0;
    double max_im_phi = 0.0;        // max |Im phi| over the populated fields
    DevBuf<double> mim_part;        // its device reduction (256 block maxima + result)
    PinBuf<double> mim_h;
    DevBuf<double2> phi;            // [nfield][nt][nx]
    DevBuf<double2> P64, Pinv64;    // phases
    DevBuf<float2> P32;

    VecScratch vsc;
    DevBuf<unsigned char> work;     // apply temporaries
    SolverWork sw;
    uint64_t pop_gen = 0;           // incremented by every populate()
    // time-slice direct solvers (FP64, FP32) of the populated fields, built on
    // first use and refactorised after a populate (tss.hip).  nc <= 0 picks
    // the chunk count by tss_chunk: the direct solve's bound, or, for a
    // preconditioner (precond) in FP64, the wider one.
    std::unique_ptr<TimeSliceSolver> tss[2];
    // whether the time-slice applies may use the persistent recurrence
    // launch (which needs the device to itself); off while several streams
    // drive this engine at once (batched and served time-slice solves)
    bool chains_ok = true;
    TimeSliceSolver& time_slice_solver(int prec, int nc, bool precond, hipStream_t s);
    // delta ||K|| + |delta mu| + max |Im phi|: the log of a slice's growth bound
    double slice_growth() const;
    DdWork ddw;

    int sgn() const { return species == NSLX_PARTICLE ? +1 : -1; }
    double delta() const { return beta / nt; }
    // |delta mu|: the largest change of log |P| the chemical potential makes
    double mut_abs() const { return std::hypot(delta() * mu, delta() * mu_im); }
    int64_t len() const { return (int64_t)nt * nx; }

    EEntry& E(double a, hipStream_t s);      // memoised exp(a K)
    EEntry& Elin(double a, double id, hipStream_t s);   // memoised id I + a K (HubbardDiag)
    std::map<std::pair<double, double>, std::unique_ptr<EEntry>> memo_lin;
    // the hop's dense matrix: exp(s delta K) (EXP) or s delta K (DIAG)
    const void* Eptr(int prec, hipStream_t s);
    // whether Eptr would find it memoised (no allocation)
    bool has_hop_matrix() const {
        return disc == DISC_DIAG ? memo_lin.count({sgn() * delta(), 0.0}) != 0 : memo.count(sgn() * delta()) != 0;
    }
    const void* Pptr(int prec) const { return prec == NSLX_F64 ? (const void*)P64.p : (const void*)P32.p; }

    void populate(const void* phi_dev, int64_t nf, int sp, hipStream_t s);
    void require_populated() const;
    // one hop (PLAIN) with the fm's E and phases
    void hop(int prec, int op, const void* in, void* out, int64_t nb, hipStream_t s);
    void apply(int op, const void* in, void* out, int64_t nb, int prec, hipStream_t s);
    HopArgs hop_args(int prec, int64_t nb, hipStream_t s);
    // tuned or default tile; HubbardDiag runs on the classic tiles 0-3 only
    int tile(int prec, int64_t nb) {
        const int t = ctx->tiles.get(prec, nx, nt, nb);
        return (disc == DISC_DIAG && t >= 4) ? default_tile(prec, nx, nt, nb) : t;
    }
};

// HubbardExp(species, lat, Nt, beta, mu) on ctx's device (fm.cpp)
void fm_init(FermionMatrix& m, Ctx* ctx, int nx, int nt, const double* K, double beta, double mu, int species,
             int disc = DISC_EXP);

// Solvers (solver.cpp)
int solve(FermionMatrix& fm, int op, int method, const void* b, void* x, const void* x0, int64_t nb, double tol,
          int tol_kind, int64_t maxit, int prec, int check, nslx_solve_report* rep, hipStream_t s);
// Reliable-update CG (recurrence in prec) with x in double-double and the
// replacements' true residual in double-double near the FP64 floor, to
// |b - A (x + xlo)| <= tol |b| per RHS (xlo may be null).
int solve_dd(FermionMatrix& fm, int op, const void* b, void* x, void* xlo, const void* x0, int64_t nb, double tol,
             int64_t maxit, int prec, nslx_solve_report* rep, hipStream_t s);
// Outer CG on fm preconditioned by an inner CG on fp (CGpreconditioned).
int solve_precond(FermionMatrix& fm, FermionMatrix& fp, int op, const void* b, void* x, const void* x0, int64_t nb,
                  double tol, int tol_kind, int64_t maxit, int inner_prec, nslx_solve_report* rep, hipStream_t s);

// CG (for M, M^dag) and BiCGStab (any op) in FP64 with every recurrence
// scalar on the device (krylov.hip); the reference's semantics with
// NSLX_TOL_ABS_BATCH, per-RHS scalars and stops otherwise.
int solve_generic_dev(FermionMatrix& fm, int op, int method, const void* b, void* x, const void* x0, int64_t nb,
                      double tol, int tol_kind, int64_t maxit, nslx_solve_report* rep, hipStream_t s);

// CG on MMdagger preconditioned by the time-slice solve in precision pprec
// (FP64 / FP32), double-double defect correction to |b - A x| <= tol |b| per
// RHS; x = high part, xlo (device c128 or null) = low part (tss_pcg.hip)
int solve_tss(FermionMatrix& fm, int op, const void* b, void* x, void* xlo, const void* x0, int64_t nb, double tol,
              int64_t maxit, int pprec, int nchunks, nslx_solve_report* rep, hipStream_t s);

// Batched time-slice-preconditioned solves over nstreams worker streams
// (tss_batch.cpp); report fields as nslx_batch_solve_tss documents.
int batch_solve_tss(Ctx& ctx, int64_t nitems, const nslx_batch_item* items, void* const* x_lo, double tol,
                    int pprec, int64_t maxit, int nstreams, nslx_batch_report* rep);

// log det / force (logdet.cpp)
void fm_logdet(FermionMatrix& fm, int64_t field, int nchunks, double* out2, hipStream_t s);
void fm_grad_logdet(FermionMatrix& fm, int64_t field, void* out, hipStream_t s);

// Allocation-free application (noalloc.hip): noalloc_init looks up the
// current device's scratch array (called by nslx_ctx_create); apply_noalloc
// is nslx_fm_apply with the two-hop intermediate in that array.
void noalloc_init(int device);
void apply_noalloc(FermionMatrix& fm, int op, const void* in, void* out, int64_t nb, int prec, hipStream_t s);

// hop tile autotuner (tune.cpp)
int fm_tune(FermionMatrix& fm, int prec, int64_t nb, int reps, int mode, int* best_tile, double* best_us,
            hipStream_t s);
// The tile the table holds for (prec, nx, nt, nb), tuned in CG mode on first
// use when NSLX_AUTOTUNE is on and the block is large (tune.cpp).
int tuned_tile(FermionMatrix& fm, int prec, int64_t nb, hipStream_t s);

// set by NSLX_INNER_TOL etc.; see solver.cpp
double env_double(const char* name, double dflt);

}  // namespace nslx

struct nslx_ctx_s {
    nslx::Ctx c;
};
struct nslx_fm_s {
    nslx::FermionMatrix m;
};
