// Time-slice hop of the HubbardExp fermion matrix as one MFMA GEMM.
//
//   (M psi)_t    = psi_t - B_t  E (P_{t-1} (.) psi_{t-1}),   B_0 = -1
//   (M^dag psi)_t = psi_t - conj(P_t) (.) (B'_t E psi_{t+1}),  B'_{Nt-1} = -1
//
// (reference: src/NSL/FermionMatrix/Impl/hubbardExp.tpp:12-75).  The field
// block (nbatch, nt, nx) is the GEMM's row space (nbatch*nt rows of nx
// sites); E (nx x nx, real) is the shared right operand.  The phase multiply
// is done in the tile loader, the time shift is a row gather, and the sign /
// identity / conj(P) terms are fused into the store epilogue.  CG modes fuse
// the Krylov recurrences into the same passes (see MODE_* below).
#pragma once

#include "common.hpp"

namespace nslx {

enum HopOp { HOP_M = 0, HOP_MD = 1 };

// Iteration i of the fused CG on A = outer(inner(.)) is two launches:
// CGA_i: rr_{i-1} = sum of the |r|^2 tile partials of CGB_{i-1} (rpart[(i-1)&1]),
//        rr_{i-2} likewise; stop b if rr_{i-1} <= tol2 (not in reliable-update
//        mode); beta = rr_{i-1}/rr_{i-2} (0 for i = 1); p_new = X(=r) +
//        beta p_old (written once, by column tile 0); Y(=u) = inner(p_new);
//        |u|^2 tile partials -> upart.
// CGB_i: alpha = rr_{i-1} / sum(upart); t = outer(X(=u)) in registers;
//        x += alpha p; r -= alpha t; |r|^2 tile partials -> rpart[i&1].
// CGI:   t = outer(X(=u)); r = b - t; p = r; |r|^2 partials -> rpart[0].
// PLAIN: Y = op(X).
// Each workgroup reduces the few partials of its own right-hand side in a
// fixed order, so no cross-workgroup synchronisation is needed inside a launch.
enum HopMode { MODE_PLAIN = 0, MODE_CGA = 1, MODE_CGB = 2, MODE_CGI = 3 };

// Device-side state of the fused CG (one slot per right-hand side; in
// batch-global mode slot 0 carries the batch decision).
struct CgDev {
    double* upart = nullptr;                 // [nb][nparts] |u|^2 tile partials
    double* rpart[2] = {nullptr, nullptr};   // ping-pong |r|^2 tile partials
    double* tol2 = nullptr;                  // stop when rr <= tol2
    int* done = nullptr;                     // -1 active, else the converging iteration
    int global = 0;
    // reliable-update mode: CGA does not stop a right-hand side; it sets
    // want[b] = 1 once rr_{i-1} <= delta2 * rr_ref[b] (rr_ref: |r|^2 of the last
    // FP64 residual replacement) or rr_{i-1} <= tol2[b].
    int ru = 0;
    double delta2 = 0.01;
    double* rr_ref = nullptr;
    int* want = nullptr;
};

// Discretisation of the fermion matrix behind a hop launch.
//   DISC_EXP  (HubbardExp):  (M psi)_t = psi_t - b_t E (P_{t-1} . psi_{t-1}),      E = exp(s delta K)
//   DISC_DIAG (HubbardDiag): (M psi)_t = psi_t - E psi_t - b_t P_{t-1} . psi_{t-1}, E = s delta K
// (b_0 = -1, else +1; M^dagger correspondingly with conj(P_t) and psi_{t+1}).
enum HopDisc { DISC_EXP = 0, DISC_DIAG = 1 };

struct HopArgs {
    int nx = 0, nt = 0, nb = 0, rpf = 1;   // rpf: right-hand sides per field
    const void* E = nullptr;   // E[y][x] in compute precision (BF16X3: hi plane then lo plane)
    const void* P = nullptr;   // phases [nfield][nt][nx], complex, field precision
    const void* X = nullptr;
    void* Y = nullptr;
    const void* Pold = nullptr;
    void* Pnew = nullptr;
    void* Xs = nullptr;
    void* R = nullptr;
    const void* Bv = nullptr;
    CgDev cg;
    int iter = 0;
    int disc = 0;              // DISC_EXP or DISC_DIAG (classic tiles only)
};

// One problem of a grouped launch: a field block with its own lattice size,
// E, phases and Krylov vectors (vector slots v[], selected per launch by
// GroupArgs::role).  nrt/nct must match the launch's tile shape.  rhs0 and
// part0 offset this problem's right-hand sides in the CG scalar arrays and
// its tiles in the partial-sum array.
struct HopProblem {
    int nx = 0, nt = 0, nb = 0, rpf = 1;
    int nrt = 0, nct = 0, ldE = 0, rhs0 = 0;
    long long planeE = 0, part0 = 0;
    const void* E = nullptr;
    const void* P = nullptr;
    void* v[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

// Vector slots of HopProblem::v used by the batched solver.
enum VecSlot { VS_R = 0, VS_P0 = 1, VS_P1 = 2, VS_U = 3, VS_X = 4, VS_B = 5, VS_IN = 6, VS_OUT = 7 };

struct GroupArgs {
    const HopProblem* probs = nullptr;   // device array
    int nprob = 0;
    const int4* tiles = nullptr;         // device array of (problem, b, rt, ct)
    int ntiles = 0;
    // slot index (or -1) for the roles X, Y, Pold, Pnew, Xs, R, Bv
    signed char role[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    CgDev cg;
    int iter = 0;
    // when set, the iteration number is *iter_dev + iter and the tile count is
    // read from device memory (so that one launch sequence built as a graph
    // can be replayed, iter then being the iteration's offset in the sequence)
    const int* iter_dev = nullptr;
    const int* ntiles_dev = nullptr;
};

// Layout of E in device memory: row y (output site), column x, row stride
// e_ld(nx), e_rows(nx) rows, zero padded so that every tile of the menu reads
// whole 16-byte vectors in bounds.  BF16X3 stores the hi plane, then the lo
// plane, each e_rows*e_ld elements.
inline int e_ld(int nx) { return (nx + 63) / 64 * 64; }
inline int e_rows(int nx) { return (nx + 191) / 192 * 192; }

// Tile menu entry (rows x cols of the output block, K chunk).
struct TileInfo {
    int id, bm, bn, bk, nthreads;
};

// Launch one hop.  prec: NSLX_F64 / NSLX_F32 / NSLX_BF16X3; op: HOP_M/HOP_MD;
// tile: index into tile_menu(prec).  Throws on bad arguments.
void launch_hop(int prec, int op, int mode, const HopArgs& a, int tile, hipStream_t s);
// Grouped persistent launch over g.ntiles tiles with `grid` workgroups; every
// problem of one launch has the discretisation disc (DISC_DIAG: the CG pair
// HOP_MD/MODE_CGA and HOP_M/MODE_CGB only).
void launch_hop_grouped(int prec, int op, int mode, int tile, int disc, const GroupArgs& g, int grid, hipStream_t s);
// The kernel launch_hop_grouped would launch (function and workgroup size),
// for building graph kernel nodes; it takes one GroupArgs by value and no
// dynamic LDS.
struct GroupedKernel {
    const void* fn;
    int nthreads;
};
GroupedKernel hop_grouped_kernel(int prec, int op, int mode, int tile, int disc);
// Host-side tile list for problems (nrt/nct filled in from the tile shape):
// problems with larger nx first (longer K loops), and the column tiles of a
// row block 8 entries apart so that a persistent grid that is a multiple of 8
// keeps them on one XCD.
std::vector<int4> build_tile_list(std::vector<HopProblem>& probs, int prec, int tile);
const std::vector<TileInfo>& tile_menu(int prec);
// Whether tile `tile` of tile_menu(prec) can run a lattice of nx sites (the
// slice tiles keep a column slice of E in LDS, which bounds nx).
bool tile_fits(int prec, int tile, int nx);
int hop_nparts(int prec, int tile, int nx, int nt);
// Heuristic tile choice when no tuned entry exists.
int default_tile(int prec, int nx, int nt, int64_t nb);

}  // namespace nslx

namespace nslx {
// Pack E (nx x nx, ld nx, FP64) into the padded per-precision layouts the hop
// kernels read: E64 (double), E32 (float), Ebf (bf16 hi plane, then lo plane
// with lo = bf16(E - hi)).  Any output pointer may be null.
void pack_E(int nx, const double* E, void* E64, void* E32, void* Ebf, hipStream_t s);
}  // namespace nslx
