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

// PLAIN: Y = op(X).
// CGA:   p_new = X(=r) + beta_b * p_old (written once, by column tile 0),
//        Y(=u) = op(p_new), partial |u|^2; last block: alpha_b = rr_b/|u_b|^2.
// CGB:   t = op(X(=u)) kept in registers; x += alpha_b p; r -= alpha_b t;
//        partial |r|^2; last block: beta_b, rr_b, done_b.
// CGI:   t = op(X(=u)); r = b - t; p = r; partial |r|^2; last block: rr_b,
//        done_b, beta_b = 0.
enum HopMode { MODE_PLAIN = 0, MODE_CGA = 1, MODE_CGB = 2, MODE_CGI = 3 };

// Device-side scalar state of the fused CG (one slot per right-hand side; in
// batch-global mode slot 0 carries the batch value).
struct CgDev {
    double* rr = nullptr;       // current |r|^2
    double* alpha = nullptr;
    double* beta = nullptr;
    double* tol2 = nullptr;     // stop when rr <= tol2
    int* done = nullptr;        // -1 active, else iteration count at convergence
    int* status = nullptr;      // [0] active rhs, [1] last iteration finalised
    double* part = nullptr;     // per-tile partial sums [nb][nparts]
    unsigned* counter = nullptr;// last-block counter, zero between launches
    int global = 0;
};

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
const std::vector<TileInfo>& tile_menu(int prec);
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
