/*
 * nslx.h — C ABI of the MI355X-native HubbardExp fermion-matrix and
 * Krylov-solver engine (library: nsl_amd/libnslx.so).
 *
 * The reference (Marcel-Rodekamp/NSL) is a header-only C++ template library
 * on libtorch and has no C ABI; every entry point below replaces a C++
 * member function or free function of the reference, cited on its
 * declaration (paths relative to the reference's repository root).
 *
 * Conventions
 *  - Complex fields are interleaved (re, im) pairs, row-major with shape
 *    (nbatch, nt, nx), x fastest: the byte layout of a contiguous
 *    NSL::Tensor<complex<double>> of shape (..., Nt, Nx)
 *    (src/NSL/Tensor/Impl/base.tpp:42-59).  c128 = 2 x double, c64 = 2 x float.
 *  - Field pointers are DEVICE pointers on the context's device unless a
 *    parameter says "host".  `stream` is a hipStream_t passed as void*
 *    (NULL = the null stream); work is stream-ordered.
 *  - Every function returns an nslx_status.  On error nslx_last_error()
 *    returns a message for the calling thread.  NSLX_NOT_CONVERGED is not a
 *    hard error: the solution buffer holds the last iterate, as the
 *    reference's solvers return x after logging (CG.tpp:90-93).
 *  - A handle is not re-entrant (like the reference's objects, which hold
 *    mutable scratch); distinct handles may be used from distinct threads.
 */
#ifndef NSLX_H
#define NSLX_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */
typedef enum {
    NSLX_OK = 0,
    NSLX_E_ARG = 1,          /* invalid argument (shape, enum, null)         */
    NSLX_E_DEVICE = 2,       /* HIP runtime error                            */
    NSLX_E_OOM = 3,          /* device allocation failed                     */
    NSLX_E_STATE = 4,        /* call order (e.g. apply before populate)      */
    NSLX_NOT_CONVERGED = 5,  /* solver hit maxIter; x holds the last iterate */
    NSLX_E_INTERNAL = 6
} nslx_status;

/* MatrixCombination (src/NSL/LinAlg/Solver/applicationType.tpp:18) */
typedef enum { NSLX_OP_M = 0, NSLX_OP_MDAGGER = 1, NSLX_OP_MMDAGGER = 2,
               NSLX_OP_MDAGGERM = 3 } nslx_op;

/* NSL::Hubbard::Species (src/NSL/Action/Implementations/hubbard.tpp:7) */
typedef enum { NSLX_PARTICLE = 0, NSLX_HOLE = 1 } nslx_species;

/* arithmetic of an apply / of a solver's inner iteration */
typedef enum {
    NSLX_F64 = 0,      /* complex128 fields, FP64 MFMA                        */
    NSLX_F32 = 1,      /* complex64 fields, FP32 MFMA, FP32 accumulate        */
    NSLX_BF16X3 = 2    /* complex64 fields, E.X as 3 BF16 MFMA products
                          (hi*hi + hi*lo + lo*hi), FP32 accumulate            */
} nslx_precision;

typedef enum { NSLX_CG = 0, NSLX_BICGSTAB = 1 } nslx_method;

typedef enum {
    /* the reference's criterion: one alpha/beta for the whole batch and an
       absolute stop sum_all |r|^2 <= tol^2 on the recursive residual
       (CG.hpp:34, CG.tpp:70, inner_product.tpp:31-33)                      */
    NSLX_TOL_ABS_BATCH = 0,
    /* per right-hand side alpha/beta; stop each RHS when its FP64 true
       relative residual |b_k - A x_k| <= tol |b_k| (reliable updates), or
       at the FP64 attainable accuracy c eps ||A|| |x_k| when that lies above
       tol |b_k| -- such a floor stop is reported (report.floor_stops) and
       the call returns NSLX_NOT_CONVERGED (nslx_solve_dd reaches tol)      */
    NSLX_TOL_REL_PER_RHS = 1,
    /* as NSLX_TOL_REL_PER_RHS, but a floor stop counts as converged          */
    NSLX_TOL_REL_PER_RHS_FLOOR = 2
} nslx_tol_kind;

typedef struct nslx_ctx_s* nslx_ctx;
typedef struct nslx_fm_s* nslx_fm;

/* ---- library ------------------------------------------------------------ */
int         nslx_version(void);
const char* nslx_status_string(int status);
const char* nslx_last_error(void);

/* Device context: device ordinal, tuned tile table, scratch pool.
 * Replaces NSL::Device (src/NSL/device.tpp:22-126) for this path. */
int nslx_ctx_create(int device, nslx_ctx* out);
int nslx_ctx_destroy(nslx_ctx ctx);

/* ---- host RNG (Philox-4x32-10, Salmon et al. SC'11) -----------------------
 * The reference draws from torch's global generator
 * (src/NSL/Tensor/Impl/factory.tpp:38-41, src/NSL/MarkovChain/HMC.tpp:189-195);
 * the engine uses a counter RNG so that CPU and GPU runs see the same numbers.
 * Block i of stream (seed, offset) = philox(ctr=(i, off_lo, off_hi, 0),
 * key=(seed_lo, seed_hi)).  Host memory. */
int nslx_rng_philox4x32(uint64_t seed, uint64_t offset, int64_t nblocks, uint32_t* out4);
int nslx_rng_uniform(uint64_t seed, uint64_t offset, int64_t n, double* out);
int nslx_rng_normal(uint64_t seed, uint64_t offset, int64_t n, double* out);

/* ---- lattices (host) -----------------------------------------------------
 * K is written row-major nx*nx (real, symmetric); sites (optional, may be
 * NULL) nx*2 (honeycomb) or nx*3 (ring, nanotube). */
/* Honeycomb<T>(L, kappa): src/NSL/Lattice/Implementations/honeycomb.tpp:72-138 */
int nslx_lattice_honeycomb(int L0, int L1, double kappa, double* K, double* sites);
/* Ring<T>(n, kappa): src/NSL/Lattice/Implementations/ring.tpp:8-42 */
int nslx_lattice_ring(int n, double kappa, double* K, double* sites);
/* (n,n) armchair nanotube with open ends: no counterpart in the reference
 * (project lattice for mixed-size batches); nx = 4*n*cells. */
int nslx_lattice_armchair_nanotube(int n, int cells, double kappa, double* K, double* sites);

/* ---- fermion matrix (HubbardExp) -----------------------------------------
 * Replaces NSL::FermionMatrix::HubbardExp<complex<double>, Lattice>.
 * ctor (species, lat, Nt, beta, mu): src/NSL/FermionMatrix/Impl/hubbardExp.hpp:64-76.
 * K: host, nx*nx real (the lattice's hopping matrix, lattice.tpp:46-49).
 * E(s*delta) = exp(s*delta*K) is computed in FP64 on the device (Taylor with
 * scaling and squaring) and memoised per exponent, like
 * SpatialLattice::exp_hopping_matrix (lattice.tpp:51-58).
 * beta = beta_re + i beta_im and mu = mu_re + i mu_im stand for the
 * reference's `const Type & beta, const Type & mu` with Type complex.  mu
 * may be complex (it enters only the phases P = exp(i s phi + s delta mu),
 * hubbardExp.hpp:116-128, so an imaginary chemical potential is a shift of
 * the field).  beta must be real: beta_im != 0 returns NSLX_E_ARG (a complex
 * delta makes exp(s delta K) complex and M^dag's E(s conj(delta)) differ
 * from it, hubbardExp.tpp:62-65, which the engine's real-E kernels do not
 * hold). */
int nslx_fm_create(nslx_ctx ctx, int nx, int nt, const double* K, double beta_re, double beta_im,
                   double mu_re, double mu_im, int species, nslx_fm* out);

/* Replaces NSL::FermionMatrix::HubbardDiag<complex<double>, Lattice>, ctor
 * (species, lat, Nt, beta, mu): src/NSL/FermionMatrix/Impl/hubbardDiag.hpp.
 * The hop's dense matrix is s*delta*K on the same time slice and the phases
 * sit on the time shift: (M psi)_t = psi_t - s delta K psi_t
 * - b_t P_{t-1} psi_{t-1} (hubbardDiag.tpp:10-36), M^dagger :39-72, MMdagger
 * and MdaggerM as M(M^dagger .) and M^dagger(M .) (:75-165 expand to exactly
 * these).  Every other nslx_fm_* call, nslx_solve, nslx_solve_precond,
 * nslx_correlator_two_point and the HMC entries take the handle; logdet
 * follows hubbardDiag.tpp:166-190 (i s sum(phi) + log det(1 + prod_t
 * (1 - s delta K) diag(Pinv_{t+1})), stable chunked form, no mu term like the
 * reference) and grad_logdet :193-252.  nslx_batch_solve takes HubbardDiag
 * handles too (mixed with HubbardExp ones), and a server request selects its
 * discretisation (nslx_request.discretisation).  beta and mu as in
 * nslx_fm_create (complex mu, real beta). */
int nslx_fm_create_diag(nslx_ctx ctx, int nx, int nt, const double* K, double beta_re, double beta_im,
                        double mu_re, double mu_im, int species, nslx_fm* out);
int nslx_fm_destroy(nslx_fm fm);

/* populate(phi[, species]): hubbardExp.hpp:103-128.  phi: device c128,
 * nfield fields of shape (nt, nx), contiguous.  RHS k of an apply/solve with
 * nbatch RHS uses field k / (nbatch / nfield).  species < 0 keeps the current
 * species. */
int nslx_fm_populate(nslx_fm fm, const void* phi, int64_t nfield, int species, void* stream);

/* M, Mdagger, MMdagger, MdaggerM applied to nbatch fields
 * (fermionMatrix.hpp:40-56; hubbardExp.tpp:38-125).  in/out: device, shape
 * (nbatch, nt, nx); c128 for NSLX_F64, c64 for NSLX_F32 / NSLX_BF16X3.
 * in and out must not overlap.  nbatch must be a multiple of nfield. */
int nslx_fm_apply(nslx_fm fm, int op, const void* in, void* out, int64_t nbatch,
                  int precision, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): nslx_fm_apply with the same arguments, results and
 * kernels, allocating nothing.  The MMdagger / MdaggerM intermediate lives in
 * a library-owned device array of NSLX_NOALLOC_SCRATCH_BYTES per GPU (held
 * until the process ends); NSLX_E_ARG when nbatch*nt*nx*(16 for NSLX_F64,
 * else 8) exceeds it.  Calls that overlap on one GPU share that array and
 * race.  NSLX_E_STATE before populate (which builds the hop matrix) or while
 * kernel timing (nslx_prof_enable) is on. */
#define NSLX_NOALLOC_SCRATCH_BYTES (64ull << 20)
int nslx_fm_apply_noalloc(nslx_fm fm, int op, const void* in, void* out, int64_t nbatch,
                          int precision, void* stream);

/* logDetM() for field `field` (hubbardExp.tpp:128-150), FP64.  Computed
 * as log det of the block-cyclic matrix of chunked slice products with
 * partial pivoting (stable where the reference's det overflows); nchunks <= 0
 * picks the count from the spectral growth.  out[0] = log|det M|, out[1] =
 * arg det M in (-pi, pi]. Synchronous. */
int nslx_fm_logdet(nslx_fm fm, int64_t field, int nchunks, double* out_host2, void* stream);

/* gradLogDetM() for field `field` (hubbardExp.tpp:152-211), FP64.
 * out: device c128 (nt, nx). */
int nslx_fm_grad_logdet(nslx_fm fm, int64_t field, void* out, void* stream);

/* ---- solvers --------------------------------------------------------------- */
typedef struct {
    int64_t iterations;        /* max over RHS of inner (Krylov) iterations   */
    int64_t iterations_total;  /* sum of inner iterations over outer steps     */
    int64_t outer_steps;       /* FP64 defect-correction steps (0 for F64)    */
    int64_t converged_rhs;     /* RHS that met the criterion                   */
    double  rsqr;              /* final |r|^2 (batch sum, recursive, FP64 run) */
    double  max_rel_residual;  /* max_k |b_k - A x_k| / |b_k|, FP64, if checked */
    double  seconds;           /* host wall time of the call                    */
    /* NSLX_TOL_REL_PER_RHS only: RHS that stopped at the attainable accuracy
     * of an FP64 residual (|r| <= c eps ||A|| |x|, above tol |b|), and the
     * largest such bound relative to |b| over the RHS                        */
    int64_t floor_stops;
    double  max_floor_rel;
    /* reliable-update solvers: sum over iterations of the right-hand sides
     * still active, in the low precision and in FP64 (counted at the
     * solver's polls of the device flags, every few iterations)             */
    int64_t rhs_iters_low;
    int64_t rhs_iters_f64;
} nslx_solve_report;

/* CG::operator()(b[, x0]) / BiCGStab::operator()(b[, x0]) on the operator
 * `op` of fm (Solver.hpp:86-96; CG.tpp:14-113; BiCGStab.tpp:14-150).
 * b, x: device c128 (nbatch, nt, nx).  x0: device c128 or NULL (NULL means
 * x0 = b, as the reference).  `precision` != NSLX_F64 runs the inner solve
 * in that precision inside an FP64 defect-correction loop (CG only,
 * Hermitian ops only) driven to the FP64 true relative residual tol; it
 * requires NSLX_TOL_REL_PER_RHS.  check_true_residual != 0 fills
 * max_rel_residual. */
int nslx_solve(nslx_fm fm, int op, int method, const void* b, void* x, const void* x0,
               int64_t nbatch, double tol, int tol_kind, int64_t maxit, int precision,
               int check_true_residual, nslx_solve_report* report, void* stream);

/* CG::operator()(b[, x0]) on MMdagger / MdaggerM (CG.tpp:14-113) per RHS to
 * |b_k - A (x_k + x_lo_k)| <= tol |b_k| with the solution in double-double:
 * the reliable-update recurrence of nslx_solve (NSLX_TOL_REL_PER_RHS) in
 * `precision`, its increments accumulated into x + x_lo with TwoSum, and,
 * once an FP64 true residual comes within 1e4 of the FP64 floor
 * eps ||A|| |x|, every later residual replacement evaluated in double-double
 * (nslx_dd_residual).  Reaches targets that an FP64 x cannot represent (the
 * ill-conditioned MM^dag of random fields at large beta).  x: the solution
 * rounded to FP64; x_lo (device c128 or NULL): the remainder.  x0 NULL
 * means x0 = 0.  HubbardExp only.  report.max_rel_residual is the final
 * true residual (double-double once switched), report.floor_stops = 0. */
int nslx_solve_dd(nslx_fm fm, int op, const void* b, void* x, void* x_lo, const void* x0, int64_t nbatch,
                  double tol, int64_t maxit, int precision, nslx_solve_report* report, void* stream);

/* Preconditioned CG.  Replaces NSL::LinAlg::CGpreconditioned<Type>
 * (src/NSL/LinAlg/Solver/Impl/CGpreconditioned.hpp:34-41 (std::function M,
 * Mprec), :83-91 (two fermion matrices + MatrixCombination), :133-141,
 * :183-190; solve_ CGpreconditioned.tpp:14-100, operator() :103-118): outer CG
 * on op of fm, each z = P^-1 r an inner CG on op of fm_prec started from r
 * with the same tol and maxit (the reference's innerCG_).  x0 NULL means
 * x0 = b.  tol_kind NSLX_TOL_ABS_BATCH reproduces the reference (batch-global
 * inner products, |r|^2 <= tol^2, both levels); NSLX_TOL_REL_PER_RHS runs
 * both levels per RHS.  inner_precision NSLX_F32 / NSLX_BF16X3 runs the inner
 * solves with the reliable-update solver (then op must be MMDAGGER or
 * MDAGGERM and tol_kind NSLX_TOL_REL_PER_RHS).  fm_prec may be fm itself; it
 * must have fm's Nx, Nt and device.  Report: iterations = outer iterations,
 * iterations_total = inner iterations summed over all inner solves,
 * outer_steps = number of inner solves, max_rel_residual = FP64 true
 * residual max_k |b_k - A x_k| / |b_k| (always computed).  Returns
 * NSLX_NOT_CONVERGED (x written) like nslx_solve. */
int nslx_solve_precond(nslx_fm fm, nslx_fm fm_prec, int op, const void* b, void* x, const void* x0,
                       int64_t nbatch, double tol, int tol_kind, int64_t maxit, int inner_precision,
                       nslx_solve_report* report, void* stream);

/* ---- time-slice direct solve and the preconditioned solver -------------------
 * No counterpart in the reference (it applies M, hubbardExp.tpp:12-125, and
 * solves only by Krylov iteration, CG.tpp:14-113).  M x = y is the recurrence
 * x_t = y_t + b_t E (P_{t-1} . x_{t-1}) around the time circle; cutting the
 * Nt slices into chunks of c (c delta ||K|| <= 3) and eliminating the inner
 * slices leaves a block-cyclic system of Nt/c blocks of Nx, factorised per
 * field by Gauss-Jordan elimination with partial pivoting (HubbardExp only).
 *
 * nslx_fm_solve_direct: out = M^{-1} in (op NSLX_OP_M), M^{-dag} in
 * (NSLX_OP_MDAGGER), (M M^dag)^{-1} in (NSLX_OP_MMDAGGER) or (M^dag M)^{-1}
 * in (NSLX_OP_MDAGGERM) for nbatch RHS;
 * precision NSLX_F64 (c128 in/out) or NSLX_F32 (c64 in/out).  The first call
 * after a populate factorises (nchunks <= 0: chosen from the growth per slice).
 * Synchronous.  The in-chunk recurrences run as one persistent launch per
 * pass where the device holds its grid (NSLX_CHAINS=0: one launch per slice);
 * NSLX_E_INTERNAL if such a launch stopped waiting for its workgroups. */
int nslx_fm_solve_direct(nslx_fm fm, int op, const void* in, void* out, int64_t nbatch, int precision,
                         int nchunks, void* stream);

/* CG on MMdagger or MdaggerM (op; the CG of CG.tpp:14-113 on
 * MatrixCombination MMdagger / MdaggerM, Solver.hpp:86-96), preconditioned
 * by the time-slice solve (of the same product) in
 * precond_precision (NSLX_F64 or NSLX_F32: the preconditioner's factorisation
 * and applies on FP32 matrix cores, the CG recurrence in FP64), wrapped in a
 * double-double defect correction: every RHS stops when
 * |b - MM^dag (x + x_lo)| <= tol |b| with the residual evaluated in
 * double-double on the FP64 operator.  This reaches targets below the FP64
 * floor eps ||MM^dag|| |x| (where |x| / |b| is large); x receives the
 * double-double solution rounded to FP64, x_lo (device c128, may be NULL) the
 * remainder.  x0 NULL means x0 = 0.  report: iterations = inner CG
 * iterations, iterations_total = preconditioner applications, outer_steps =
 * double-double residual evaluations after the start, max_rel_residual =
 * max_k |b_k - A (x_k + x_lo_k)| / |b_k| (double-double).  NSLX_NOT_CONVERGED
 * (x written) when maxit inner iterations did not suffice.  nchunks <= 0
 * picks the chunk length c from the growth per slice g = delta ||K|| +
 * |delta mu| + max |Im phi|: the largest divisor of Nt with c g <= 3 (at most
 * 16) for an FP32 preconditioner, c g <= 12 (at most 64) for an FP64 one,
 * whose wider chunks keep a near-exact preconditioner with fewer block
 * steps. */
int nslx_solve_tss(nslx_fm fm, int op, const void* b, void* x, void* x_lo, const void* x0, int64_t nbatch,
                   double tol, int64_t maxit, int precond_precision, int nchunks,
                   nslx_solve_report* report, void* stream);

/* Double-double residual (a measurement aid; no counterpart in the
 * reference): r = b - A (x + x_lo), A = MMdagger or MdaggerM of a HubbardExp
 * fm, every product and sum in double-double on the FP64 numbers of
 * exp(s delta K) and of the phase table; rel_host[k] = |r_k| / |b_k|.  x_lo
 * and r may be NULL.  Synchronous. */
int nslx_dd_residual(nslx_fm fm, int op, const void* b, const void* x, const void* x_lo, void* r,
                     double* rel_host, int64_t nbatch, void* stream);

/* The FP64 operator data the engine applies (for checking residuals on the
 * host bit for bit): E_host = exp(s delta K) (nx*nx, row-major) of the
 * current species, P_host = the phase table P = exp(i s phi + s delta mu),
 * c128 (nfield, nt, nx).  Either may be NULL.  Synchronous. */
int nslx_fm_export(nslx_fm fm, double* E_host, void* P_host, void* stream);

/* ---- batched ragged solve ----------------------------------------------------------
 * No single counterpart in the reference (it solves one HubbardExp at a
 * time, CG.tpp:14-113); this is the reference's CG on MMdagger or MdaggerM
 * (MatrixCombination, Solver.hpp:86-96) for many fermion matrices at once.
 * Each item is a populated fm (own lattice, Nt, beta, fields; HubbardExp or
 * HubbardDiag) with its own b and x (device c128, nbatch x nt x nx) and its
 * operator (op: NSLX_OP_MMDAGGER, or NSLX_OP_MDAGGERM for HubbardExp items);
 * all fms must belong to ctx.  Every Krylov iteration is one grouped
 * persistent hop launch per phase over the tiles of all items (per operator /
 * discretisation / precision / tile class);
 * with use_graph the iteration loop is built once as a HIP graph and
 * replayed.  Per-RHS stop: |b - A x| <= tol |b| in FP64 or the FP64 attainable
 * accuracy (c eps ||A|| |x|).  precision F32 / BF16X3 runs the low-precision
 * recurrence with FP64 reliable updates (x accumulated in FP64, the true
 * residual replacing the recursive one at replay boundaries); RHS whose
 * low-precision recurrence loses track, stalls or projects past its budget
 * are finished from x in FP64.
 * An item's iterates do not depend on the other items.  Synchronous. */
typedef struct {
    nslx_fm     fm;
    const void* b;
    void*       x;
    int64_t     nbatch;
    int         op;         /* NSLX_OP_MMDAGGER or NSLX_OP_MDAGGERM */
} nslx_batch_item;

typedef struct {
    int64_t iterations;         /* global Krylov iterations run               */
    int64_t replays;            /* graph replays (or launch rounds)           */
    int64_t restarts;           /* FP64 residual restarts over all items      */
    int64_t escalations;        /* items switched from low precision to FP64  */
    int64_t converged_items;
    int64_t floor_stops;        /* RHS stopped at the attainable accuracy     */
    double  max_rel_residual;   /* max over RHS of |b - A x| / |b|, FP64      */
    double  max_floor_rel;
    double  seconds;
} nslx_batch_report;

int nslx_batch_solve(nslx_ctx ctx, int64_t nitems, const nslx_batch_item* items, double tol, int precision,
                     int64_t maxit, int use_graph, nslx_batch_report* report, void* stream);

/* Time-slice-preconditioned batch (no counterpart in the reference, whose
 * systems are solved one CG at a time, CG.tpp:14-113): every item's MMdagger
 * x = b (or MdaggerM x = b, item.op) is solved as nslx_solve_tss does (CG preconditioned by the item's
 * own time-slice solve in precond_precision, double-double correction to
 * |b - A (x + x_lo)| <= tol |b| per RHS; x receives the FP64 rounding,
 * x_lo[i] (NULL, or an array of nitems device c128 buffers shaped like x,
 * entries may be NULL) the remainder).
 * Items (HubbardExp, each fermion matrix at most once) are taken largest
 * first by nstreams host threads (<= 0: 8), each driving its own stream, so
 * the items' factorisations and applies overlap on the device; with more
 * than one stream the in-chunk recurrences run one launch per slice.  An
 * item's result does not depend on the others.  report: iterations = max
 * inner iterations over the items, replays = preconditioner applications,
 * restarts = double-double residual evaluations, converged_items,
 * max_rel_residual (double-double), seconds.  Synchronous (the device is
 * synchronised at entry). */
int nslx_batch_solve_tss(nslx_ctx ctx, int64_t nitems, const nslx_batch_item* items, void* const* x_lo,
                         double tol, int precond_precision, int64_t maxit, int nstreams,
                         nslx_batch_report* report);

/* ---- measurements ----------------------------------------------------------------
 * TwoPointCorrelator::measure(N)
 * (src/NSL/Measurements/Hubbard/twoPointCorrelationFunction.tpp:116-162) on
 * fm's populated field (nfield must be 1; species as populated): point
 * sources S[x][tsrc][x] = 1 at tsrc = 0, Nt/N, 2 Nt/N, ... < Nt, X = CG on
 * MMdagger to eps, G = Mdagger X, G shifted by -tsrc in time with the
 * antiperiodic sign, corr[t][y][x] += G[x][t][y], corr /= N.
 * corr: device c128 (nt, nx, nx).  tol_kind NSLX_TOL_ABS_BATCH reproduces
 * the reference (one CG per time source over its nx sources, x0 = b,
 * sum |r|^2 <= eps^2); NSLX_TOL_REL_PER_RHS solves up to max_rhs sources per
 * batch (<= 0: 4096) with per-RHS relative stops and allows `precision`
 * F32 / BF16X3 (reliable updates, FP64 true residual).  report: aggregate
 * over the solves (iterations = max).  Synchronous. */
int nslx_correlator_two_point(nslx_fm fm, int64_t n_tsrc, double eps, int64_t maxit, int tol_kind,
                              int precision, int64_t max_rhs, void* corr, nslx_solve_report* report,
                              void* stream);

/* The measurement loop of TwoPointCorrelator::measure()
 * (twoPointCorrelationFunction.tpp:241-315) computes measure(N) once per
 * configuration of an ensemble; here every populated field of a HubbardExp
 * fm is one configuration and all of them are measured together: the point
 * sources of every field go through one time-slice-preconditioned CG on
 * MMdagger (nslx_solve_tss: one factorisation per field, every RHS to
 * |b - MM^dag x| <= tol |b| in double-double), then G = Mdagger X, re-centred
 * and averaged per field as measure(N) does.  corr: device c128 (nfield,
 * nt, nx, nx).  max_rhs bounds the RHS of one solve (<= 0: 8192; one time
 * source of every field at least).  precond_precision NSLX_F64 or NSLX_F32.
 * No counterpart in the reference (which measures one configuration at a
 * time; the loop over the stored configurations is nsl_amd.measure_ensemble).
 * Synchronous. */
int nslx_correlator_two_point_fields(nslx_fm fm, int64_t n_tsrc, double tol, int64_t maxit,
                                     int precond_precision, int64_t max_rhs, void* corr,
                                     nslx_solve_report* report, void* stream);

/* TwoPointCorrelator::measureK(k, N) (twoPointCorrelationFunction.tpp:169-233)
 * for the k-th set of wall sources: wall = wallSources[k], device c128
 * (nw, nx).  For tsrc = 0, Nt/N, ... < Nt: S[s][tsrc][x] = wall[s][x],
 * X = CG on MMdagger, G = Mdagger X shifted by -tsrc (antiperiodic),
 * corrK += G; corrK /= N; corrKblock[t][s1][s2] = sum_x conj(wall[s1][x])
 * corrK[s2][t][x].  corrK: device c128 (nw, nt, nx) and block: device c128
 * (nt, nw, nw); either may be NULL.  tol_kind / precision as
 * nslx_correlator_two_point (NSLX_TOL_REL_PER_RHS batches all time sources
 * in one solve).  Synchronous. */
int nslx_correlator_two_point_wall(nslx_fm fm, const void* wall, int64_t nw, int64_t n_tsrc, double eps,
                                   int64_t maxit, int tol_kind, int precision, void* corrK, void* block,
                                   nslx_solve_report* report, void* stream);

/* Momentum quadruples of TwoBodyCorrelator::measure(N)
 * (twoBodyCorrelationFunction.tpp:139-151, eq_modBZ_ :121-135): every
 * (w, x, z, y) in [0, kdim)^4, looped w, x, y, z, with k_w + k_x = k_z + k_y
 * modulo the honeycomb reciprocal vectors (4 pi/3, 0), (-2 pi/3, 2 pi/sqrt 3)
 * (to 2e-12).  momenta: host (kdim, 2).  *nquad = count; quads (host int,
 * 4 per quadruple, in key order w, x, z, y) is written when not NULL and
 * *nquad on entry is at least the count. */
int nslx_two_body_quads(const double* momenta, int kdim, int64_t* nquad, int* quads);

/* TwoBodyCorrelator::measure(N) (twoBodyCorrelationFunction.tpp:244-300,
 * channels :481-1000) on fm's populated field: wall = wallSources, device c128
 * (kdim, bdim, nx).  For tsrc = 0, Nt/N, ... < Nt and species particle, hole:
 * CG on MMdagger of the wall sources and on MdaggerM of their conjugates,
 * back-multiplied by Mdagger / M, re-centred on tsrc, projected on the walls
 * (corrPool_, corrPoolDag_), then the twelve channels I1S1Iz1Sz1,
 * I1S1Iz1Sz0, I1S1Iz1Szn1, I1S1Iz0Sz1, I1S1Iz0Szn1, I1S1Izn1Sz1,
 * I1S1Izn1Sz0, I1S1Izn1Szn1, I0S1Iz0Sz1, I0S1Iz0Szn1, I1S0Iz1Sz0,
 * I1S0Izn1Sz0 accumulate into out: device c128 (12, nquad, nt, bdim^2,
 * bdim^2), element [ch][q][t][i*bdim+j][l*bdim+k] for quadruple q =
 * quads[4q..4q+3] = (w, x, z, y) (the key of the reference's maps).  The
 * field's species is restored on return.  tol_kind / precision as
 * nslx_correlator_two_point.  Synchronous. */
int nslx_correlator_two_body(nslx_fm fm, const void* wall, int kdim, int bdim, int64_t nquad, const int* quads,
                             int64_t n_tsrc, double eps, int64_t maxit, int tol_kind, int precision, void* out,
                             nslx_solve_report* report, void* stream);

/* ---- correlator server (continuous batching) ----------------------------------
 * No counterpart in the reference (its measurements run one configuration
 * at a time, twoPointCorrelationFunction.tpp:116-162).  A server owns a
 * worker thread and a batched CG pool on its own stream.  A request is a
 * lattice (K, Nx), Nt, beta, mu, species, one field phi (host c128 (nt, nx),
 * copied at submit) and a number of time sources N; its result is the
 * two-point correlator of TwoPointCorrelator::measure(N) (N x Nx point
 * sources, MMdagger solves to tol per RHS, M^dag back-multiply, antiperiodic
 * re-centring, average), kept on the device until fetched.
 * Requests wait in FIFO order and are admitted while their device footprint
 * fits hbm_budget and their RHS fit max_rhs; admitted requests join the
 * running batch at the next Krylov-iteration boundary (every
 * iters_per_replay iterations) and leave it as they finish, so requests that
 * arrive while others are on the GPU share every hop launch with them.  A
 * request's result does not depend on what else is in the batch.
 * Latency = completion - submit, host clock.  Thread-safe. */
typedef struct nslx_server_s* nslx_server;

/* Fermion matrix of a request: HubbardExp (hubbardExp.hpp) or HubbardDiag
 * (hubbardDiag.hpp), as nslx_fm_create / nslx_fm_create_diag. */
typedef enum { NSLX_DISC_EXP = 0, NSLX_DISC_DIAG = 1 } nslx_discretisation;

typedef struct {
    int64_t max_rhs;            /* RHS slots of the pool (<= 0: 65536)          */
    double  hbm_budget;         /* bytes of device memory for admitted work     */
    int     use_graph;          /* replay the iteration loop as a HIP graph     */
    int     iters_per_replay;   /* even; <= 0: 32                               */
    /* > 0: time-slice-preconditioned serving (no counterpart in the
     * reference) with that many worker streams instead of the Krylov pool.
     * Each admitted request is solved as nslx_solve_tss does (CG on MMdagger
     * preconditioned by its own time-slice solve, double-double correction
     * to tol per RHS; request.precision = the preconditioner's, F64 or F32;
     * HubbardExp only), on the first free worker whose stream then runs its
     * factorisation, solve and correlator beside the other workers'
     * requests; admission is FIFO while the footprint (factorisation
     * included) fits hbm_budget.  iterations in the result = inner
     * preconditioned iterations; use_graph and iters_per_replay unused. */
    int     precond_workers;
} nslx_server_config;

typedef struct {
    int           nx, nt;
    const double* K;            /* host, nx*nx, copied at submit                */
    double        beta, mu;
    int           species;
    const void*   phi;          /* host c128 (nt, nx), copied at submit         */
    int64_t       n_tsrc;       /* time sources N (1 <= N <= nt)                */
    double        tol;          /* per-RHS relative FP64 residual               */
    int           precision;    /* NSLX_F64 / NSLX_F32 / NSLX_BF16X3            */
    int64_t       maxit;
    int           discretisation; /* NSLX_DISC_EXP (HubbardExp, 0) or NSLX_DISC_DIAG */
} nslx_request;

typedef struct {
    int     status;             /* NSLX_OK, NSLX_NOT_CONVERGED or an error      */
    int64_t iterations;         /* Krylov iterations of the request             */
    int     escalated;          /* low precision gave way to FP64               */
    double  max_rel_residual;   /* FP64 true residual, max over its RHS         */
    double  queue_ms;           /* submit -> admitted                           */
    double  latency_ms;         /* submit -> result ready                       */
} nslx_result;

typedef struct {
    int64_t submitted, completed, rhs_solved, admitted_peak;
    double  p50_ms, p99_ms, max_ms;   /* latency over completed requests          */
    double  busy_seconds;             /* worker time with work on the device      */
    double  bytes_peak;               /* largest admitted footprint               */
    int64_t iterations, replays;
} nslx_server_stats;

int nslx_server_create(nslx_ctx ctx, const nslx_server_config* cfg, nslx_server* out);
int nslx_server_destroy(nslx_server srv);     /* waits for admitted work, drops the rest */
int nslx_server_submit(nslx_server srv, const nslx_request* req, int64_t* id);
/* Block until request id is done (timeout_ms < 0: no limit); on success copy
 * the correlator (device c128 (nt, nx, nx), corr[t][y][x]) to corr (may be
 * NULL) and release it.  NSLX_E_STATE on timeout. */
int nslx_server_wait(nslx_server srv, int64_t id, double timeout_ms, void* corr, nslx_result* result);
int nslx_server_stats_get(nslx_server srv, nslx_server_stats* stats);

/* ---- HMC ---------------------------------------------------------------------
 * Hubbard action S = S_gauge + S_fermion and its force for one real field
 * (src/NSL/Action/Implementations/hubbardGaugeAction.tpp:59-67,
 * hubbardFermiAction.tpp:68-113; U~ = (beta/Nt) U, hubbard.tpp:10-19).
 * phi: device c128 (nt, nx).  The fm handle is re-populated. */
int nslx_hmc_action(nslx_fm fm, const void* phi, double U, int nchunks, double* out_host2,
                    void* stream);
int nslx_hmc_force(nslx_fm fm, const void* phi, double U, void* force_out, void* stream);

/* One HMC trajectory with LeapfrogRealForce (leapfrogRealForce.tpp:35-74)
 * and the Metropolis step of HMC::generate_ (HMC.tpp:185-241).  phi (in/out)
 * and momentum (in): device c128 (nt, nx), momentum real-valued; S0 = action
 * at phi (host, 2 doubles).  On accept phi is replaced and S0 updated. */
typedef struct {
    double dH_re, dH_im;       /* H1 - H0                                      */
    double acc_prob;           /* exp(Re(H0 - H1))                             */
    int    accepted;
    double S_re, S_im;         /* action of the returned field                 */
    double seconds;
} nslx_hmc_report;

int nslx_hmc_trajectory(nslx_fm fm, void* phi, const void* momentum, double* S0_host2,
                        double U, double traj_length, int nmd, double uniform,
                        int nchunks, nslx_hmc_report* report, void* stream);

/* ntraj trajectories from phi with momenta and uniforms drawn from Philox:
 * momentum of trajectory k = nslx_rng_normal(seed, 2k, nt*nx), uniform =
 * nslx_rng_uniform(seed, 2k+1, 1).  reports: host array of ntraj. */
int nslx_hmc_chain(nslx_fm fm, void* phi, double U, double traj_length, int nmd,
                   int64_t ntraj, uint64_t seed, int nchunks, nslx_hmc_report* reports,
                   void* stream);

/* HMC::generate<Chain>(state, Nconf, saveFrequency) (src/NSL/MarkovChain/HMC.tpp:57-155)
 * without its HDF5 read/write, with the MarkovState record of
 * src/NSL/MarkovChain/markovState.tpp:13-96.  The start state is phi (device
 * c128 (nt, nx), in; on return the last state's configuration) at markov time
 * markov_time0, with its action computed here.  NSLX_CHAIN_ALL: states[0] =
 * the start state (accepted = 1, acc_prob = 1) and states[n], n < nconf, the
 * state after saveFrequency more trajectories (HMC.tpp:98-137);
 * NSLX_CHAIN_LAST: Nconf * saveFrequency - 1 trajectories and states[0] =
 * the last state (:144-152).  save_frequency <= 0 counts as 1 (:61-63).
 * The trajectory that makes markov time m + 1 takes its momenta from
 * nslx_rng_normal(seed, 2m, nt*nx) and its uniform from
 * nslx_rng_uniform(seed, 2m+1, 1), so a chain resumed from any saved state
 * (its configuration and markov time) continues bit for bit.  states: host
 * array (nconf for ALL, 1 for LAST) or NULL; configs: device c128 (nconf, nt,
 * nx) with every saved configuration, or NULL. */
typedef enum { NSLX_CHAIN_ALL = 0, NSLX_CHAIN_LAST = 1 } nslx_chain;
typedef struct {
    double  S_re, S_im;        /* actionValue                                 */
    double  acc_prob;          /* acceptanceProbability                       */
    int64_t markov_time;       /* markovTime                                  */
    int     accepted;
    double  dH_re, dH_im;      /* H1 - H0 of the trajectory that made the state */
} nslx_markov_state;
int nslx_hmc_generate(nslx_fm fm, void* phi, int64_t markov_time0, double U, double traj_length,
                      int nmd, int64_t nconf, int64_t save_frequency, int chain, uint64_t seed,
                      int nchunks, void* configs, nslx_markov_state* states, void* stream);

/* ---- hop tiles and autotuning ------------------------------------------------
 * No counterpart in the reference (its GEMMs are whatever torch::matmul
 * dispatches, Tensor/Impl/matmul.tpp:15-17).  The hop kernels come in tile
 * variants (ids 0 .. count-1 per precision); the context keeps one choice per
 * (precision, nx, nt, power-of-two class of nbatch), filled by nslx_fm_tune
 * or nslx_ctx_set_tile, else a built-in heuristic.
 * nslx_fm_tune times, on fm's populated fields and nbatch vectors, either
 * reps x (M + Mdagger) plain applies for every tile that fits nx
 * (NSLX_TUNE_APPLY), or reps iterations of the fused CG on MMdagger -- the
 * CGA and CGB launches the solvers and the batching engine run -- for the
 * classic tiles 0-3 (NSLX_TUNE_CG), records the fastest in fm's context and
 * returns it (best_us: mean time of one pair).  The solvers (nslx_solve's
 * fused CG, nslx_batch_solve) tune in CG mode on their own when the table has
 * no entry for a block of at least 512 workgroups of the 32 x 96 tile
 * (environment NSLX_AUTOTUNE=0 turns that off).  Synchronous; allocates
 * temporary vectors. */
typedef enum { NSLX_TUNE_APPLY = 0, NSLX_TUNE_CG = 1 } nslx_tune_mode;
int nslx_tile_count(int precision, int* count);
int nslx_ctx_set_tile(nslx_ctx ctx, int precision, int nx, int nt, int64_t nbatch, int tile);
int nslx_ctx_get_tile(nslx_ctx ctx, int precision, int nx, int nt, int64_t nbatch, int* tile);
int nslx_fm_tune(nslx_fm fm, int precision, int64_t nbatch, int reps, int mode, int* best_tile, double* best_us,
                 void* stream);

/* ---- kernel timing ----------------------------------------------------------
 * No counterpart in the reference (its only profiling is a host stopwatch,
 * src/NSL/logger.hpp:174-191).  nslx_prof_enable(n), n > 0: every n-th
 * launch of each kernel class (hop, batched GEMM, double-double hop) is
 * bracketed by HIP events on its own stream (n = 1: every launch; 0 turns
 * timing off).  nslx_prof_read writes one line per kernel class,
 * "name launches total_ms timed flops bytes\n" (launches counted;
 * milliseconds, algorithmic flops and bytes summed over the timed launches;
 * flops and bytes are 0 for the hop, whose work the caller knows), into buf
 * (NUL terminated) and returns the number of classes.  Process-wide; not for
 * concurrent use. */
int nslx_prof_enable(int on);
int nslx_prof_reset(void);
int nslx_prof_read(char* buf, int64_t size);

#ifdef __cplusplus
}
#endif
#endif /* NSLX_H */
