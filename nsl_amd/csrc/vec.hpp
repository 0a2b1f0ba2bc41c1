// Vector (BLAS-1 style) kernels with per-right-hand-side deterministic
// reductions.  A field batch is nb vectors of len = nt*nx complex elements.
// Reductions write per-chunk partials and a second launch sums them in index
// order, so results do not depend on scheduling or on the rest of the
// batch.
#pragma once

#include "common.hpp"
#include "hop.hpp"

namespace nslx {

// Scratch for vec reductions: one partial (2 doubles) per chunk of a RHS.
struct VecScratch {
    DevBuf<double> part;       // nb * nchunk * 2
    void ensure(int64_t nb, int64_t len, hipStream_t s);
};

// P = exp(i s phi + s mu~), Pinv = exp(-i s phi - s mu~) in FP64, plus the
// FP32 copy of P; mu~ = mut + i mut_im.  phi: c128 [n].  (hubbardExp.hpp:116-128)
void phase_tables(const void* phi, int64_t n, int sgn, double mut, double mut_im, void* P64, void* Pinv64,
                  void* P32, hipStream_t s);

// out[0] = max |Im phi| over n c128 values (device); part: 256 doubles of scratch.
void max_abs_im(const void* phi, int64_t n, double* part, double* out, hipStream_t s);

// out[b] = sum |v_b|^2 (double), S = double (c128) or float (c64).
void norm2_perb(int prec_is_f64, const void* v, int64_t nb, int64_t len, double* out,
                VecScratch& sc, hipStream_t s);
// tol2[b] = tol*tol*(relative ? bn2[b] : 1); in batch-global mode slot 0 only.
void set_tol2(const double* bn2, double* tol2, int64_t nb, double tol, int relative, int global,
              hipStream_t s);
// Defect correction helpers (FP64 outer loop around a low-precision solve).
// r64 = b - t64; rn2[b] = |r64_b|^2.
void resid64(const void* b, const void* t, void* r, int64_t nb, int64_t len, double* rn2,
             VecScratch& sc, hipStream_t s);
// r32 = r64 / |r64_b| for b with rn2[b] > otol2[b] (else 0).
void scale_to32(const void* r64, void* r32, const double* rn2, const double* otol2, int64_t nb,
                int64_t len, hipStream_t s);
// x64 += |r64_b| * d32 for b with rn2[b] > otol2[b].
void update64(void* x64, const void* d32, const double* rn2, const double* otol2, int64_t nb,
              int64_t len, hipStream_t s);
// Reliable-update CG (FP64 residual replacement around a CG recurrence in
// the "low" storage: c64 for F32/BF16X3, c128 for F64).
// rs = low(r64) for every RHS; with init also p = rs and xs = 0.
void ru_residual_to_low(int low_is_f64, const void* r64, void* rs, void* p, void* xs, int init, int64_t nb,
                        int64_t len, hipStream_t s);
// x64 += xs, xs = 0.
void ru_accumulate(int low_is_f64, void* x64, void* xs, int64_t nb, int64_t len, hipStream_t s);
// Per-RHS bookkeeping after a true residual rn2 (single block): a RHS is
// converged when rn2 <= tol2 or rn2 <= floor2 * xn2 (the attainable accuracy
// of an FP64 residual, (c eps ||A||)^2 |x|^2; xn2 may be null); converged RHS
// get done = iter, the others restart their reference |r|^2 and
// their |r|^2 partials (rpart[cur], nparts per RHS) are replaced by rn2 so
// that the next CGA continues the recurrence from the true residual.
// With stall_max > 0 (a low-precision phase) a RHS gets done = RU_ESCALATE
// (the kernels skip it; the solver finishes it in FP64) when its true |r|^2
// exceeds 100x the recursive |r|^2 held in rpart[cur] (the low-precision
// recurrence has lost track of the residual), or when its true |r|^2 has not
// fallen below a quarter of its best value over stall_max replacements that
// it asked for (want[b] set: its recursive |r| fell by delta), or
// when budget > 0 and, 100 or more iterations after the phase started at
// iteration it0, its average rate so far (from rstart, the |r|^2 at the
// phase start) projects more than budget iterations to |r|^2 <= tol^2.
// init re-initialises the RHS with done < 0 or done == RU_ESCALATE (and
// records rstart).  best/stall/rstart: nb entries.
constexpr int RU_ESCALATE = 1 << 30;
void ru_scalars(const double* rn2, const double* xn2, double floor2, const CgDev& cg, int64_t nb, int nparts,
                int cur, int iter, int init, double* best, int* stall, int stall_max, double* rstart, int it0,
                int budget, hipStream_t s);

// Element type conversions.
void c128_to_c64(const void* in, void* out, int64_t n, hipStream_t s);
void c64_to_c128(const void* in, void* out, int64_t n, hipStream_t s);

// Generic per-RHS complex dots: out[2b], out[2b+1] = sum conj(a_b) b_b.
// If global, out[0..1] = sum over all b (fixed order).
void dot_perb(int prec_is_f64, const void* a, const void* bv, int64_t nb, int64_t len, double* out,
              int global, VecScratch& sc, hipStream_t s);

// z = x + c_b * y  with complex per-b scalar c (cs[2b], cs[2b+1]); sign = +1/-1
// applied to c; if global, c = cs[0..1] for all b.  x, y, z may alias
// elementwise (same index).
void axpy_perb(int prec_is_f64, const void* x, const void* y, void* z, const double* cs, double sign,
               int64_t nb, int64_t len, int global, const int* done, hipStream_t s);

}  // namespace nslx
