// Two-point correlator (TwoPointCorrelator::measure) on the device; see corr.hip.
#pragma once

#include "engine.hpp"

namespace nslx {

// corr_out: device c128 (nt, nx, nx), corr[t][y][x].  max_rhs bounds the
// right-hand sides of one batched solve in the per-RHS mode (<= 0: 4096).
int fm_two_point(FermionMatrix& fm, int64_t ntsrc, double eps, int64_t maxit, int tol_kind, int prec, int64_t max_rhs,
                 void* corr_out, nslx_solve_report* rep, hipStream_t s);

}  // namespace nslx
