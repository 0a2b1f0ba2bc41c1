// Two-point correlator (TwoPointCorrelator::measure) on the device; see corr.hip.
#pragma once

#include <vector>

#include "engine.hpp"

namespace nslx {

// corr_out: device c128 (nt, nx, nx), corr[t][y][x].  max_rhs bounds the
// right-hand sides of one batched solve in the per-RHS mode (<= 0: 4096).
int fm_two_point(FermionMatrix& fm, int64_t ntsrc, double eps, int64_t maxit, int tol_kind, int prec, int64_t max_rhs,
                 void* corr_out, nslx_solve_report* rep, hipStream_t s);

// measure(N) for every populated field of fm (nfield configurations at
// once), solved by the time-slice-preconditioned CG to tol per RHS:
// corr_out (nfield, nt, nx, nx).  max_rhs bounds the RHS of one solve
// (<= 0: 8192; at least one time source per solve).
int fm_two_point_fields(FermionMatrix& fm, int64_t ntsrc, double tol, int64_t maxit, int pprec, int64_t max_rhs,
                        void* corr_out, nslx_solve_report* rep, hipStream_t s);

// TwoPointCorrelator::measureK(k, N) for one set of nw wall sources
// (wall: device c128 nw x nx): corrK_out (nw, nt, nx) and block_out
// (nt, nw, nw), either may be null (corr_wall.hip).
int fm_two_point_wall(FermionMatrix& fm, const void* wall, int64_t nw, int64_t ntsrc, double eps, int64_t maxit,
                      int tol_kind, int prec, void* corrK_out, void* block_out, nslx_solve_report* rep,
                      hipStream_t s);
// Momentum quadruples (w, x, z, y) with k_w + k_x = k_z + k_y modulo the
// honeycomb reciprocal lattice, in the reference's loop order.
void two_body_quads(const double* momenta, int kdim, std::vector<int4>& out);
// TwoBodyCorrelator::measure(N): out (12, nquad, nt, b*b, b*b) c128.
int fm_two_body(FermionMatrix& fm, const void* wall, int kdim, int bdim, int64_t nquad, const int* quads_host,
                int64_t ntsrc, double eps, int64_t maxit, int tol_kind, int prec, void* out, nslx_solve_report* rep,
                hipStream_t s);

// Pieces of the measurement for other drivers (the correlator server):
// S (c128, ng*nx RHS of nt*nx) = point sources at tsrc0 + g*step;
// corr (c128 nt*nx*nx) = 0; corr += re-centred G of those sources; corr *= f.
void corr_sources(void* S, int ng, int tsrc0, int step, int nt, int nx, hipStream_t s);
void corr_zero(void* corr, int nt, int nx, hipStream_t s);
void corr_accumulate(void* corr, const void* G, int ng, int tsrc0, int step, int nt, int nx, hipStream_t s);
void corr_scale(void* corr, int nt, int nx, double f, hipStream_t s);

}  // namespace nslx
