// Two-point correlator of the Hubbard model on the device
// (reference: src/NSL/Measurements/Hubbard/twoPointCorrelationFunction.tpp:116-162).
//
// For time sources tsrc = 0, step, 2 step, ... < Nt (step = Nt / N):
//   S[x][t][y] = delta_{t,tsrc} delta_{y,x}           (Nx point sources, :136)
//   X = CG_{MM^dag}(S)                                  (:139)
//   G = M^dag X                                         (:143)
//   corr[t][y][x] += sign * G[x][(t + tsrc) mod Nt][y]  (shift by -tsrc with
//                    antiperiodic wrap, then transpose, :146-152)
// corr /= N (:160; the reference divides by N even when step does not divide Nt).
// Several time sources are solved as one batch of RHS unless the reference's
// batch-global criterion is asked for, in which case each time source is its
// own solve, as in the reference.
#include "engine.hpp"
#include "corr.hpp"

#include <algorithm>

namespace nslx {
namespace {

// S for the time sources tsrc[0..ng): RHS (i, x) is a point source at (tsrc[i], x).
__global__ void k_sources(double2* S, int ng, int tsrc0, int step, int nt, int nx) {
    const int64_t len = (int64_t)nt * nx;
    const int64_t n = (int64_t)ng * nx * len;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t rhs = i / len, e = i - rhs * len;
        const int g = (int)(rhs / nx), x = (int)(rhs - (int64_t)g * nx);
        const int t = (int)(e / nx), y = (int)(e - (int64_t)t * nx);
        S[i] = make_double2((t == tsrc0 + g * step && y == x) ? 1.0 : 0.0, 0.0);
    }
}

// corr[t][y][x] += scale * sum_g sign * G[(g, x)][(t + tsrc_g) mod nt][y], g in order.
__global__ void k_accumulate(double2* corr, const double2* __restrict__ G, int ng, int tsrc0, int step, int nt, int nx,
                             double scale) {
    const int64_t n = (int64_t)nt * nx * nx;
    const int64_t len = (int64_t)nt * nx;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int t = (int)(i / ((int64_t)nx * nx));
        const int64_t r = i - (int64_t)t * nx * nx;
        const int y = (int)(r / nx), x = (int)(r - (int64_t)y * nx);
        double2 acc = corr[i];
        for (int g = 0; g < ng; ++g) {
            int tt = t + tsrc0 + g * step;
            double sg = 1.0;
            if (tt >= nt) { tt -= nt; sg = -1.0; }
            const double2 v = G[((int64_t)g * nx + x) * len + (int64_t)tt * nx + y];
            acc.x += scale * sg * v.x;
            acc.y += scale * sg * v.y;
        }
        corr[i] = acc;
    }
}

__global__ void k_fill0(double2* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = make_double2(0.0, 0.0);
}

__global__ void k_scale(double2* p, int64_t n, double s) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        p[i].x *= s;
        p[i].y *= s;
    }
}

inline unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 65536); }

}  // namespace

void corr_sources(void* S, int ng, int tsrc0, int step, int nt, int nx, hipStream_t s) {
    const int64_t n = (int64_t)ng * nx * nt * nx;
    hipLaunchKernelGGL(k_sources, dim3(grid_for(n)), dim3(256), 0, s, (double2*)S, ng, tsrc0, step, nt, nx);
    NSLX_LAUNCH_CHECK();
}

void corr_zero(void* corr, int nt, int nx, hipStream_t s) {
    const int64_t n = (int64_t)nt * nx * nx;
    hipLaunchKernelGGL(k_fill0, dim3(grid_for(n)), dim3(256), 0, s, (double2*)corr, n);
    NSLX_LAUNCH_CHECK();
}

void corr_accumulate(void* corr, const void* G, int ng, int tsrc0, int step, int nt, int nx, hipStream_t s) {
    const int64_t n = (int64_t)nt * nx * nx;
    hipLaunchKernelGGL(k_accumulate, dim3(grid_for(n)), dim3(256), 0, s, (double2*)corr, (const double2*)G, ng, tsrc0,
                       step, nt, nx, 1.0);
    NSLX_LAUNCH_CHECK();
}

void corr_scale(void* corr, int nt, int nx, double f, hipStream_t s) {
    const int64_t n = (int64_t)nt * nx * nx;
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, s, (double2*)corr, n, f);
    NSLX_LAUNCH_CHECK();
}

int fm_two_point(FermionMatrix& fm, int64_t ntsrc, double eps, int64_t maxit, int tol_kind, int prec, int64_t max_rhs,
                 void* corr_out, nslx_solve_report* rep, hipStream_t s) {
    fm.require_populated();
    NSLX_REQUIRE(fm.nfield == 1, "two_point: populate exactly one field");
    NSLX_REQUIRE(corr_out, "two_point: null output");
    NSLX_REQUIRE(ntsrc >= 1 && ntsrc <= fm.nt, "two_point: need 1 <= number of time sources <= Nt");
    NSLX_REQUIRE(tol_kind == NSLX_TOL_ABS_BATCH || tol_kind == NSLX_TOL_REL_PER_RHS, "two_point: bad tol_kind");
    NSLX_REQUIRE(prec == NSLX_F64 || tol_kind == NSLX_TOL_REL_PER_RHS,
                 "two_point: low precision needs NSLX_TOL_REL_PER_RHS");
    const int nt = fm.nt, nx = fm.nx;
    const int step = (int)(nt / ntsrc);
    const int nsrc = (nt + step - 1) / step;       // tsrc = 0, step, ... < nt
    int group = 1;
    if (tol_kind == NSLX_TOL_REL_PER_RHS) {
        const int64_t cap = max_rhs > 0 ? max_rhs : 4096;
        group = (int)std::max<int64_t>(1, std::min<int64_t>(nsrc, cap / nx));
    }
    const int64_t len = fm.len();
    const size_t vbytes = (size_t)group * nx * len;   // complex elements per batch vector
    DevBuf<double2> S(vbytes), X(vbytes), G(vbytes);
    double2* corr = static_cast<double2*>(corr_out);
    const int64_t ncorr = (int64_t)nt * nx * nx;
    hipLaunchKernelGGL(k_fill0, dim3(grid_for(ncorr)), dim3(256), 0, s, corr, ncorr);
    NSLX_LAUNCH_CHECK();

    nslx_solve_report agg{};
    int status = NSLX_OK;
    for (int g0 = 0; g0 < nsrc; g0 += group) {
        const int ng = std::min(group, nsrc - g0);
        const int64_t nb = (int64_t)ng * nx;
        const int64_t n = nb * len;
        const int tsrc0 = g0 * step;
        hipLaunchKernelGGL(k_sources, dim3(grid_for(n)), dim3(256), 0, s, S.p, ng, tsrc0, step, nt, nx);
        NSLX_LAUNCH_CHECK();
        nslx_solve_report r{};
        // the reference's CG starts from x0 = b (CG.tpp:97-103); the per-RHS
        // reliable-update solver starts from zero
        // per-RHS solves accept a stop at the FP64 attainable accuracy (counted in floor_stops)
        const int tk = tol_kind == NSLX_TOL_REL_PER_RHS ? NSLX_TOL_REL_PER_RHS_FLOOR : tol_kind;
        const int st = solve(fm, NSLX_OP_MMDAGGER, NSLX_CG, S.p, X.p, nullptr, nb, eps, tk, maxit, prec, 0, &r, s);
        if (st != NSLX_OK) status = st;
        fm.apply(NSLX_OP_MDAGGER, X.p, G.p, nb, NSLX_F64, s);
        hipLaunchKernelGGL(k_accumulate, dim3(grid_for(ncorr)), dim3(256), 0, s, corr, G.p, ng, tsrc0, step, nt, nx,
                           1.0);
        NSLX_LAUNCH_CHECK();
        agg.iterations = std::max(agg.iterations, r.iterations);
        agg.iterations_total += r.iterations_total;
        agg.outer_steps += r.outer_steps;
        agg.converged_rhs += r.converged_rhs;
        agg.rsqr += r.rsqr;
        agg.max_rel_residual = std::max(agg.max_rel_residual, r.max_rel_residual);
        agg.seconds += r.seconds;
        agg.floor_stops += r.floor_stops;
        agg.max_floor_rel = std::max(agg.max_floor_rel, r.max_floor_rel);
        agg.rhs_iters_low += r.rhs_iters_low;
        agg.rhs_iters_f64 += r.rhs_iters_f64;
    }
    hipLaunchKernelGGL(k_scale, dim3(grid_for(ncorr)), dim3(256), 0, s, corr, ncorr, 1.0 / (double)ntsrc);
    NSLX_LAUNCH_CHECK();
    NSLX_HIP(hipStreamSynchronize(s));   // S, X, G are freed on return
    if (rep) *rep = agg;
    return status;
}

// The same measurement for every populated field of fm at once (an ensemble
// of configurations): the point sources of a group of time sources are laid
// out per field (RHS f * ng * nx + g * nx + x uses field f), all fields'
// systems MM^dag X = S go through one time-slice-preconditioned solve (one
// factorisation per field, double-double correction to tol per RHS), then
// G = M^dag X and each field's re-centred G is added to its own corr.
int fm_two_point_fields(FermionMatrix& fm, int64_t ntsrc, double tol, int64_t maxit, int pprec, int64_t max_rhs,
                        void* corr_out, nslx_solve_report* rep, hipStream_t s) {
    fm.require_populated();
    NSLX_REQUIRE(corr_out, "two_point_fields: null output");
    NSLX_REQUIRE(ntsrc >= 1 && ntsrc <= fm.nt, "two_point_fields: need 1 <= number of time sources <= Nt");
    NSLX_REQUIRE(fm.disc == DISC_EXP, "two_point_fields: HubbardExp only (the time-slice solver)");
    const int nt = fm.nt, nx = fm.nx;
    const int64_t nf = fm.nfield;
    const int step = (int)(nt / ntsrc);
    const int nsrc = (nt + step - 1) / step;
    const int64_t cap = max_rhs > 0 ? max_rhs : 8192;
    const int group = (int)std::max<int64_t>(1, std::min<int64_t>(nsrc, cap / (nf * nx)));
    const int64_t len = fm.len();
    const size_t vbytes = (size_t)nf * group * nx * len;
    DevBuf<double2> S(vbytes), X(vbytes), G(vbytes);
    double2* corr = static_cast<double2*>(corr_out);
    const int64_t ncorr = (int64_t)nt * nx * nx;
    hipLaunchKernelGGL(k_fill0, dim3(grid_for(nf * ncorr)), dim3(256), 0, s, corr, nf * ncorr);
    NSLX_LAUNCH_CHECK();
    nslx_solve_report agg{};
    int status = NSLX_OK;
    for (int g0 = 0; g0 < nsrc; g0 += group) {
        const int ng = std::min(group, nsrc - g0);
        const int64_t rpf = (int64_t)ng * nx;          // RHS per field
        const int64_t nb = nf * rpf;
        const int tsrc0 = g0 * step;
        for (int64_t f = 0; f < nf; ++f) {
            const int64_t n = rpf * len;
            hipLaunchKernelGGL(k_sources, dim3(grid_for(n)), dim3(256), 0, s, S.p + f * n, ng, tsrc0, step, nt, nx);
            NSLX_LAUNCH_CHECK();
        }
        nslx_solve_report r{};
        const int st = solve_tss(fm, NSLX_OP_MMDAGGER, S.p, X.p, nullptr, nullptr, nb, tol, maxit, pprec, 0, &r, s);
        if (st != NSLX_OK) status = st;
        fm.apply(NSLX_OP_MDAGGER, X.p, G.p, nb, NSLX_F64, s);
        for (int64_t f = 0; f < nf; ++f) {
            hipLaunchKernelGGL(k_accumulate, dim3(grid_for(ncorr)), dim3(256), 0, s, corr + f * ncorr,
                               G.p + f * rpf * len, ng, tsrc0, step, nt, nx, 1.0);
            NSLX_LAUNCH_CHECK();
        }
        agg.iterations = std::max(agg.iterations, r.iterations);
        agg.iterations_total += r.iterations_total;
        agg.outer_steps += r.outer_steps;
        agg.converged_rhs += r.converged_rhs;
        agg.rsqr += r.rsqr;
        agg.max_rel_residual = std::max(agg.max_rel_residual, r.max_rel_residual);
        agg.seconds += r.seconds;
        agg.rhs_iters_low += r.rhs_iters_low;
        agg.rhs_iters_f64 += r.rhs_iters_f64;
    }
    hipLaunchKernelGGL(k_scale, dim3(grid_for(nf * ncorr)), dim3(256), 0, s, corr, nf * ncorr, 1.0 / (double)ntsrc);
    NSLX_LAUNCH_CHECK();
    NSLX_HIP(hipStreamSynchronize(s));   // S, X, G are freed on return
    if (rep) *rep = agg;
    return status;
}

}  // namespace nslx
