// Wall-source correlators of the Hubbard model on the device:
//   TwoPointCorrelator::measureK(k, N)
//       (src/NSL/Measurements/Hubbard/twoPointCorrelationFunction.tpp:164-233)
//   TwoBodyCorrelator::measure(N)
//       (src/NSL/Measurements/Hubbard/twoBodyCorrelationFunction.tpp:244-300,
//        contraction channels :481-1000)
//
// A wall source w (nx complex numbers) placed on time slice tsrc is the RHS
// S[t][x] = delta_{t,tsrc} w[x].  For every time source the RHS of all walls
// are one batched solve; the propagators are re-centred on tsrc with the
// antiperiodic sign (Tensor::shift(-tsrc, -2, -1), shift.tpp:63-68) and
// projected on the walls: P[a][b][t] = sum_x conj(w_a[x]) G_b[t][x]
// (inner_product(., ., dim), inner_product.tpp:36-38).
#include "engine.hpp"
#include "corr.hpp"

#include <algorithm>
#include <cmath>

namespace nslx {
namespace {

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

// S[g][w][t][x] = (t == tsrc0 + g step) ? (conj ? conj(wall[w][x]) : wall[w][x]) : 0
__global__ void k_wall_sources(double2* S, const double2* __restrict__ wall, int nw, int ng, int tsrc0, int step,
                               int nt, int nx, int conj) {
    const int64_t n = (int64_t)ng * nw * nt * nx;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % nx);
    const int t = (int)((i / nx) % nt);
    const int w = (int)((i / ((int64_t)nx * nt)) % nw);
    const int g = (int)(i / ((int64_t)nx * nt * nw));
    double2 v = make_double2(0.0, 0.0);
    if (t == tsrc0 + g * step) {
        v = wall[(size_t)w * nx + x];
        if (conj) v.y = -v.y;
    }
    S[i] = v;
}

// corrK[w][t][x] += sum_g sign * G[g][w][(t + tsrc_g) mod nt][x]
__global__ void k_wall_accumulate(double2* corrK, const double2* __restrict__ G, int nw, int ng, int tsrc0, int step,
                                  int nt, int nx) {
    const int64_t n = (int64_t)nw * nt * nx;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % nx);
    const int t = (int)((i / nx) % nt);
    const int w = (int)(i / ((int64_t)nx * nt));
    double2 acc = corrK[i];
    for (int g = 0; g < ng; ++g) {
        const int tsrc = tsrc0 + g * step;
        int ts = t + tsrc;
        double sg = 1.0;
        if (ts >= nt) { ts -= nt; sg = -1.0; }
        const double2 v = G[(((size_t)g * nw + w) * nt + ts) * nx + x];
        acc.x += sg * v.x;
        acc.y += sg * v.y;
    }
    corrK[i] = acc;
}

__global__ void k_scale_c(double2* p, int64_t n, double f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { p[i].x *= f; p[i].y *= f; }
}

// Projection on the walls, written in pool order:
//   out[((ks * kdim + kr) * nt + t) * b * b + ss * b + sr]
//     = f( sum_x g(w[ks*b+ss][x]) * sign * G[kr*b+sr][(t + tshift) mod nt][x] )
// g = conj unless conj_w == 0; f = conj if conj_out.  tshift wraps with the
// antiperiodic sign (tshift = 0: no re-centring).  One wavefront per output.
__global__ void __launch_bounds__(256) k_project(double2* out, const double2* __restrict__ wall,
                                                 const double2* __restrict__ G, int kdim, int bdim, int nt, int nx,
                                                 int tshift, int conj_w, int conj_out) {
    const int nwall = kdim * bdim;
    const int64_t nout = (int64_t)nwall * nwall * nt;
    const int64_t o = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (o >= nout) return;
    // o enumerates (a = sink wall, bsrc = source wall, t)
    const int t = (int)(o % nt);
    const int bsrc = (int)((o / nt) % nwall);
    const int a = (int)(o / ((int64_t)nt * nwall));
    int ts = t + tshift;
    double sg = 1.0;
    if (ts >= nt) { ts -= nt; sg = -1.0; }
    const double2* wa = wall + (size_t)a * nx;
    const double2* gb = G + ((size_t)bsrc * nt + ts) * nx;
    double re = 0.0, im = 0.0;
    for (int x = lane; x < nx; x += 64) {
        double2 w = wa[x];
        if (conj_w) w.y = -w.y;
        const double2 g = gb[x];
        re += w.x * g.x - w.y * g.y;
        im += w.x * g.y + w.y * g.x;
    }
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off, 64);
        im += __shfl_xor(im, off, 64);
    }
    if (lane == 0) {
        const int ks = a / bdim, ss = a % bdim, kr = bsrc / bdim, sr = bsrc % bdim;
        const size_t idx = (((size_t)ks * kdim + kr) * nt + t) * bdim * bdim + (size_t)ss * bdim + sr;
        out[idx] = make_double2(sg * re, sg * (conj_out ? -im : im));
    }
}

// The twelve channels of twoBodyCorrelationFunction.tpp:481-1000.  Pools:
// 0 = C particle, 1 = C hole, 2 = Cdag particle, 3 = Cdag hole (corrPool_,
// corrPoolDag_).  A term is s * X(...) * Y(...) with the direct index pattern
// X(w,z,t,i,l) Y(x,y,t,j,k) or the exchange pattern X(w,y,t,i,k) Y(x,z,t,j,l);
// a channel is alpha * sum of its terms, added as alpha * terms / N.
struct TbTerm { signed char x, y, exch, sign; };
struct TbChannel { double alpha; int nterm; TbTerm t[4]; };
enum { PC = 0, HC = 1, PD = 2, HD = 3 };
__constant__ TbChannel c_tb[12] = {
    // I1S1Iz1Sz1
    {1.0, 2, {{PD, PD, 0, 1}, {PD, PD, 1, -1}, {0, 0, 0, 0}, {0, 0, 0, 0}}},
    // I1S1Iz1Sz0
    {-0.5, 4, {{PD, HC, 0, 1}, {PD, HC, 1, -1}, {HC, PD, 1, -1}, {HC, PD, 0, 1}}},
    // I1S1Iz1Szn1
    {1.0, 2, {{HC, HC, 0, 1}, {HC, HC, 1, -1}, {0, 0, 0, 0}, {0, 0, 0, 0}}},
    // I1S1Iz0Sz1
    {0.5, 4, {{PD, HD, 0, 1}, {PD, HD, 1, -1}, {HD, PD, 1, -1}, {HD, PD, 0, 1}}},
    // I1S1Iz0Szn1
    {0.5, 4, {{PC, HC, 0, 1}, {PC, HC, 1, -1}, {HC, PC, 1, -1}, {HC, PC, 0, 1}}},
    // I1S1Izn1Sz1
    {1.0, 2, {{HD, HD, 0, 1}, {HD, HD, 1, -1}, {0, 0, 0, 0}, {0, 0, 0, 0}}},
    // I1S1Izn1Sz0
    {-0.5, 4, {{PC, HD, 0, 1}, {PC, HD, 1, -1}, {HD, PC, 1, -1}, {HD, PC, 0, 1}}},
    // I1S1Izn1Szn1
    {1.0, 2, {{PC, PC, 0, 1}, {PC, PC, 1, -1}, {0, 0, 0, 0}, {0, 0, 0, 0}}},
    // I0S1Iz0Sz1
    {0.5, 4, {{PD, HD, 0, 1}, {PD, HD, 1, 1}, {HD, PD, 1, 1}, {HD, PD, 0, 1}}},
    // I0S1Iz0Szn1
    {0.5, 4, {{PC, HC, 0, 1}, {PC, HC, 1, 1}, {HC, PC, 1, 1}, {HC, PC, 0, 1}}},
    // I1S0Iz1Sz0
    {-0.5, 4, {{PD, HC, 0, 1}, {PD, HC, 1, 1}, {HC, PD, 1, 1}, {HC, PD, 0, 1}}},
    // I1S0Izn1Sz0
    {-0.5, 4, {{PC, HD, 0, 1}, {PC, HD, 1, 1}, {HD, PC, 1, 1}, {HD, PC, 0, 1}}},
};

struct TbPools { const double2* p[4]; };

// out[ch][q][t][i*b+j][l*b+k] += alpha_ch / N * sum_terms ...
__global__ void k_two_body(double2* out, TbPools pools, const int4* __restrict__ quads, int nquad, int kdim, int bdim,
                           int nt, double inv_n) {
    const int b2 = bdim * bdim;
    const int64_t per_q = (int64_t)nt * b2 * b2;
    const int64_t n = 12 * (int64_t)nquad * per_q;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 >= n) return;
    const int lk = (int)(i0 % b2);
    const int ij = (int)((i0 / b2) % b2);
    const int t = (int)((i0 / b2 / b2) % nt);
    const int64_t cq = i0 / per_q;
    const int q = (int)(cq % nquad);
    const int ch = (int)(cq / nquad);
    const int i = ij / bdim, j = ij % bdim, l = lk / bdim, k = lk % bdim;
    const int4 Q = quads[q];           // (w, x, z, y)
    const int w = Q.x, x = Q.y, z = Q.z, y = Q.w;
    auto at = [&](int pool, int ka, int kb, int sa, int sb) {
        return pools.p[pool][(((size_t)ka * kdim + kb) * nt + t) * b2 + (size_t)sa * bdim + sb];
    };
    const TbChannel& C = c_tb[ch];
    double re = 0.0, im = 0.0;
    for (int m = 0; m < C.nterm; ++m) {
        const TbTerm T = C.t[m];
        double2 a, c;
        if (!T.exch) { a = at(T.x, w, z, i, l); c = at(T.y, x, y, j, k); }
        else { a = at(T.x, w, y, i, k); c = at(T.y, x, z, j, l); }
        re += T.sign * (a.x * c.x - a.y * c.y);
        im += T.sign * (a.x * c.y + a.y * c.x);
    }
    const double f = C.alpha * inv_n;
    double2 v = out[i0];
    v.x += f * re;
    v.y += f * im;
    out[i0] = v;
}

void aggregate(nslx_solve_report& agg, const nslx_solve_report& r) {
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

// per-RHS solves accept a stop at the FP64 attainable accuracy (counted in floor_stops)
int floor_kind(int tol_kind) { return tol_kind == NSLX_TOL_REL_PER_RHS ? NSLX_TOL_REL_PER_RHS_FLOOR : tol_kind; }

void check_common(FermionMatrix& fm, int64_t ntsrc, int tol_kind, int prec) {
    fm.require_populated();
    NSLX_REQUIRE(fm.nfield == 1, "correlator: populate exactly one field");
    NSLX_REQUIRE(ntsrc >= 1 && ntsrc <= fm.nt, "correlator: need 1 <= number of time sources <= Nt");
    NSLX_REQUIRE(tol_kind == NSLX_TOL_ABS_BATCH || tol_kind == NSLX_TOL_REL_PER_RHS, "correlator: bad tol_kind");
    NSLX_REQUIRE(prec == NSLX_F64 || prec == NSLX_F32 || prec == NSLX_BF16X3, "correlator: bad precision");
    NSLX_REQUIRE(prec == NSLX_F64 || tol_kind == NSLX_TOL_REL_PER_RHS,
                 "correlator: low precision needs NSLX_TOL_REL_PER_RHS");
}

}  // namespace

int fm_two_point_wall(FermionMatrix& fm, const void* wall, int64_t nw, int64_t ntsrc, double eps, int64_t maxit,
                      int tol_kind, int prec, void* corrK_out, void* block_out, nslx_solve_report* rep,
                      hipStream_t s) {
    check_common(fm, ntsrc, tol_kind, prec);
    NSLX_REQUIRE(wall && nw >= 1, "two_point_wall: no wall sources");
    NSLX_REQUIRE(corrK_out || block_out, "two_point_wall: no output");
    const int nt = fm.nt, nx = fm.nx;
    const int step = (int)(nt / ntsrc);
    const int nsrc = (nt + step - 1) / step;       // tsrc = 0, step, ... < nt (measureK :185)
    // the reference solves each time source on its own (batch-global CG);
    // the per-RHS mode batches all time sources
    const int group = tol_kind == NSLX_TOL_REL_PER_RHS ? nsrc : 1;
    const int64_t len = fm.len();
    const size_t vel = (size_t)group * nw * len;
    DevBuf<double2> S(vel), X(vel), G(vel), corrK((size_t)nw * len);
    const int64_t nk = nw * len;
    NSLX_HIP(hipMemsetAsync(corrK.p, 0, nk * sizeof(double2), s));
    nslx_solve_report agg{};
    int status = NSLX_OK;
    for (int g0 = 0; g0 < nsrc; g0 += group) {
        const int ng = std::min(group, nsrc - g0);
        const int64_t nb = (int64_t)ng * nw;
        const int tsrc0 = g0 * step;
        hipLaunchKernelGGL(k_wall_sources, dim3(blocks_for(nb * len)), dim3(256), 0, s, S.p,
                           static_cast<const double2*>(wall), (int)nw, ng, tsrc0, step, nt, nx, 0);
        NSLX_LAUNCH_CHECK();
        nslx_solve_report r{};
        const int st = solve(fm, NSLX_OP_MMDAGGER, NSLX_CG, S.p, X.p, nullptr, nb, eps, floor_kind(tol_kind), maxit, prec, 0, &r, s);
        if (st != NSLX_OK) status = st;
        fm.apply(NSLX_OP_MDAGGER, X.p, G.p, nb, NSLX_F64, s);
        hipLaunchKernelGGL(k_wall_accumulate, dim3(blocks_for(nk)), dim3(256), 0, s, corrK.p, G.p, (int)nw, ng, tsrc0,
                           step, nt, nx);
        NSLX_LAUNCH_CHECK();
        aggregate(agg, r);
    }
    // corrK /= N (:205)
    hipLaunchKernelGGL(k_scale_c, dim3(blocks_for(nk)), dim3(256), 0, s, corrK.p, nk, 1.0 / (double)ntsrc);
    NSLX_LAUNCH_CHECK();
    if (corrK_out)
        NSLX_HIP(hipMemcpyAsync(corrK_out, corrK.p, nk * sizeof(double2), hipMemcpyDeviceToDevice, s));
    if (block_out) {
        // corrKblock(t, s1, s2) = sum_x conj(w_s1[x]) corrK(s2, t, x)  (:210-214)
        const int64_t nout = nw * nw * (int64_t)nt;
        hipLaunchKernelGGL(k_project, dim3((unsigned)((nout + 3) / 4)), dim3(256), 0, s,
                           static_cast<double2*>(block_out), static_cast<const double2*>(wall), corrK.p, 1, (int)nw,
                           nt, nx, 0, 1, 0);
        NSLX_LAUNCH_CHECK();
    }
    NSLX_HIP(hipStreamSynchronize(s));
    if (rep) *rep = agg;
    return status;
}

void two_body_quads(const double* momenta, int kdim, std::vector<int4>& out) {
    // eq_modBZ_ (twoBodyCorrelationFunction.tpp:121-135): k - q is a
    // combination m C + n D of the reciprocal vectors C = (4 pi / 3, 0),
    // D = (-2 pi / 3, 2 pi / sqrt 3) with integer m, n (to 2e-12)
    const double Cx = 4.0 * M_PI / 3.0, Cy = 0.0, Dx = -2.0 * M_PI / 3.0, Dy = 2.0 * M_PI / std::sqrt(3.0);
    const double den = Cy * Dx - Cx * Dy;
    auto eq = [&](double kx, double ky, double qx, double qy) {
        const double dx = kx - qx, dy = ky - qy;
        const double m = (dy * Dx - dx * Dy) / den;
        const double n = (dx * Cy - dy * Cx) / den;
        return std::fabs(std::round(m) - m) < 2e-12 && std::fabs(std::round(n) - n) < 2e-12;
    };
    out.clear();
    // loop order and key of measure(N) :139-151: w, x, y, z; key [w][x][z][y]
    for (int w = 0; w < kdim; ++w)
        for (int x = 0; x < kdim; ++x)
            for (int y = 0; y < kdim; ++y)
                for (int z = 0; z < kdim; ++z) {
                    const double* mw = momenta + 2 * w;
                    const double* mx = momenta + 2 * x;
                    const double* my = momenta + 2 * y;
                    const double* mz = momenta + 2 * z;
                    if (eq(mw[0] + mx[0], mw[1] + mx[1], mz[0] + my[0], mz[1] + my[1]))
                        out.push_back(make_int4(w, x, z, y));
                }
}

int fm_two_body(FermionMatrix& fm, const void* wall, int kdim, int bdim, int64_t nquad, const int* quads_host,
                int64_t ntsrc, double eps, int64_t maxit, int tol_kind, int prec, void* out, nslx_solve_report* rep,
                hipStream_t s) {
    check_common(fm, ntsrc, tol_kind, prec);
    NSLX_REQUIRE(wall && kdim >= 1 && bdim >= 1, "two_body: no wall sources");
    NSLX_REQUIRE(nquad >= 0 && (nquad == 0 || quads_host), "two_body: bad momentum quadruples");
    NSLX_REQUIRE(out || nquad == 0, "two_body: null output");
    for (int64_t q = 0; q < 4 * nquad; ++q)
        NSLX_REQUIRE(quads_host[q] >= 0 && quads_host[q] < kdim, "two_body: momentum index out of range");
    const int nt = fm.nt, nx = fm.nx;
    const int step = (int)(nt / ntsrc);
    const int nsrc = (nt + step - 1) / step;
    const int nw = kdim * bdim;
    const int64_t len = fm.len();
    const size_t vel = (size_t)nw * len;
    const size_t npool = (size_t)kdim * kdim * nt * bdim * bdim;
    DevBuf<double2> S(vel), Sc(vel), X(vel), G(vel), pool(4 * npool), phi(len);
    DevBuf<int4> quads((size_t)std::max<int64_t>(nquad, 1));
    const int64_t nout = 12 * nquad * (int64_t)nt * bdim * bdim * bdim * bdim;
    if (nquad > 0) {
        NSLX_HIP(hipMemcpyAsync(quads.p, quads_host, nquad * sizeof(int4), hipMemcpyHostToDevice, s));
        NSLX_HIP(hipMemsetAsync(out, 0, nout * sizeof(double2), s));
    }
    // populate(phi_, species) per species (:262): keep phi and the caller's species
    NSLX_HIP(hipMemcpyAsync(phi.p, fm.phi.p, len * sizeof(double2), hipMemcpyDeviceToDevice, s));
    const int species0 = fm.species;
    nslx_solve_report agg{};
    int status = NSLX_OK;
    const double2* w = static_cast<const double2*>(wall);
    try {
        for (int g = 0; g < nsrc; ++g) {
            const int tsrc = g * step;
            hipLaunchKernelGGL(k_wall_sources, dim3(blocks_for((int64_t)vel)), dim3(256), 0, s, S.p, w, nw, 1, tsrc, step,
                               nt, nx, 0);
            hipLaunchKernelGGL(k_wall_sources, dim3(blocks_for((int64_t)vel)), dim3(256), 0, s, Sc.p, w, nw, 1, tsrc,
                               step, nt, nx, 1);
            NSLX_LAUNCH_CHECK();
            for (int sp = 0; sp < 2; ++sp) {
                const int species = sp == 0 ? NSLX_PARTICLE : NSLX_HOLE;
                fm.populate(phi.p, 1, species, s);
                nslx_solve_report r{};
                // invMMdag = cg_(src); corrK = M^dag invMMdag (:263-266)
                int st = solve(fm, NSLX_OP_MMDAGGER, NSLX_CG, S.p, X.p, nullptr, nw, eps, floor_kind(tol_kind), maxit, prec, 0, &r, s);
                if (st != NSLX_OK) status = st;
                aggregate(agg, r);
                fm.apply(NSLX_OP_MDAGGER, X.p, G.p, nw, NSLX_F64, s);
                // corrKPool(kSink,kSrc,t,sSink,sSrc) = <w(kSink,sSink), shifted corrK(kSrc,sSrc)> (:271-280)
                hipLaunchKernelGGL(k_project, dim3((unsigned)(((int64_t)nw * nw * nt + 3) / 4)), dim3(256), 0, s,
                                   pool.p + (size_t)sp * npool, w, G.p, kdim, bdim, nt, nx, tsrc, 1, 0);
                NSLX_LAUNCH_CHECK();
                // invMdagM = cgDag_(conj(src)); corrKDag = M invMdagM; pool = conj(<conj(w), .>) (:264-281)
                r = nslx_solve_report{};
                st = solve(fm, NSLX_OP_MDAGGERM, NSLX_CG, Sc.p, X.p, nullptr, nw, eps, floor_kind(tol_kind), maxit, prec, 0, &r, s);
                if (st != NSLX_OK) status = st;
                aggregate(agg, r);
                fm.apply(NSLX_OP_M, X.p, G.p, nw, NSLX_F64, s);
                hipLaunchKernelGGL(k_project, dim3((unsigned)(((int64_t)nw * nw * nt + 3) / 4)), dim3(256), 0, s,
                                   pool.p + (size_t)(2 + sp) * npool, w, G.p, kdim, bdim, nt, nx, tsrc, 0, 1);
                NSLX_LAUNCH_CHECK();
            }
            if (nquad > 0) {
                TbPools P{{pool.p, pool.p + npool, pool.p + 2 * npool, pool.p + 3 * npool}};
                hipLaunchKernelGGL(k_two_body, dim3(blocks_for(nout)), dim3(256), 0, s, static_cast<double2*>(out), P,
                                   quads.p, (int)nquad, kdim, bdim, nt, 1.0 / (double)ntsrc);
                NSLX_LAUNCH_CHECK();
            }
        }
    } catch (...) {
        fm.populate(phi.p, 1, species0, s);
        throw;
    }
    fm.populate(phi.p, 1, species0, s);
    NSLX_HIP(hipStreamSynchronize(s));
    if (rep) *rep = agg;
    return status;
}

}  // namespace nslx
