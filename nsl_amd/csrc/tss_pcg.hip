// CG on MM^dag (or M^dag M) preconditioned by the time-slice direct solve,
// inside a double-double defect-correction loop (nslx_solve_tss).
//
// Outer loop (per right-hand side, until |b - A x| <= tol |b| in double-
// double): r = b - A (x_hi + x_lo) in double-double, rounded to FP64; an inner
// flexible PCG on A d = r from d = 0 with FP64 vectors and A applied in FP64
// on the hop kernels, preconditioned by z = A^{-1} r (A = M M^dag or M^dag M) from the
// time-slice solve in FP64 or FP32; x += d in double-double.  The inner solve
// stops a right-hand side once its recursive residual has dropped by rho
// (1e-6) or below 0.3 tol |b|: below that the FP64 recurrence no longer tracks
// the true residual of the correction (eps ||A|| |d|), and the next outer
// step starts from an exact residual instead.  Beta is Polak-Ribiere's,
// <r_new - r_old, z_new> / <r_old, z_old> = -alpha <t, z_new> / <r_old, z_old>,
// which keeps the recurrence convergent when the FP32 preconditioner is not
// exactly linear.  The reference's CGpreconditioned (CGpreconditioned.tpp:14-100)
// is the same outer recurrence with an inner CG as the preconditioner.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "engine.hpp"

namespace nslx {
namespace {

// alpha_b = rz_b / <p,t>_b (0 for a finished RHS)
__global__ void k_alpha(int64_t nb, const double* rz, const double* pt, const int* done, double* alpha) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double ar = 0.0, ai = 0.0;
    if (done[b] < 0) {
        const double pr = pt[2 * b], pi = pt[2 * b + 1], d = pr * pr + pi * pi;
        if (d > 0.0) { ar = rz[b] * pr / d; ai = -rz[b] * pi / d; }
    }
    alpha[2 * b] = ar;
    alpha[2 * b + 1] = ai;
}

// *nactive = the count of c over one block of 256 threads (no memset, no
// atomics: the one block writes it)
__device__ inline void block_count(int c, int* nactive) {
    __shared__ int red[256];
    red[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *nactive = red[0];
}

// done_b = it once |r_b|^2 <= max(rho2 |r0_b|^2, abs2 |b_b|^2); counts the
// rest (one block)
__global__ void __launch_bounds__(256) k_inner_check(int64_t nb, const double* rn2, const double* r0n2,
                                                     const double* bn2, double rho2, double abs2, int it, int* done,
                                                     int* nactive) {
    int c = 0;
    for (int64_t b = threadIdx.x; b < nb; b += 256) {
        if (done[b] >= 0) continue;
        if (rn2[b] <= rho2 * r0n2[b] || rn2[b] <= abs2 * bn2[b] || !(rn2[b] == rn2[b])) done[b] = it;
        else ++c;
    }
    block_count(c, nactive);
}

// beta_b = Re(-conj(alpha_b) <t, z_new>_b) / rz_old_b; rz_old <- Re <r, z_new>
__global__ void k_beta(int64_t nb, const double* alpha, const double* tz, const double* rzn, double* rz,
                       const int* done, double* beta) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double be = 0.0;
    if (done[b] < 0 && rz[b] != 0.0) {
        const double ar = alpha[2 * b], ai = alpha[2 * b + 1];
        const double tr = tz[2 * b], ti = tz[2 * b + 1];
        be = -(ar * tr + ai * ti) / rz[b];   // Re(conj(alpha) tz) = ar tr + ai ti
    }
    beta[2 * b] = be;
    beta[2 * b + 1] = 0.0;
    rz[b] = rzn[2 * b];
}

// outer test on the double-double residual: done_b = round when |r|^2 <=
// tol2 |b|^2; counts the rest (one block)
__global__ void __launch_bounds__(256) k_outer_check(int64_t nb, const double* rn2, const double* bn2, double tol2,
                                                     int round, int* done, int* nactive) {
    int c = 0;
    for (int64_t b = threadIdx.x; b < nb; b += 256) {
        if (done[b] >= 0) continue;
        if (rn2[b] <= tol2 * bn2[b]) done[b] = round;
        else ++c;
    }
    block_count(c, nactive);
}

// start of an outer step: done_in = done_out, r0n2 = rn2
__global__ void k_outer_begin(int64_t nb, const int* done_out, int* done_in, const double* rn2, double* r0n2) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    done_in[b] = done_out[b];
    r0n2[b] = rn2[b];
}

// start of the solve: every RHS open; rn2 = bn2 when x starts at zero
__global__ void k_solve_begin(int64_t nb, const double* bn2, double* rn2, int* done_out) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    done_out[b] = -1;
    if (bn2) rn2[b] = bn2[b];
}

__global__ void k_rz_init(int64_t nb, const double* rzn, double* rz) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) rz[b] = rzn[2 * b];
}

inline unsigned g1(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int solve_tss(FermionMatrix& fm, int op, const void* b, void* x, void* xlo, const void* x0, int64_t nb, double tol,
              int64_t maxit, int pprec, int nchunks, nslx_solve_report* rep, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    NSLX_REQUIRE(b && x, "solve_tss: null buffer");
    NSLX_REQUIRE(b != x && x0 != x && (!xlo || (xlo != x && xlo != b)), "solve_tss: x must not alias b or x0");
    NSLX_REQUIRE(op == NSLX_OP_MMDAGGER || op == NSLX_OP_MDAGGERM,
                 "solve_tss: the time-slice preconditioner serves MMdagger and MdaggerM systems");
    NSLX_REQUIRE(pprec == NSLX_F64 || pprec == NSLX_F32, "solve_tss: preconditioner precision must be F64 or F32");
    NSLX_REQUIRE(tol > 0.0 && maxit >= 1, "solve_tss: bad tol/maxit");
    NSLX_REQUIRE(fm.disc == DISC_EXP, "solve_tss: HubbardExp only");
    fm.require_populated();
    NSLX_REQUIRE(nb >= 1 && nb % fm.nfield == 0, "solve_tss: nbatch must be a positive multiple of nfield");
    nslx_solve_report local{};
    if (!rep) rep = &local;
    *rep = nslx_solve_report{};
    const int64_t len = fm.len();
    const size_t bytes = (size_t)nb * len * 16;
    SolverWork& w = fm.sw;
    w.ensure_scalars(nb, 1, s);
    w.r64.ensure(bytes); w.t64.ensure(bytes); w.p0.ensure(bytes); w.u.ensure(bytes); w.x32.ensure(bytes);
    w.v1.ensure(bytes);
    w.rr_ref.ensure(nb); w.rr_prev.ensure(nb); w.xn2.ensure(nb); w.want.ensure(nb); w.best.ensure(2 * nb);
    w.rstart.ensure(2 * nb);
    DevBuf<double> xl_own;
    double2* xl = static_cast<double2*>(xlo);
    if (!xl) {
        xl_own.alloc(2 * (size_t)nb * len);
        xl = reinterpret_cast<double2*>(xl_own.p);
    }
    void* r = w.r64.p;       // residual of the current outer step's system
    void* z = w.t64.p;       // preconditioned residual
    void* p = w.p0.p;        // search direction
    void* t = w.u.p;         // A p
    void* d = w.x32.p;       // correction of this outer step (FP64)
    double* bn2 = w.bn2.p;
    double* rn2 = w.rn2.p;
    double* r0n2 = w.rr_ref.p;
    double* rz = w.rr_prev.p;
    double* scal = w.scal.p;            // complex dots: pt, tz, rzn (2 nb each)
    double* pt = scal;
    double* tz = scal + 2 * nb;
    double* rzn = scal + 4 * nb;
    double* alpha = w.alpha.p;
    double* beta = w.beta.p;
    int* done_in = w.done.p;
    int* done_out = w.want.p;
    int* nact = w.status.p;
    int* nact_h = w.status_h.p;

    TimeSliceSolver& ts = fm.time_slice_solver(pprec, nchunks, true, s);
    norm2_perb(1, b, nb, len, bn2, fm.vsc, s);
    if (x0) {
        NSLX_HIP(hipMemcpyAsync(x, x0, bytes, hipMemcpyDeviceToDevice, s));
        NSLX_HIP(hipMemsetAsync(xl, 0, bytes, s));
        dd_residual(fm, op, b, x, xl, r, rn2, nb, fm.ddw, s);
    } else {
        NSLX_HIP(hipMemsetAsync(x, 0, bytes, s));
        NSLX_HIP(hipMemsetAsync(xl, 0, bytes, s));
        NSLX_HIP(hipMemcpyAsync(r, b, bytes, hipMemcpyDeviceToDevice, s));
    }
    hipLaunchKernelGGL(k_solve_begin, dim3(g1(nb)), dim3(256), 0, s, nb, x0 ? nullptr : (const double*)bn2, rn2,
                       done_out);
    NSLX_LAUNCH_CHECK();
    auto check_outer = [&](int round) {
        hipLaunchKernelGGL(k_outer_check, dim3(1), dim3(256), 0, s, nb, rn2, bn2, tol * tol, round, done_out, nact);
        NSLX_LAUNCH_CHECK();
    };
    auto count_outer = [&](int round) {
        check_outer(round);
        NSLX_HIP(hipMemcpyAsync(nact_h, nact, sizeof(int), hipMemcpyDeviceToHost, s));
        NSLX_HIP(hipStreamSynchronize(s));
        ts.check_singular(s, true);
        return *nact_h;
    };
    // round 0 marks the RHS already solved on the device; the host takes
    // every RHS as open until the first inner test's count (no wait here)
    check_outer(0);
    int active = (int)nb;
    const double rho = env_double("NSLX_TSS_RHO", 1e-6);
    const int max_outer = (int)env_double("NSLX_TSS_OUTER", 24.0);
    const int max_inner = (int)env_double("NSLX_TSS_INNER", 64.0);
    int64_t inner_total = 0, inner_max = 0, nouter = 0, napply = 0;
    while (active > 0 && nouter < max_outer && inner_total < maxit) {
        ++nouter;
        hipLaunchKernelGGL(k_outer_begin, dim3(g1(nb)), dim3(256), 0, s, nb, done_out, done_in, rn2, r0n2);
        NSLX_LAUNCH_CHECK();
        NSLX_HIP(hipMemsetAsync(d, 0, bytes, s));
        ts.apply(op, r, z, nb, true, s);
        ++napply;
        NSLX_HIP(hipMemcpyAsync(p, z, bytes, hipMemcpyDeviceToDevice, s));
        dot_perb(1, r, z, nb, len, rzn, 0, fm.vsc, s);
        hipLaunchKernelGGL(k_rz_init, dim3(g1(nb)), dim3(256), 0, s, nb, rzn, rz);   // rz = Re <r, z>
        NSLX_LAUNCH_CHECK();
        int it = 0, inner_active = active;
        while (inner_active > 0 && it < max_inner && inner_total < maxit) {
            ++it;
            ++inner_total;
            fm.apply(op, p, t, nb, NSLX_F64, s);
            dot_perb(1, p, t, nb, len, pt, 0, fm.vsc, s);
            hipLaunchKernelGGL(k_alpha, dim3(g1(nb)), dim3(256), 0, s, nb, rz, pt, done_in, alpha);
            NSLX_LAUNCH_CHECK();
            axpy_perb(1, d, p, d, alpha, 1.0, nb, len, 0, done_in, s);     // d += alpha p
            axpy_perb(1, r, t, r, alpha, -1.0, nb, len, 0, done_in, s);    // r -= alpha t
            norm2_perb(1, r, nb, len, rn2, fm.vsc, s);
            hipLaunchKernelGGL(k_inner_check, dim3(1), dim3(256), 0, s, nb, rn2, r0n2, bn2, rho * rho,
                               0.09 * tol * tol, it, done_in, nact);
            NSLX_LAUNCH_CHECK();
            NSLX_HIP(hipMemcpyAsync(nact_h, nact, sizeof(int), hipMemcpyDeviceToHost, s));
            NSLX_HIP(hipStreamSynchronize(s));
            ts.check_singular(s, true);
            inner_active = *nact_h;
            if (inner_active == 0) break;
            ts.apply(op, r, z, nb, true, s);
            ++napply;
            dot_perb(1, t, z, nb, len, tz, 0, fm.vsc, s);
            dot_perb(1, r, z, nb, len, rzn, 0, fm.vsc, s);
            hipLaunchKernelGGL(k_beta, dim3(g1(nb)), dim3(256), 0, s, nb, alpha, tz, rzn, rz, done_in, beta);
            NSLX_LAUNCH_CHECK();
            axpy_perb(1, z, p, p, beta, 1.0, nb, len, 0, done_in, s);      // p = z + beta p
        }
        inner_max = std::max<int64_t>(inner_max, it);
        dd_accumulate(x, xl, d, nb, len, done_out, s);
        dd_residual(fm, op, b, x, xl, r, rn2, nb, fm.ddw, s);
        active = count_outer((int)nouter);
    }
    // report
    // per-RHS residuals, norms and flags into pinned memory, one wait
    w.dbl_h.ensure(2 * (size_t)nb);
    w.int_h.ensure((size_t)nb);
    double* hr = w.dbl_h.p;
    double* hb = w.dbl_h.p + nb;
    int* hd = w.int_h.p;
    NSLX_HIP(hipMemcpyAsync(hr, rn2, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipMemcpyAsync(hb, bn2, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipMemcpyAsync(hd, done_out, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    ts.queue_pass_checks(s);
    NSLX_HIP(hipStreamSynchronize(s));
    ts.check_singular(s, true);
    ts.check_passes_synced(s);
    int64_t conv = 0;
    double maxrel = 0.0, tot = 0.0;
    for (int64_t i = 0; i < nb; ++i) {
        if (hd[i] >= 0) ++conv;
        maxrel = std::max(maxrel, hb[i] > 0 ? std::sqrt(hr[i] / hb[i]) : std::sqrt(hr[i]));
        tot += hr[i];
    }
    rep->iterations = inner_total;
    rep->iterations_total = napply;
    rep->outer_steps = nouter;
    rep->converged_rhs = conv;
    rep->rsqr = tot;
    rep->max_rel_residual = maxrel;
    rep->floor_stops = 0;
    rep->max_floor_rel = 0.0;
    rep->rhs_iters_low = pprec == NSLX_F64 ? 0 : inner_total * nb;
    rep->rhs_iters_f64 = inner_total * nb;
    rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return conv == nb ? NSLX_OK : NSLX_NOT_CONVERGED;
}

}  // namespace nslx
