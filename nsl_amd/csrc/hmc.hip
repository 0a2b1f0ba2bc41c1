// Hubbard action, force, LeapfrogRealForce and the HMC step on the device.
//   U~ = (beta/Nt) U                      src/NSL/Action/Implementations/hubbard.tpp:10-19
//   S_G = sum phi*phi / (2 U~), F_G = -phi/U~   hubbardGaugeAction.tpp:59-67
//   S_F = -(logdet M_p + logdet M_h), grad = gradLogDetM_p - gradLogDetM_h,
//   F_F = -grad                            hubbardFermiAction.tpp:68-113
//   leapfrog with Im(F) zeroed             src/NSL/Integrator/Impl/leapfrogRealForce.tpp:35-74
//   H = S + sum p*p / 2, accept iff u <= exp(Re(H0 - H1))   src/NSL/MarkovChain/HMC.tpp:185-241
#include <chrono>
#include <cmath>

#include "engine.hpp"
#include "hmc.hpp"

namespace nslx {
namespace {

constexpr int HNT = 256;

// out = sum a*b (unconjugated, complex), deterministic single-block reduction
__global__ void k_sumprod(const double2* a, const double2* b, int64_t n, double* out) {
    __shared__ double sr[HNT], si[HNT];
    double r = 0.0, im = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += HNT) {
        const double2 x = a[i], y = b[i];
        r += x.x * y.x - x.y * y.y;
        im += x.x * y.y + x.y * y.x;
    }
    sr[threadIdx.x] = r;
    si[threadIdx.x] = im;
    __syncthreads();
    for (int s = HNT / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) { sr[threadIdx.x] += sr[threadIdx.x + s]; si[threadIdx.x] += si[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = sr[0]; out[1] = si[0]; }
}

// F = -(phi/U~ + gp - gh), then Im(F) = 0 if real_force
__global__ void k_force(const double2* phi, const double2* gp, const double2* gh, double invU, int64_t n, int real_force,
                        double2* F) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 f = phi[i], a = gp[i], b = gh[i];
    double2 o = make_double2(-(f.x * invU + a.x - b.x), -(f.y * invU + a.y - b.y));
    if (real_force) o.y = 0.0;
    F[i] = o;
}

// y += c * x (real c)
__global__ void k_axpy_r(double2* y, const double2* x, double c, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2 v = y[i];
    const double2 u = x[i];
    v.x += c * u.x;
    v.y += c * u.y;
    y[i] = v;
}

__global__ void k_real_to_c(const double* in, double2* out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_double2(in[i], 0.0);
}

inline unsigned g1(int64_t n) { return (unsigned)((n + 255) / 256); }

struct HmcWork {
    DevBuf<double2> gp, gh, F, q, p, p0;
    DevBuf<double> red;
};

std::pair<double, double> sumprod(const double2* a, const double2* b, int64_t n, double* red, hipStream_t s) {
    hipLaunchKernelGGL(k_sumprod, dim3(1), dim3(HNT), 0, s, a, b, n, red);
    NSLX_LAUNCH_CHECK();
    double h[2];
    NSLX_HIP(hipMemcpyAsync(h, red, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    return {h[0], h[1]};
}

}  // namespace

void hmc_action(FermionMatrix& fm, const void* phi, double U, int nchunks, double* out2, hipStream_t s) {
    const int64_t n = fm.len();
    const double Ut = fm.delta() * U;
    DevBuf<double> red(2);
    const auto g = sumprod((const double2*)phi, (const double2*)phi, n, red.p, s);
    double lp[2], lh[2];
    fm.populate(phi, 1, NSLX_PARTICLE, s);
    fm_logdet(fm, 0, nchunks, lp, s);
    fm.populate(phi, 1, NSLX_HOLE, s);
    fm_logdet(fm, 0, nchunks, lh, s);
    out2[0] = g.first / (2.0 * Ut) - (lp[0] + lh[0]);
    out2[1] = g.second / (2.0 * Ut) - (lp[1] + lh[1]);
}

static void force_into(FermionMatrix& fm, const void* phi, double U, int real_force, double2* F, double2* gp, double2* gh,
                       hipStream_t s) {
    const int64_t n = fm.len();
    fm.populate(phi, 1, NSLX_PARTICLE, s);
    fm_grad_logdet(fm, 0, gp, s);
    fm.populate(phi, 1, NSLX_HOLE, s);
    fm_grad_logdet(fm, 0, gh, s);
    hipLaunchKernelGGL(k_force, dim3(g1(n)), dim3(256), 0, s, (const double2*)phi, gp, gh, 1.0 / (fm.delta() * U), n,
                       real_force, F);
    NSLX_LAUNCH_CHECK();
}

void hmc_force(FermionMatrix& fm, const void* phi, double U, void* out, hipStream_t s) {
    const int64_t n = fm.len();
    DevBuf<double2> gp(n), gh(n);
    force_into(fm, phi, U, 0, (double2*)out, gp.p, gh.p, s);
    NSLX_HIP(hipStreamSynchronize(s));
}

void hmc_trajectory(FermionMatrix& fm, void* phi, const void* mom, double* S0, double U, double T, int nmd, double uni,
                    int nchunks, nslx_hmc_report* rep, hipStream_t s) {
    NSLX_REQUIRE(nmd >= 1, "hmc: nmd must be >= 1");
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t n = fm.len();
    HmcWork w;
    w.gp.alloc(n); w.gh.alloc(n); w.F.alloc(n); w.q.alloc(n); w.p.alloc(n); w.p0.alloc(n); w.red.alloc(2);
    NSLX_HIP(hipMemcpyAsync(w.q.p, phi, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    NSLX_HIP(hipMemcpyAsync(w.p.p, mom, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    NSLX_HIP(hipMemcpyAsync(w.p0.p, mom, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    const double eps = T / nmd;
    auto kick = [&](double c) {
        force_into(fm, w.q.p, U, 1, w.F.p, w.gp.p, w.gh.p, s);
        hipLaunchKernelGGL(k_axpy_r, dim3(g1(n)), dim3(256), 0, s, w.p.p, (const double2*)w.F.p, c, n);
        NSLX_LAUNCH_CHECK();
    };
    auto drift = [&]() {
        hipLaunchKernelGGL(k_axpy_r, dim3(g1(n)), dim3(256), 0, s, w.q.p, (const double2*)w.p.p, eps, n);
        NSLX_LAUNCH_CHECK();
    };
    kick(0.5 * eps);
    drift();
    for (int k = 0; k < nmd - 1; ++k) { kick(eps); drift(); }
    kick(0.5 * eps);
    double S1[2];
    hmc_action(fm, w.q.p, U, nchunks, S1, s);
    const auto pp0 = sumprod(w.p0.p, w.p0.p, n, w.red.p, s);
    const auto pp1 = sumprod(w.p.p, w.p.p, n, w.red.p, s);
    const double H0r = S0[0] + 0.5 * pp0.first, H0i = S0[1] + 0.5 * pp0.second;
    const double H1r = S1[0] + 0.5 * pp1.first, H1i = S1[1] + 0.5 * pp1.second;
    const double acc = std::exp(H0r - H1r);
    const bool accept = uni <= acc;
    if (accept) {
        NSLX_HIP(hipMemcpyAsync(phi, w.q.p, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
        S0[0] = S1[0];
        S0[1] = S1[1];
    }
    NSLX_HIP(hipStreamSynchronize(s));
    if (rep) {
        rep->dH_re = H1r - H0r;
        rep->dH_im = H1i - H0i;
        rep->acc_prob = acc;
        rep->accepted = accept ? 1 : 0;
        rep->S_re = S0[0];
        rep->S_im = S0[1];
        rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
}

void hmc_chain(FermionMatrix& fm, void* phi, double U, double T, int nmd, int64_t ntraj, uint64_t seed, int nchunks,
               nslx_hmc_report* reps, hipStream_t s) {
    const int64_t n = fm.len();
    double S0[2];
    hmc_action(fm, phi, U, nchunks, S0, s);
    std::vector<double> z(n);
    DevBuf<double> zd(n);
    DevBuf<double2> mom(n);
    for (int64_t k = 0; k < ntraj; ++k) {
        host_normal(seed, 2 * (uint64_t)k, n, z.data());
        double u;
        host_uniform(seed, 2 * (uint64_t)k + 1, 1, &u);
        NSLX_HIP(hipMemcpyAsync(zd.p, z.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_real_to_c, dim3(g1(n)), dim3(256), 0, s, (const double*)zd.p, mom.p, n);
        NSLX_LAUNCH_CHECK();
        hmc_trajectory(fm, phi, mom.p, S0, U, T, nmd, u, nchunks, reps ? reps + k : nullptr, s);
    }
}

}  // namespace nslx

namespace nslx {

void hmc_generate(FermionMatrix& fm, void* phi, int64_t mt0, double U, double T, int nmd, int64_t nconf,
                  int64_t save, int chain, uint64_t seed, int nchunks, void* configs, nslx_markov_state* states,
                  hipStream_t s) {
    NSLX_REQUIRE(phi && nconf >= 1 && mt0 >= 0, "hmc_generate: bad argument");
    NSLX_REQUIRE(chain == NSLX_CHAIN_ALL || chain == NSLX_CHAIN_LAST, "hmc_generate: bad chain kind");
    if (save <= 0) save = 1;   // HMC.tpp:61-63
    const int64_t n = fm.len();
    double S[2];
    hmc_action(fm, phi, U, nchunks, S, s);
    std::vector<double> z(n);
    DevBuf<double> zd(n);
    DevBuf<double2> mom(n);
    int64_t mt = mt0;
    nslx_markov_state cur{S[0], S[1], 1.0, mt, 1, 0.0, 0.0};   // the start state
    // the trajectory that makes markov time m + 1 draws its momenta from
    // Philox counter 2m and its Metropolis uniform from 2m + 1
    auto step = [&]() {
        host_normal(seed, 2 * (uint64_t)mt, n, z.data());
        double u;
        host_uniform(seed, 2 * (uint64_t)mt + 1, 1, &u);
        NSLX_HIP(hipMemcpyAsync(zd.p, z.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_real_to_c, dim3(g1(n)), dim3(256), 0, s, (const double*)zd.p, mom.p, n);
        NSLX_LAUNCH_CHECK();
        nslx_hmc_report r{};
        hmc_trajectory(fm, phi, mom.p, S, U, T, nmd, u, nchunks, &r, s);
        ++mt;
        cur = nslx_markov_state{S[0], S[1], r.acc_prob, mt, r.accepted, r.dH_re, r.dH_im};
    };
    auto record = [&](int64_t i) {
        if (states) states[i] = cur;
        if (configs)
            NSLX_HIP(hipMemcpyAsync(static_cast<double2*>(configs) + i * n, phi, n * sizeof(double2),
                                    hipMemcpyDeviceToDevice, s));
    };
    if (chain == NSLX_CHAIN_ALL) {
        // MC[0] = state; MC[n] = generate_ applied save times to MC[n-1] (HMC.tpp:98-137)
        record(0);
        for (int64_t i = 1; i < nconf; ++i) {
            for (int64_t m = 0; m < save; ++m) step();
            record(i);
        }
    } else {
        // LastState: Nconf * saveFrequency - 1 steps from state (HMC.tpp:144-152)
        for (int64_t k = 1; k < nconf * save; ++k) step();
        record(0);
    }
    NSLX_HIP(hipStreamSynchronize(s));
}

}  // namespace nslx
