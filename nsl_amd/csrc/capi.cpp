// extern "C" boundary of libnslx (declarations and reference citations in
// include/nslx.h).  Every entry point converts exceptions to status codes.
#include <cmath>
#include <cstring>

#include <algorithm>

#include "engine.hpp"
#include "corr.hpp"
#include "pool.hpp"
#include "serve.hpp"
#include "hmc.hpp"
#include "prof.hpp"

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    // A failed runtime call leaves its code as the thread's last error; clear
    // it so that the next call's launch check does not report it again.
    try {
        return f();
    } catch (const nslx::Error& e) {
        g_err = e.what();
        (void)hipGetLastError();
        return e.status;
    } catch (const std::bad_alloc& e) {
        g_err = std::string("host allocation failed: ") + e.what();
        (void)hipGetLastError();
        return NSLX_E_OOM;
    } catch (const std::exception& e) {
        g_err = e.what();
        (void)hipGetLastError();
        return NSLX_E_INTERNAL;
    }
}

inline nslx::FermionMatrix& FM(nslx_fm fm) {
    if (!fm) throw nslx::Error(NSLX_E_ARG, "null fermion-matrix handle");
    return fm->m;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        NSLX_HIP(hipGetDevice(&prev));
        if (prev != dev) NSLX_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};
}  // namespace

extern "C" {

int nslx_version(void) { return 10000; }

const char* nslx_status_string(int s) {
    switch (s) {
        case NSLX_OK: return "ok";
        case NSLX_E_ARG: return "invalid argument";
        case NSLX_E_DEVICE: return "device error";
        case NSLX_E_OOM: return "out of memory";
        case NSLX_E_STATE: return "invalid state";
        case NSLX_NOT_CONVERGED: return "not converged";
        case NSLX_E_INTERNAL: return "internal error";
        default: return "unknown status";
    }
}

const char* nslx_last_error(void) { return g_err.c_str(); }

int nslx_ctx_create(int device, nslx_ctx* out) {
    return guard([&] {
        NSLX_REQUIRE(out, "ctx_create: null out");
        int n = 0;
        NSLX_HIP(hipGetDeviceCount(&n));
        NSLX_REQUIRE(device >= 0 && device < n, "ctx_create: no such device");
        {
            // with the device current: its copy of the no-alloc scratch array
            DeviceGuard dg(device);
            nslx::noalloc_init(device);
        }
        auto* c = new nslx_ctx_s();
        c->c.device = device;
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, device) == hipSuccess) c->c.num_cu = p.multiProcessorCount;
        *out = c;
        return NSLX_OK;
    });
}

int nslx_ctx_destroy(nslx_ctx ctx) {
    delete ctx;
    return NSLX_OK;
}

int nslx_rng_philox4x32(uint64_t seed, uint64_t offset, int64_t nblocks, uint32_t* out4) {
    return guard([&] {
        NSLX_REQUIRE(out4 && nblocks >= 0, "rng: bad arguments");
        nslx::host_philox(seed, offset, nblocks, out4);
        return NSLX_OK;
    });
}
int nslx_rng_uniform(uint64_t seed, uint64_t offset, int64_t n, double* out) {
    return guard([&] {
        NSLX_REQUIRE(out && n >= 0, "rng: bad arguments");
        nslx::host_uniform(seed, offset, n, out);
        return NSLX_OK;
    });
}
int nslx_rng_normal(uint64_t seed, uint64_t offset, int64_t n, double* out) {
    return guard([&] {
        NSLX_REQUIRE(out && n >= 0, "rng: bad arguments");
        nslx::host_normal(seed, offset, n, out);
        return NSLX_OK;
    });
}

int nslx_lattice_honeycomb(int L0, int L1, double kappa, double* K, double* sites) {
    return guard([&] { nslx::lattice_honeycomb(L0, L1, kappa, K, sites); return NSLX_OK; });
}
int nslx_lattice_ring(int n, double kappa, double* K, double* sites) {
    return guard([&] { nslx::lattice_ring(n, kappa, K, sites); return NSLX_OK; });
}
int nslx_lattice_armchair_nanotube(int n, int cells, double kappa, double* K, double* sites) {
    return guard([&] { nslx::lattice_nanotube(n, cells, kappa, K, sites); return NSLX_OK; });
}

namespace {
// beta: real (a complex beta would make exp(s delta K) complex, which the
// engine's real-E kernels do not hold); mu: complex
void check_beta_mu(double beta_re, double beta_im, double mu_re, double mu_im) {
    NSLX_REQUIRE(std::isfinite(beta_re) && std::isfinite(beta_im) && std::isfinite(mu_re) && std::isfinite(mu_im),
                 "fm_create: beta and mu must be finite");
    NSLX_REQUIRE(beta_im == 0.0, "fm_create: beta must be real (Im beta != 0 is not supported)");
}
}  // namespace

int nslx_fm_create(nslx_ctx ctx, int nx, int nt, const double* K, double beta_re, double beta_im, double mu_re,
                   double mu_im, int species, nslx_fm* out) {
    return guard([&] {
        NSLX_REQUIRE(ctx && out && K, "fm_create: null argument");
        NSLX_REQUIRE(nx >= 1 && nt >= 2, "fm_create: need nx >= 1 and nt >= 2");
        NSLX_REQUIRE(species == NSLX_PARTICLE || species == NSLX_HOLE, "fm_create: bad species");
        check_beta_mu(beta_re, beta_im, mu_re, mu_im);
        DeviceGuard dg(ctx->c.device);
        auto f = std::make_unique<nslx_fm_s>();
        nslx::fm_init(f->m, &ctx->c, nx, nt, K, beta_re, mu_re, species);
        f->m.mu_im = mu_im;
        *out = f.release();
        return NSLX_OK;
    });
}

int nslx_fm_create_diag(nslx_ctx ctx, int nx, int nt, const double* K, double beta_re, double beta_im,
                        double mu_re, double mu_im, int species, nslx_fm* out) {
    return guard([&] {
        NSLX_REQUIRE(ctx && out && K, "fm_create_diag: null argument");
        check_beta_mu(beta_re, beta_im, mu_re, mu_im);
        DeviceGuard dg(ctx->c.device);
        auto f = std::make_unique<nslx_fm_s>();
        nslx::fm_init(f->m, &ctx->c, nx, nt, K, beta_re, mu_re, species, nslx::DISC_DIAG);
        f->m.mu_im = mu_im;
        *out = f.release();
        return NSLX_OK;
    });
}

int nslx_fm_destroy(nslx_fm fm) {
    if (fm) {
        (void)hipDeviceSynchronize();
        delete fm;
    }
    return NSLX_OK;
}

int nslx_fm_populate(nslx_fm fm, const void* phi, int64_t nfield, int species, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        m.populate(phi, nfield, species, nslx::as_stream(stream));
        return NSLX_OK;
    });
}

int nslx_fm_apply(nslx_fm fm, int op, const void* in, void* out, int64_t nbatch, int precision, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        m.apply(op, in, out, nbatch, precision, nslx::as_stream(stream));
        return NSLX_OK;
    });
}

int nslx_fm_apply_noalloc(nslx_fm fm, int op, const void* in, void* out, int64_t nbatch, int precision,
                          void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        nslx::apply_noalloc(m, op, in, out, nbatch, precision, nslx::as_stream(stream));
        return NSLX_OK;
    });
}

int nslx_fm_logdet(nslx_fm fm, int64_t field, int nchunks, double* out_host2, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        nslx::fm_logdet(m, field, nchunks, out_host2, nslx::as_stream(stream));
        return NSLX_OK;
    });
}

int nslx_fm_grad_logdet(nslx_fm fm, int64_t field, void* out, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        nslx::fm_grad_logdet(m, field, out, nslx::as_stream(stream));
        return NSLX_OK;
    });
}

int nslx_solve(nslx_fm fm, int op, int method, const void* b, void* x, const void* x0, int64_t nbatch, double tol,
               int tol_kind, int64_t maxit, int precision, int check_true_residual, nslx_solve_report* report,
               void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        return nslx::solve(m, op, method, b, x, x0, nbatch, tol, tol_kind, maxit, precision, check_true_residual,
                           report, nslx::as_stream(stream));
    });
}

int nslx_solve_dd(nslx_fm fm, int op, const void* b, void* x, void* x_lo, const void* x0, int64_t nbatch, double tol,
                  int64_t maxit, int precision, nslx_solve_report* report, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        return nslx::solve_dd(m, op, b, x, x_lo, x0, nbatch, tol, maxit, precision, report, nslx::as_stream(stream));
    });
}

int nslx_solve_precond(nslx_fm fm, nslx_fm fm_prec, int op, const void* b, void* x, const void* x0, int64_t nbatch,
                       double tol, int tol_kind, int64_t maxit, int inner_precision, nslx_solve_report* report,
                       void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        auto& mp = FM(fm_prec);
        DeviceGuard dg(m.ctx->device);
        return nslx::solve_precond(m, mp, op, b, x, x0, nbatch, tol, tol_kind, maxit, inner_precision, report,
                                   nslx::as_stream(stream));
    });
}

int nslx_fm_solve_direct(nslx_fm fm, int op, const void* in, void* out, int64_t nbatch, int precision, int nchunks,
                         void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        NSLX_REQUIRE(in && out && in != out, "solve_direct: null or aliased buffers");
        NSLX_REQUIRE(precision == NSLX_F64 || precision == NSLX_F32, "solve_direct: precision must be F64 or F32");
        m.require_populated();
        auto s = nslx::as_stream(stream);
        auto& ts = m.time_slice_solver(precision, nchunks, false, s);
        ts.apply(op, in, out, nbatch, false, s);
        ts.check_singular(s, false);
        ts.check_passes(s);
        return NSLX_OK;
    });
}

int nslx_solve_tss(nslx_fm fm, int op, const void* b, void* x, void* x_lo, const void* x0, int64_t nbatch, double tol,
                   int64_t maxit, int precond_precision, int nchunks, nslx_solve_report* report, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        return nslx::solve_tss(m, op, b, x, x_lo, x0, nbatch, tol, maxit, precond_precision, nchunks, report,
                               nslx::as_stream(stream));
    });
}

int nslx_dd_residual(nslx_fm fm, int op, const void* b, const void* x, const void* x_lo, void* r, double* rel_host,
                     int64_t nbatch, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        NSLX_REQUIRE(b && x, "dd_residual: null buffer");
        auto s = nslx::as_stream(stream);
        const size_t bytes = (size_t)nbatch * m.len() * 16;
        nslx::DevBuf<unsigned char> rr;
        void* rp = r;
        if (!rp) {
            rr.alloc(bytes);
            rp = rr.p;
        }
        nslx::DevBuf<double> rn2(nbatch > 0 ? nbatch : 1), bn2(nbatch > 0 ? nbatch : 1);
        nslx::dd_residual(m, op, b, x, x_lo, rp, rn2.p, nbatch, m.ddw, s);
        nslx::norm2_perb(1, b, nbatch, m.len(), bn2.p, m.vsc, s);
        std::vector<double> hr(nbatch), hb(nbatch);
        NSLX_HIP(hipMemcpyAsync(hr.data(), rn2.p, nbatch * sizeof(double), hipMemcpyDeviceToHost, s));
        NSLX_HIP(hipMemcpyAsync(hb.data(), bn2.p, nbatch * sizeof(double), hipMemcpyDeviceToHost, s));
        NSLX_HIP(hipStreamSynchronize(s));
        if (rel_host)
            for (int64_t k = 0; k < nbatch; ++k) rel_host[k] = hb[k] > 0 ? std::sqrt(hr[k] / hb[k]) : std::sqrt(hr[k]);
        return NSLX_OK;
    });
}

int nslx_fm_export(nslx_fm fm, double* E_host, void* P_host, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        m.require_populated();
        auto s = nslx::as_stream(stream);
        if (E_host) {
            const double a = m.disc == nslx::DISC_DIAG ? 0.0 : m.sgn() * m.delta();
            NSLX_REQUIRE(m.disc == nslx::DISC_EXP, "fm_export: HubbardExp only");
            const double* e = m.E(a, s).dense.p;
            NSLX_HIP(hipMemcpyAsync(E_host, e, (size_t)m.nx * m.nx * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        if (P_host)
            NSLX_HIP(hipMemcpyAsync(P_host, m.P64.p, (size_t)m.nfield * m.len() * sizeof(double2), hipMemcpyDeviceToHost, s));
        NSLX_HIP(hipStreamSynchronize(s));
        return NSLX_OK;
    });
}

int nslx_batch_solve(nslx_ctx ctx, int64_t nitems, const nslx_batch_item* items, double tol, int precision,
                     int64_t maxit, int use_graph, nslx_batch_report* report, void* stream) {
    return guard([&] {
        NSLX_REQUIRE(ctx, "null context handle");
        NSLX_REQUIRE(precision == NSLX_F64 || precision == NSLX_F32 || precision == NSLX_BF16X3,
                     "batch_solve: bad precision");
        DeviceGuard dg(ctx->c.device);
        return nslx::batch_solve(ctx->c, nitems, items, tol, precision, maxit, use_graph, report,
                                 nslx::as_stream(stream));
    });
}

int nslx_batch_solve_tss(nslx_ctx ctx, int64_t nitems, const nslx_batch_item* items, void* const* x_lo,
                         double tol, int precond_precision, int64_t maxit, int nstreams,
                         nslx_batch_report* report) {
    return guard([&] {
        NSLX_REQUIRE(ctx, "null context handle");
        DeviceGuard dg(ctx->c.device);
        return nslx::batch_solve_tss(ctx->c, nitems, items, x_lo, tol, precond_precision, maxit, nstreams, report);
    });
}

int nslx_server_create(nslx_ctx ctx, const nslx_server_config* cfg, nslx_server* out) {
    return guard([&] {
        NSLX_REQUIRE(ctx && out, "server_create: null argument");
        nslx_server_config c{};
        if (cfg) c = *cfg;
        DeviceGuard dg(ctx->c.device);
        auto s = std::make_unique<nslx_server_s>();
        s->s = std::make_unique<nslx::Server>(&ctx->c, c);
        *out = s.release();
        return NSLX_OK;
    });
}

int nslx_server_destroy(nslx_server srv) {
    return guard([&] {
        delete srv;
        return NSLX_OK;
    });
}

int nslx_server_submit(nslx_server srv, const nslx_request* req, int64_t* id) {
    return guard([&] {
        NSLX_REQUIRE(srv && req && id, "server_submit: null argument");
        *id = srv->s->submit(*req);
        return NSLX_OK;
    });
}

int nslx_server_wait(nslx_server srv, int64_t id, double timeout_ms, void* corr, nslx_result* result) {
    return guard([&] {
        NSLX_REQUIRE(srv, "server_wait: null server");
        return srv->s->wait(id, timeout_ms, corr, result);
    });
}

int nslx_server_stats_get(nslx_server srv, nslx_server_stats* stats) {
    return guard([&] {
        NSLX_REQUIRE(srv && stats, "server_stats: null argument");
        srv->s->stats(stats);
        return NSLX_OK;
    });
}

int nslx_correlator_two_point(nslx_fm fm, int64_t n_tsrc, double eps, int64_t maxit, int tol_kind, int precision,
                              int64_t max_rhs, void* corr, nslx_solve_report* report, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        return nslx::fm_two_point(m, n_tsrc, eps, maxit, tol_kind, precision, max_rhs, corr, report,
                                  nslx::as_stream(stream));
    });
}

int nslx_correlator_two_point_fields(nslx_fm fm, int64_t n_tsrc, double tol, int64_t maxit, int precond_precision,
                                     int64_t max_rhs, void* corr, nslx_solve_report* report, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        return nslx::fm_two_point_fields(m, n_tsrc, tol, maxit, precond_precision, max_rhs, corr, report,
                                         nslx::as_stream(stream));
    });
}

int nslx_correlator_two_point_wall(nslx_fm fm, const void* wall, int64_t nw, int64_t n_tsrc, double eps,
                                   int64_t maxit, int tol_kind, int precision, void* corrK, void* block,
                                   nslx_solve_report* report, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        return nslx::fm_two_point_wall(m, wall, nw, n_tsrc, eps, maxit, tol_kind, precision, corrK, block, report,
                                       nslx::as_stream(stream));
    });
}

int nslx_two_body_quads(const double* momenta, int kdim, int64_t* nquad, int* quads) {
    return guard([&] {
        NSLX_REQUIRE(momenta && nquad && kdim >= 1, "two_body_quads: bad argument");
        std::vector<int4> q;
        nslx::two_body_quads(momenta, kdim, q);
        if (quads) {
            NSLX_REQUIRE(*nquad >= (int64_t)q.size(), "two_body_quads: quads holds fewer than the count");
            for (size_t i = 0; i < q.size(); ++i) {
                quads[4 * i] = q[i].x;
                quads[4 * i + 1] = q[i].y;
                quads[4 * i + 2] = q[i].z;
                quads[4 * i + 3] = q[i].w;
            }
        }
        *nquad = (int64_t)q.size();
        return NSLX_OK;
    });
}

int nslx_correlator_two_body(nslx_fm fm, const void* wall, int kdim, int bdim, int64_t nquad, const int* quads,
                             int64_t n_tsrc, double eps, int64_t maxit, int tol_kind, int precision, void* out,
                             nslx_solve_report* report, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        return nslx::fm_two_body(m, wall, kdim, bdim, nquad, quads, n_tsrc, eps, maxit, tol_kind, precision, out,
                                 report, nslx::as_stream(stream));
    });
}

int nslx_hmc_action(nslx_fm fm, const void* phi, double U, int nchunks, double* out_host2, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        NSLX_REQUIRE(phi && out_host2, "hmc_action: null argument");
        DeviceGuard dg(m.ctx->device);
        nslx::hmc_action(m, phi, U, nchunks, out_host2, nslx::as_stream(stream));
        return NSLX_OK;
    });
}

int nslx_hmc_force(nslx_fm fm, const void* phi, double U, void* force_out, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        NSLX_REQUIRE(phi && force_out, "hmc_force: null argument");
        DeviceGuard dg(m.ctx->device);
        nslx::hmc_force(m, phi, U, force_out, nslx::as_stream(stream));
        return NSLX_OK;
    });
}

int nslx_hmc_trajectory(nslx_fm fm, void* phi, const void* momentum, double* S0_host2, double U, double traj_length,
                        int nmd, double uniform, int nchunks, nslx_hmc_report* report, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        NSLX_REQUIRE(phi && momentum && S0_host2, "hmc_trajectory: null argument");
        DeviceGuard dg(m.ctx->device);
        nslx::hmc_trajectory(m, phi, momentum, S0_host2, U, traj_length, nmd, uniform, nchunks, report,
                             nslx::as_stream(stream));
        return NSLX_OK;
    });
}

int nslx_hmc_chain(nslx_fm fm, void* phi, double U, double traj_length, int nmd, int64_t ntraj, uint64_t seed,
                   int nchunks, nslx_hmc_report* reports, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        NSLX_REQUIRE(phi && ntraj >= 0, "hmc_chain: bad argument");
        DeviceGuard dg(m.ctx->device);
        nslx::hmc_chain(m, phi, U, traj_length, nmd, ntraj, seed, nchunks, reports, nslx::as_stream(stream));
        return NSLX_OK;
    });
}

int nslx_hmc_generate(nslx_fm fm, void* phi, int64_t markov_time0, double U, double traj_length, int nmd,
                      int64_t nconf, int64_t save_frequency, int chain, uint64_t seed, int nchunks, void* configs,
                      nslx_markov_state* states, void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        nslx::hmc_generate(m, phi, markov_time0, U, traj_length, nmd, nconf, save_frequency, chain, seed, nchunks,
                           configs, states, nslx::as_stream(stream));
        return NSLX_OK;
    });
}

int nslx_tile_count(int precision, int* count) {
    return guard([&] {
        NSLX_REQUIRE(count, "tile_count: null output");
        NSLX_REQUIRE(precision == NSLX_F64 || precision == NSLX_F32 || precision == NSLX_BF16X3,
                     "tile_count: bad precision");
        *count = (int)nslx::tile_menu(precision).size();
        return NSLX_OK;
    });
}

int nslx_ctx_set_tile(nslx_ctx ctx, int precision, int nx, int nt, int64_t nbatch, int tile) {
    return guard([&] {
        NSLX_REQUIRE(ctx, "null context handle");
        NSLX_REQUIRE(precision == NSLX_F64 || precision == NSLX_F32 || precision == NSLX_BF16X3,
                     "set_tile: bad precision");
        NSLX_REQUIRE(nx > 0 && nt > 0 && nbatch > 0, "set_tile: bad shape");
        if (tile >= 0) {
            NSLX_REQUIRE(tile < (int)nslx::tile_menu(precision).size(), "set_tile: bad tile id");
            NSLX_REQUIRE(nslx::tile_fits(precision, tile, nx), "set_tile: tile does not fit this lattice");
        }
        ctx->c.tiles.set(precision, nx, nt, nbatch, tile);
        return NSLX_OK;
    });
}

int nslx_ctx_get_tile(nslx_ctx ctx, int precision, int nx, int nt, int64_t nbatch, int* tile) {
    return guard([&] {
        NSLX_REQUIRE(ctx && tile, "get_tile: null argument");
        NSLX_REQUIRE(nx > 0 && nt > 0 && nbatch > 0, "get_tile: bad shape");
        *tile = ctx->c.tiles.get(precision, nx, nt, nbatch);
        return NSLX_OK;
    });
}

int nslx_fm_tune(nslx_fm fm, int precision, int64_t nbatch, int reps, int mode, int* best_tile, double* best_us,
                 void* stream) {
    return guard([&] {
        auto& m = FM(fm);
        DeviceGuard dg(m.ctx->device);
        return nslx::fm_tune(m, precision, nbatch, reps, mode, best_tile, best_us, nslx::as_stream(stream));
    });
}

int nslx_prof_enable(int on) {
    nslx::prof_enable(on);
    return NSLX_OK;
}

int nslx_prof_reset(void) {
    return guard([&] { nslx::prof_reset(); return NSLX_OK; });
}

int nslx_prof_read(char* buf, int64_t size) {
    int n = 0;
    const int st = guard([&] {
        NSLX_REQUIRE(buf && size > 0, "prof_read: bad buffer");
        const std::string r = nslx::prof_report(&n);
        const size_t m = std::min<size_t>(r.size(), (size_t)size - 1);
        std::memcpy(buf, r.data(), m);
        buf[m] = 0;
        return NSLX_OK;
    });
    return st == NSLX_OK ? n : -st;
}

}  // extern "C"
