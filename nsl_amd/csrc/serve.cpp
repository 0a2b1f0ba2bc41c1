// Correlator server: a request queue in front of the batched CG pool, with
// admission by device-memory footprint and continuous batching at Krylov-
// iteration boundaries (see nslx_server_* in nslx.h).
//
// One worker thread owns the device work: it admits queued requests in FIFO
// order while their footprint fits the budget (building the fermion matrix,
// the point sources and a pool job), runs one pool step (a replay of
// iters_per_replay iterations over every admitted request), and turns the
// jobs that finished into correlators (G = M^dag X, re-centred and averaged
// over the time sources).  Submitting and waiting never touch the device
// except for the final copy of a result.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "corr.hpp"
#include "engine.hpp"
#include "pool.hpp"
#include "serve.hpp"

namespace nslx {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    nslx_request r{};
    std::vector<double> K;
    std::vector<double2> phi;
    Clock::time_point t_submit, t_admit, t_done;
    std::unique_ptr<FermionMatrix> fm;
    DevBuf<double2> S, X, G, corr;
    int64_t nb = 0;
    int64_t rhs0 = -1;          // pool RHS slots reserved at admission
    int step = 1, nsrc = 1;
    double bytes = 0.0;
    nslx_result res{};
    bool done = false;
    std::string error;
};

namespace {
double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// device bytes a request can hold at once (FP64 recurrence, escalation included)
double footprint(const nslx_request& r) {
    const double nx = r.nx, nt = r.nt;
    const int step = std::max<int64_t>(1, r.nt / std::max<int64_t>(1, r.n_tsrc));
    const double nsrc = (r.nt + step - 1) / step;
    const double nb = nsrc * nx;
    const double vec = nb * nt * nx * 16.0;
    const double ld = (double)e_ld(r.nx), rows = (double)e_rows(r.nx);
    const double emem = nx * nx * 16.0 + rows * ld * 16.0;     // dense + padded FP64/FP32/BF16 copies
    const double phases = nt * nx * 40.0;
    const double corr = nt * nx * nx * 16.0;
    return 11.0 * vec + emem + phases + corr + (64 << 20);   // S, X, G, apply work, 4 recurrence, r64, t64, slack
}

// device bytes of a request served by the time-slice-preconditioned solver:
// the factorisation (chunk products, Gauss-Jordan blocks, conjugate
// transposes), the apply buffers, the solver's vectors, the request's own
// S, X, G and correlator; g = the growth per slice that fixes the chunking
double footprint_tss(const nslx_request& r, double g) {
    const double n = r.nx, nt = r.nt, nn = n * n;
    const int step = std::max<int64_t>(1, r.nt / std::max<int64_t>(1, r.n_tsrc));
    const double nb = (double)((r.nt + step - 1) / step) * n;
    const bool f64 = r.precision == NSLX_F64;
    const double esz = f64 ? 16.0 : 8.0;
    const int c = tss_chunk(r.nt, g, f64 ? TSS_BOUND_PRECOND_F64 : TSS_BOUND_DIRECT,
                            f64 ? TSS_CAP_PRECOND : TSS_CAP_DIRECT);
    const double nc = (double)(r.nt / c);
    const double factor = esz * (2 * nc * nn + (nc - 1) * 10 * nn + 2 * nn + (nc - 1) * 4 * nn + (nc - 1) * nn +
                                 std::max(0.0, nc - 2) * nn + 2 * nn + 16 * 8 * n);
    const double vec = nb * nt * n;
    const double apply = esz * (3 * vec + 3 * nb * nc * n + 5 * nb * n);
    const double solver = 16.0 * 8 * vec;                 // r, z, p, t, d, v, x_lo, spare
    const double request = 16.0 * 3 * vec + 16.0 * nt * nn;
    const double ld = (double)e_ld(r.nx), rows = (double)e_rows(r.nx);
    const double emem = 3 * nn * 16.0 + rows * ld * 28.0;  // dense, transposed, padded copies
    return factor + apply + solver + request + emem + nt * n * 40.0 + (64 << 20);
}
}  // namespace

Server::Server(Ctx* ctx, const nslx_server_config& cfg) : ctx_(ctx), cfg_(cfg) {
    if (cfg_.max_rhs <= 0) cfg_.max_rhs = 65536;
    if (cfg_.iters_per_replay <= 0) cfg_.iters_per_replay = 32;
    if (!(cfg_.hbm_budget > 0)) cfg_.hbm_budget = 64.0 * (1ull << 30);
    NSLX_HIP(hipSetDevice(ctx_->device));
    NSLX_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    if (cfg_.precond_workers > 0) {
        for (int w = 0; w < cfg_.precond_workers; ++w) workers_.emplace_back([this] { run_tss(); });
        return;
    }
    pool_ = std::make_unique<Pool>(ctx_, s_, cfg_.max_rhs, cfg_.use_graph, cfg_.iters_per_replay);
    th_ = std::thread([this] { run(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (th_.joinable()) th_.join();
    for (auto& t : workers_)
        if (t.joinable()) t.join();
    pool_.reset();
    (void)hipStreamDestroy(s_);
}

int64_t Server::submit(const nslx_request& r) {
    NSLX_REQUIRE(r.K && r.phi, "server_submit: null K or phi");
    NSLX_REQUIRE(r.nx >= 1 && r.nt >= 2, "server_submit: bad shape");
    NSLX_REQUIRE(r.n_tsrc >= 1 && r.n_tsrc <= r.nt, "server_submit: need 1 <= n_tsrc <= nt");
    NSLX_REQUIRE(r.tol > 0.0, "server_submit: tol must be > 0");
    NSLX_REQUIRE(r.precision == NSLX_F64 || r.precision == NSLX_F32 || r.precision == NSLX_BF16X3,
                 "server_submit: bad precision");
    NSLX_REQUIRE(r.species == NSLX_PARTICLE || r.species == NSLX_HOLE, "server_submit: bad species");
    NSLX_REQUIRE(r.discretisation == NSLX_DISC_EXP || r.discretisation == NSLX_DISC_DIAG,
                 "server_submit: bad discretisation");
    auto q = std::make_shared<Request>();
    q->r = r;
    q->K.assign(r.K, r.K + (size_t)r.nx * r.nx);
    const double2* ph = static_cast<const double2*>(r.phi);
    q->phi.assign(ph, ph + (size_t)r.nt * r.nx);
    q->r.K = nullptr;
    q->r.phi = nullptr;
    if (q->r.maxit <= 0) q->r.maxit = 100000;
    q->step = (int)std::max<int64_t>(1, r.nt / r.n_tsrc);
    q->nsrc = (r.nt + q->step - 1) / q->step;
    q->nb = (int64_t)q->nsrc * r.nx;
    if (cfg_.precond_workers > 0) {
        NSLX_REQUIRE(r.precision == NSLX_F64 || r.precision == NSLX_F32,
                     "server_submit: the preconditioned server takes precision F64 or F32 (the preconditioner's)");
        NSLX_REQUIRE(r.discretisation == NSLX_DISC_EXP, "server_submit: the preconditioned server serves HubbardExp");
        double normK = 0.0, imphi = 0.0;
        for (int i = 0; i < r.nx; ++i) {
            double rs = 0.0;
            for (int j = 0; j < r.nx; ++j) rs += std::fabs(q->K[(size_t)i * r.nx + j]);
            normK = std::max(normK, rs);
        }
        for (const double2& z : q->phi) imphi = std::max(imphi, std::fabs(z.y));
        const double delta = r.beta / r.nt;
        q->bytes = footprint_tss(r, std::fabs(delta) * normK + std::fabs(delta * r.mu) + imphi);
    } else {
        q->bytes = footprint(r);
        NSLX_REQUIRE(q->nb <= cfg_.max_rhs, "server_submit: request needs more RHS slots than the server has");
    }
    NSLX_REQUIRE(q->bytes <= cfg_.hbm_budget, "server_submit: request exceeds the server's memory budget");
    q->t_submit = Clock::now();
    int64_t id;
    {
        std::lock_guard<std::mutex> g(mu_);
        NSLX_REQUIRE(!stop_, "server_submit: server is stopping");
        NSLX_REQUIRE(fatal_.empty(), ("server_submit: server failed: " + fatal_).c_str());
        id = ++next_id_;
        q->id = id;
        queue_.push_back(q);
        all_[id] = q;
        ++submitted_;
    }
    cv_.notify_all();
    return id;
}

// Device set-up of an admitted request: fermion matrix, sources, pool job.
void Server::admit(Request& q) {
    const nslx_request& r = q.r;
    q.fm = std::make_unique<FermionMatrix>();
    fm_init(*q.fm, ctx_, r.nx, r.nt, q.K.data(), r.beta, r.mu, r.species,
            r.discretisation == NSLX_DISC_DIAG ? DISC_DIAG : DISC_EXP);
    DevBuf<double2> phi((size_t)r.nt * r.nx);
    NSLX_HIP(hipMemcpyAsync(phi.p, q.phi.data(), phi.n * sizeof(double2), hipMemcpyHostToDevice, s_));
    q.fm->populate(phi.p, 1, r.species, s_);   // deep copy; synchronises
    const size_t n = (size_t)q.nb * r.nt * r.nx;
    q.S.alloc(n);
    q.X.alloc(n);
    corr_sources(q.S.p, q.nsrc, 0, q.step, r.nt, r.nx, s_);
    auto j = std::make_unique<PoolJob>();
    j->id = q.id;
    j->fm = q.fm.get();
    j->b = q.S.p;
    j->x = q.X.p;
    j->nb = q.nb;
    j->tol = r.tol;
    j->prec = r.precision;
    j->maxit = r.maxit;
    j->rhs0 = q.rhs0;
    q.rhs0 = -1;   // the pool owns the slots from here (and frees them if admit fails)
    pool_->admit(std::move(j));
}

void Server::correlate(Request& q, hipStream_t s) {
    const nslx_request& r = q.r;
    const size_t n = (size_t)q.nb * r.nt * r.nx;
    q.G.alloc(n);
    q.fm->apply(NSLX_OP_MDAGGER, q.X.p, q.G.p, q.nb, NSLX_F64, s);
    q.corr.alloc((size_t)r.nt * r.nx * r.nx);
    corr_zero(q.corr.p, r.nt, r.nx, s);
    corr_accumulate(q.corr.p, q.G.p, q.nsrc, 0, q.step, r.nt, r.nx, s);
    corr_scale(q.corr.p, r.nt, r.nx, 1.0 / (double)r.n_tsrc, s);
    NSLX_HIP(hipStreamSynchronize(s));
    q.S.reset();
    q.X.reset();
    q.G.reset();
    q.fm.reset();
}

void Server::complete(Request& q, PoolJob& j) {
    correlate(q, s_);
    q.res.status = j.status;
    q.res.iterations = j.iters;
    q.res.escalated = j.escalated ? 1 : 0;
    q.res.max_rel_residual = j.max_rel;
    q.t_done = Clock::now();
    q.res.queue_ms = ms_between(q.t_submit, q.t_admit);
    q.res.latency_ms = ms_between(q.t_submit, q.t_done);
}

void Server::fail_all(const std::string& msg) {
    std::lock_guard<std::mutex> g(mu_);
    fatal_ = msg;
    for (auto& kv : all_) {
        Request& q = *kv.second;
        if (q.done) continue;
        q.done = true;
        q.error = msg;
        q.res.status = NSLX_E_INTERNAL;
    }
    queue_.clear();
    running_.clear();
    done_cv_.notify_all();
}

void Server::run() {
    (void)hipSetDevice(ctx_->device);
    for (;;) {
        std::vector<std::shared_ptr<Request>> admit_now;
        {
            std::unique_lock<std::mutex> g(mu_);
            cv_.wait(g, [&] { return stop_ || !queue_.empty() || !running_.empty(); });
            if (stop_ && !queue_.empty()) {   // drop what was never admitted
                for (auto& q : queue_) {
                    q->done = true;
                    q->error = "server stopped";
                    q->res.status = NSLX_E_STATE;
                }
                queue_.clear();
                done_cv_.notify_all();
            }
            if (running_.empty() && (stop_ || !fatal_.empty())) break;
            // FIFO admission while the front request fits memory and RHS slots
            while (!stop_ && !queue_.empty()) {
                auto& q = queue_.front();
                if (used_bytes_ + q->bytes > cfg_.hbm_budget) break;
                // a job needs one contiguous run of RHS slots, reserved here so
                // that the next request of this round sees it taken; a request
                // that finds none waits at the front until jobs finish
                q->rhs0 = pool_->reserve_rhs(q->nb);
                if (q->rhs0 < 0) break;
                used_bytes_ += q->bytes;
                bytes_peak_ = std::max(bytes_peak_, used_bytes_);
                q->t_admit = Clock::now();
                running_[q->id] = q;
                admitted_peak_ = std::max<int64_t>(admitted_peak_, (int64_t)running_.size());
                admit_now.push_back(q);
                queue_.pop_front();
            }
        }
        const auto t0 = Clock::now();
        try {
            for (auto& q : admit_now) {
                try {
                    admit(*q);
                } catch (const Error& e) {
                    if (q->rhs0 >= 0) {   // failed before the pool took the slots
                        pool_->release_rhs(q->rhs0, q->nb);
                        q->rhs0 = -1;
                    }
                    std::lock_guard<std::mutex> g(mu_);
                    running_.erase(q->id);
                    used_bytes_ -= q->bytes;
                    q->done = true;
                    q->error = e.what();
                    q->res.status = e.status;
                    done_cv_.notify_all();
                }
            }
            if (!pool_->idle()) {
                auto fin = pool_->step();
                {
                    std::lock_guard<std::mutex> g(mu_);
                    iterations_ = pool_->stats().iterations;
                    replays_ = pool_->stats().replays;
                }
                for (auto& j : fin) {
                    std::shared_ptr<Request> q;
                    {
                        std::lock_guard<std::mutex> g(mu_);
                        q = running_[j->id];
                    }
                    complete(*q, *j);
                    std::lock_guard<std::mutex> g(mu_);
                    running_.erase(q->id);
                    used_bytes_ -= q->bytes;
                    q->done = true;
                    lat_.push_back(q->res.latency_ms);
                    rhs_solved_ += q->nb;
                    ++completed_;
                    done_cv_.notify_all();
                }
            }
        } catch (const std::exception& e) {
            fail_all(e.what());
            break;
        }
        busy_ += std::chrono::duration<double>(Clock::now() - t0).count();
    }
}

// One request on the time-slice-preconditioned path, on the worker's stream.
void Server::solve_tss_request(Request& q, hipStream_t s) {
    const nslx_request& r = q.r;
    q.fm = std::make_unique<FermionMatrix>();
    fm_init(*q.fm, ctx_, r.nx, r.nt, q.K.data(), r.beta, r.mu, r.species, DISC_EXP);
    // the persistent recurrence launch needs the device to itself
    q.fm->chains_ok = cfg_.precond_workers == 1;
    DevBuf<double2> phi((size_t)r.nt * r.nx);
    NSLX_HIP(hipMemcpyAsync(phi.p, q.phi.data(), phi.n * sizeof(double2), hipMemcpyHostToDevice, s));
    q.fm->populate(phi.p, 1, r.species, s);
    const size_t n = (size_t)q.nb * r.nt * r.nx;
    q.S.alloc(n);
    q.X.alloc(n);
    corr_sources(q.S.p, q.nsrc, 0, q.step, r.nt, r.nx, s);
    nslx_solve_report rep{};
    const int st = solve_tss(*q.fm, NSLX_OP_MMDAGGER, q.S.p, q.X.p, nullptr, nullptr, q.nb, r.tol, r.maxit,
                             r.precision, 0, &rep, s);
    correlate(q, s);
    q.res.status = st;
    q.res.iterations = rep.iterations;
    q.res.escalated = 0;
    q.res.max_rel_residual = rep.max_rel_residual;
    q.t_done = Clock::now();
    q.res.queue_ms = ms_between(q.t_submit, q.t_admit);
    q.res.latency_ms = ms_between(q.t_submit, q.t_done);
}

void Server::run_tss() {
    hipStream_t s = nullptr;
    if (hipSetDevice(ctx_->device) != hipSuccess || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        fail_all("server: could not create a worker stream");
        return;
    }
    for (;;) {
        std::shared_ptr<Request> q;
        {
            std::unique_lock<std::mutex> g(mu_);
            // FIFO: the front request when its footprint fits beside the
            // running ones (a request over the whole budget is refused at submit)
            cv_.wait(g, [&] {
                return stop_ || !fatal_.empty() ||
                       (!queue_.empty() && used_bytes_ + queue_.front()->bytes <= cfg_.hbm_budget);
            });
            if (stop_ || !fatal_.empty()) {
                for (auto& d : queue_) {   // drop what was never admitted
                    d->done = true;
                    d->error = stop_ ? "server stopped" : fatal_;
                    d->res.status = NSLX_E_STATE;
                }
                queue_.clear();
                done_cv_.notify_all();
                break;
            }
            q = queue_.front();
            queue_.pop_front();
            used_bytes_ += q->bytes;
            bytes_peak_ = std::max(bytes_peak_, used_bytes_);
            q->t_admit = Clock::now();
            running_[q->id] = q;
            admitted_peak_ = std::max<int64_t>(admitted_peak_, (int64_t)running_.size());
        }
        const auto t0 = Clock::now();
        std::string err;
        int err_status = NSLX_OK;
        try {
            solve_tss_request(*q, s);
        } catch (const Error& e) {
            err = e.what();
            err_status = e.status;
        } catch (const std::exception& e) {
            err = e.what();
            err_status = NSLX_E_INTERNAL;
        }
        if (!err.empty()) {
            q->S.reset();
            q->X.reset();
            q->G.reset();
            q->fm.reset();
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            running_.erase(q->id);
            used_bytes_ -= q->bytes;
            q->done = true;
            if (!err.empty()) {
                q->error = err;
                q->res.status = err_status;
            } else {
                lat_.push_back(q->res.latency_ms);
                rhs_solved_ += q->nb;
                ++completed_;
                iterations_ += q->res.iterations;
            }
            busy_ += std::chrono::duration<double>(Clock::now() - t0).count();
            done_cv_.notify_all();
        }
        cv_.notify_all();   // memory freed: the next request may fit
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
}

int Server::wait(int64_t id, double timeout_ms, void* corr, nslx_result* res) {
    std::shared_ptr<Request> q;
    {
        std::unique_lock<std::mutex> g(mu_);
        auto it = all_.find(id);
        NSLX_REQUIRE(it != all_.end(), "server_wait: unknown request id (or already fetched)");
        q = it->second;
        auto ready = [&] { return q->done; };
        if (timeout_ms < 0) done_cv_.wait(g, ready);
        else if (!done_cv_.wait_for(g, std::chrono::duration<double, std::milli>(timeout_ms), ready))
            throw Error(NSLX_E_STATE, "server_wait: timed out");
        all_.erase(it);
    }
    if (res) *res = q->res;
    if (!q->error.empty()) throw Error(q->res.status ? q->res.status : NSLX_E_INTERNAL, q->error);
    if (corr && q->corr.p)
        NSLX_HIP(hipMemcpy(corr, q->corr.p, q->corr.n * sizeof(double2), hipMemcpyDeviceToDevice));
    return q->res.status;
}

void Server::stats(nslx_server_stats* st) {
    std::lock_guard<std::mutex> g(mu_);
    nslx_server_stats o{};
    o.submitted = submitted_;
    o.completed = completed_;
    o.rhs_solved = rhs_solved_;
    o.admitted_peak = admitted_peak_;
    std::vector<double> l = lat_;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
        if (l.empty()) return 0.0;
        const size_t k = std::min(l.size() - 1, (size_t)std::ceil(p * (double)l.size()) - (p > 0 ? 1 : 0));
        return l[k];
    };
    o.p50_ms = pct(0.5);
    o.p99_ms = pct(0.99);
    o.max_ms = l.empty() ? 0.0 : l.back();
    o.busy_seconds = busy_;
    o.bytes_peak = bytes_peak_;
    o.iterations = iterations_;
    o.replays = replays_;
    *st = o;
}

}  // namespace nslx
