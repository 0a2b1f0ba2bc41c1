// Correlator server (see serve.cpp and nslx_server_* in nslx.h).
#pragma once

#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"
#include "pool.hpp"

namespace nslx {

struct Request;

class Server {
   public:
    Server(Ctx* ctx, const nslx_server_config& cfg);
    ~Server();
    int64_t submit(const nslx_request& r);
    int wait(int64_t id, double timeout_ms, void* corr, nslx_result* res);
    void stats(nslx_server_stats* st);

   private:
    void run();
    void admit(Request& q);
    void complete(Request& q, PoolJob& j);
    // G = M^dag X, the correlator, and the request's buffers released
    void correlate(Request& q, hipStream_t s);
    void fail_all(const std::string& msg);
    // time-slice-preconditioned mode: one worker per stream
    void run_tss();
    void solve_tss_request(Request& q, hipStream_t s);

    Ctx* ctx_;
    nslx_server_config cfg_;
    hipStream_t s_ = nullptr;
    std::unique_ptr<Pool> pool_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::shared_ptr<Request>> queue_;
    std::unordered_map<int64_t, std::shared_ptr<Request>> all_;       // submitted, not yet fetched
    std::unordered_map<int64_t, std::shared_ptr<Request>> running_;   // admitted, by id
    double used_bytes_ = 0.0, bytes_peak_ = 0.0, busy_ = 0.0;
    int64_t next_id_ = 0, submitted_ = 0, completed_ = 0, rhs_solved_ = 0, admitted_peak_ = 0;
    int64_t iterations_ = 0, replays_ = 0;
    std::vector<double> lat_;
    std::string fatal_;
    bool stop_ = false;
    std::thread th_;
    std::vector<std::thread> workers_;
};

}  // namespace nslx

struct nslx_server_s {
    std::unique_ptr<nslx::Server> s;
};
