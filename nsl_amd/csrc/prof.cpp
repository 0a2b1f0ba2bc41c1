// HIP-event timing of kernel launches (enabled by nslx_prof_enable).
#include <map>
#include <mutex>
#include <sstream>

#include "prof.hpp"

namespace nslx {
namespace {
struct Rec {
    std::string name;
    hipEvent_t a, b;
    double flops, bytes;
};
std::mutex g_mu;
int g_every = 0;                       // 0: off; n: time every n-th launch of a class
std::map<std::string, long> g_count;   // launches per class while on
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    NSLX_HIP(hipEventCreate(&e));
    return e;
}
}  // namespace

bool prof_on() { return g_every > 0; }
int prof_every() { return g_every; }

namespace {
thread_local ProfScope* t_scope = nullptr;
}

ProfScope* prof_current() { return t_scope; }

ProfScope::ProfScope(const char* name, int prec, int mode, hipStream_t s) : active_(g_every > 0) {
    (void)s;
    outer_ = t_scope;
    t_scope = this;
    if (!active_) return;
    static const char* pn[] = {"f64", "f32", "bf16x3"};
    static const char* mn[] = {"plain", "cga", "cgb", "cgi"};
    std::lock_guard<std::mutex> g(g_mu);
    name_ = std::string(name) + "_" + pn[prec % 3] + "_" + mn[mode % 4];
    begin();
}

ProfScope::ProfScope(const std::string& full_name, hipStream_t s) : active_(g_every > 0) {
    (void)s;
    outer_ = t_scope;
    t_scope = this;
    if (!active_) return;
    std::lock_guard<std::mutex> g(g_mu);
    name_ = full_name;
    begin();
}

// with g_mu held: count the launch; every g_every-th of a class is timed
void ProfScope::begin() {
    const long n = g_count[name_]++;
    if (n % g_every != 0) {
        active_ = false;
        return;
    }
    a_ = take();
    b_ = take();
}

bool ProfScope::take_events(hipEvent_t* start, hipEvent_t* stop) {
    if (!active_ || used_) return false;
    used_ = true;
    *start = (hipEvent_t)a_;
    *stop = (hipEvent_t)b_;
    return true;
}

ProfScope::~ProfScope() {
    t_scope = outer_;
    if (!active_) return;
    std::lock_guard<std::mutex> g(g_mu);
    if (used_) {
        g_recs.push_back({name_, (hipEvent_t)a_, (hipEvent_t)b_, flops_, bytes_});
    } else {   // no kernel was launched in the scope (an error path)
        g_pool.push_back((hipEvent_t)a_);
        g_pool.push_back((hipEvent_t)b_);
    }
}

void prof_enable(int every) { g_every = every > 0 ? every : 0; }

void prof_reset() {
    std::lock_guard<std::mutex> g(g_mu);
    for (auto& r : g_recs) { g_pool.push_back(r.a); g_pool.push_back(r.b); }
    g_recs.clear();
    g_count.clear();
}

std::string prof_report(int* nclasses) {
    std::lock_guard<std::mutex> g(g_mu);
    struct Sum { long n = 0; double ms = 0.0, flops = 0.0, bytes = 0.0; };
    std::map<std::string, Sum> acc;
    for (auto& r : g_recs) {
        NSLX_HIP(hipEventSynchronize(r.b));
        float ms = 0.f;
        NSLX_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        auto& e = acc[r.name];
        e.n += 1;
        e.ms += ms;
        e.flops += r.flops;
        e.bytes += r.bytes;
    }
    std::ostringstream os;
    os.precision(17);
    for (auto& kv : acc) {
        const auto c = g_count.find(kv.first);
        os << kv.first << " " << (c == g_count.end() ? kv.second.n : c->second) << " " << kv.second.ms << " "
           << kv.second.n << " " << kv.second.flops << " " << kv.second.bytes << "\n";
    }
    if (nclasses) *nclasses = (int)acc.size();
    return os.str();
}

}  // namespace nslx
