// Batched complex NT GEMM on CDNA4 matrix cores (see BJob in tss.hpp), the
// engine of the time-slice solver: chunk products, the forward/back sweeps
// of the reduced block system, the Gauss-Jordan trailing updates and (where
// the persistent recurrence launch of chain.hip does not fit) the in-chunk
// recurrences are all launches of this kernel over job lists.  The tile
// itself (bgemm_tile) is in bgemm_tile.hpp.
#include <cmath>
#include <cstdlib>
#include <functional>

#include "bgemm_tile.hpp"
#include "prof.hpp"
#include "tss.hpp"

namespace nslx {
namespace {

template <class S, bool CB, int PH, int BM, int BN>
__global__ void __launch_bounds__(256) k_bgemm(const BJob* __restrict__ jobs, const int4* __restrict__ tiles) {
    bg::bgemm_tile<S, CB, PH, BM, BN>(jobs, tiles, blockIdx.x);
}

template <class S, bool CB, int PH>
void launch_t(int tile, const BJob* j, const int4* t, int n, hipStream_t s) {
    if (tile == 0) launch_timed(k_bgemm<S, CB, PH, 64, 64>, dim3(n), dim3(256), 0, s, j, t);
    else launch_timed(k_bgemm<S, CB, PH, 32, 32>, dim3(n), dim3(256), 0, s, j, t);
}

template <class S>
void launch_s(const BVariant& v, const BJob* j, const int4* t, int n, hipStream_t s) {
    if (v.cb) {
        if (v.ph == 0) launch_t<S, true, 0>(v.tile, j, t, n, s);
        else if (v.ph == 1) launch_t<S, true, 1>(v.tile, j, t, n, s);
        else launch_t<S, true, 2>(v.tile, j, t, n, s);
    } else {
        if (v.ph == 0) launch_t<S, false, 0>(v.tile, j, t, n, s);
        else if (v.ph == 1) launch_t<S, false, 1>(v.tile, j, t, n, s);
        else launch_t<S, false, 2>(v.tile, j, t, n, s);
    }
}

}  // namespace

int bgemm_tile_rows(int tile) { return tile == 0 ? 64 : 32; }

void bgemm_tiles(const std::vector<BJob>& jobs, int tile, std::vector<int4>& out) {
    const int bm = bgemm_tile_rows(tile), bn = bm;
    for (size_t k = 0; k < jobs.size(); ++k) {
        const BJob& j = jobs[k];
        if (j.rows <= 0 || j.cols <= 0) continue;
        const int nr = (j.rows + bm - 1) / bm, nc = (j.cols + bn - 1) / bn;
        for (int r = 0; r < nr; ++r)
            for (int c = 0; c < nc; ++c) out.push_back(make_int4((int)k, r, c, 0));
    }
}

std::string bgemm_class(const BVariant& v) {
    return std::string("bgemm_") + (v.prec == NSLX_F64 ? "f64" : "f32") + (v.cb ? "_cplx" : "_real") + "_ph" +
           std::to_string(v.ph) + (v.tile == 0 ? "_64" : "_32");
}

void bgemm_work(const std::vector<BJob>& jobs, const BVariant& v, double* flops, double* bytes) {
    const double ce = v.prec == NSLX_F64 ? 16.0 : 8.0, be = v.cb ? ce : ce / 2;
    double f = 0.0, b = 0.0;
    for (const BJob& j : jobs) {
        const double r = j.rows, c = j.cols, k = j.K;
        f += (v.cb ? 8.0 : 4.0) * r * c * k;
        b += r * k * ce + c * k * be + r * c * ce * (j.Add ? 2.0 : 1.0);
        if (v.ph) b += std::ceil(r / std::min<double>(j.phdiv, r)) * (v.ph == 1 ? k : c) * ce;
    }
    *flops = f;
    *bytes = b;
}

void launch_bgemm(const BVariant& v, const BJob* d_jobs, const int4* d_tiles, int ntiles, hipStream_t s, double flops,
                  double bytes) {
    if (ntiles <= 0) return;
    NSLX_REQUIRE(v.tile == 0 || v.tile == 1, "bgemm: bad tile");
    ProfScope prof(prof_on() ? bgemm_class(v) : std::string(), s);
    prof.work(flops, bytes);
    if (v.prec == NSLX_F64) launch_s<double>(v, d_jobs, d_tiles, ntiles, s);
    else if (v.prec == NSLX_F32) launch_s<float>(v, d_jobs, d_tiles, ntiles, s);
    else throw Error(NSLX_E_ARG, "bgemm: precision must be F64 or F32");
    NSLX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------- BPlan
void BPlan::gemm(const std::vector<BJob>& jobs, BVariant v, int tile) {
    if (tile < 0) {
        // 64 x 64 tiles when they alone give two workgroups per CU of MI355X
        std::vector<int4> t0;
        bgemm_tiles(jobs, 0, t0);
        tile = t0.size() >= 512 ? 0 : 1;
    }
    v.tile = tile;
    Op op;
    op.is_gemm = true;
    op.v = v;
    op.job0 = jobs_.size();
    op.njobs = jobs.size();
    op.tile0 = tiles_.size();
    // tile job indices are local to this launch: run() passes the launch's
    // own slice of the job array
    std::vector<int4> t;
    bgemm_tiles(jobs, tile, t);
    for (const auto& x : t) NSLX_REQUIRE(x.x >= 0 && x.x < (int)jobs.size(), "bgemm plan: tile job index out of range");
    jobs_.insert(jobs_.end(), jobs.begin(), jobs.end());
    tiles_.insert(tiles_.end(), t.begin(), t.end());
    op.ntiles = (int)t.size();
    bgemm_work(jobs, v, &op.flops, &op.bytes);
    ops_.push_back(std::move(op));
    uploaded_ = false;
}

void BPlan::chain(const std::vector<std::vector<BJob>>& steps, BVariant v, bool allow) {
    if (steps.empty()) return;
    const size_t nch = steps[0].size();
    NSLX_REQUIRE(nch >= 1, "chain plan: no jobs in a step");
    const BJob& j0 = steps[0][0];
    for (size_t s = 0; s < steps.size(); ++s) {
        NSLX_REQUIRE(steps[s].size() == nch, "chain plan: steps differ in their job count");
        for (size_t i = 0; i < nch; ++i) {
            const BJob& j = steps[s][i];
            NSLX_REQUIRE(j.B == j0.B && j.ldb == j0.ldb && j.rows == j0.rows && j.cols == j0.cols && j.K == j0.K,
                         "chain plan: steps differ in B or shape");
            NSLX_REQUIRE(j.K <= j.kseg, "chain plan: split A rows");
            if (s > 0) {
                const BJob& p = steps[s - 1][i];
                NSLX_REQUIRE(j.A0 == p.Y && j.lda0 == p.ldy && j.K == j.cols,
                             "chain plan: a step does not read the previous step's rows");
            }
        }
    }
    const char* env = std::getenv("NSLX_CHAINS");   // read per plan: tests compare both paths
    const bool enabled = !(env && env[0] == '0');
    v.tile = 1;
    if (!allow || !enabled || v.cb || !chain_supported(v, (int)steps.size(), (int)nch, j0.rows, j0.cols, j0.K)) {
        for (const auto& st : steps) gemm(st, v);
        return;
    }
    Op op;
    op.is_chain = true;
    op.v = v;
    op.job0 = jobs_.size();
    op.njobs = steps.size() * nch;
    op.steps = (int)steps.size();
    op.nchain = (int)nch;
    op.rows = j0.rows;
    op.cols = j0.cols;
    for (const auto& st : steps) {
        jobs_.insert(jobs_.end(), st.begin(), st.end());
        double f = 0.0, b = 0.0;
        bgemm_work(st, v, &f, &b);
        op.flops += f;
        op.bytes += b;
    }
    sync_words_ = std::max(sync_words_, chain_sync_words(op.nchain, op.rows));
    ops_.push_back(std::move(op));
    uploaded_ = false;
}

bool BPlan::chain_failed(hipStream_t s) {
    if (!err_.p) return false;
    int e = 0;
    NSLX_HIP(hipMemcpyAsync(&e, err_.p, sizeof(int), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    if (!e) return false;
    chain_reset(s);
    return true;
}

void BPlan::chain_reset(hipStream_t s) {
    if (!err_.p) return;
    NSLX_HIP(hipMemsetAsync(err_.p, 0, sizeof(int), s));
    NSLX_HIP(hipMemsetAsync(sync_.p, 0, sync_words_ * sizeof(unsigned), s));
    NSLX_HIP(hipStreamSynchronize(s));
}

void BPlan::gemm_with(const std::vector<BJob>& jobs, BVariant v, int tile, GemmLauncher launch) {
    gemm(jobs, v, tile);
    ops_.back().fused = std::move(launch);
}

void BPlan::custom(std::function<void(hipStream_t)> fn) {
    Op op;
    op.fn = std::move(fn);
    ops_.push_back(std::move(op));
}

namespace {
// Extents (in bytes) of every buffer a job touches, checked against the
// library's own device allocations when NSLX_CHECK_PLANS is set: a plan whose
// indexing is wrong fails on the host before anything is launched.
void check_job(const BJob& j, const BVariant& v, size_t idx) {
    const size_t ce = v.prec == NSLX_F64 ? 16 : 8;         // complex element
    const size_t be = v.cb ? ce : ce / 2;                   // B element
    auto need = [&](const void* p, long long last_elem, size_t es, const char* what) {
        if (!p) throw Error(NSLX_E_INTERNAL, std::string("bgemm plan: null ") + what);
        if (last_elem < 0) return;
        if (!alloc_covers(p, (size_t)(last_elem + 1) * es))
            throw Error(NSLX_E_INTERNAL, "bgemm plan: job " + std::to_string(idx) + " reads/writes " + what +
                                             " outside any device allocation");
    };
    if (j.rows <= 0 || j.cols <= 0 || j.K <= 0) return;
    const int k0 = std::min(j.K, j.kseg);
    need(j.A0, (long long)(j.rows - 1) * j.lda0 + k0 - 1, ce, "A0");
    if (j.K > j.kseg) need(j.A1, (long long)(j.rows - 1) * j.lda1 + (j.K - j.kseg) - 1, ce, "A1");
    need(j.B, (long long)(j.cols - 1) * j.ldb + j.K - 1, be, "B");
    need(j.Y, (long long)(j.rows - 1) * j.ldy + j.cols - 1, ce, "Y");
    if (j.Add) need(j.Add, (long long)(j.rows - 1) * j.ldadd + j.cols - 1, ce, "Add");
    if (v.ph) {
        const long long prow = (long long)((j.rows - 1) / j.phdiv) * j.ldph;
        need(j.ph, prow + (v.ph == 1 ? j.K : j.cols) - 1, ce, "phase");
    }
}
}  // namespace

void BPlan::upload(hipStream_t s) {
    if (alloc_checks_on()) {
        for (const Op& op : ops_)
            if (op.is_gemm) {
                const size_t nj = (op.job0 < jobs_.size()) ? jobs_.size() - op.job0 : 0;
                for (int t = 0; t < op.ntiles; ++t) {
                    const int4 tl = tiles_[op.tile0 + t];
                    NSLX_REQUIRE(tl.x >= 0 && (size_t)tl.x < nj, "bgemm plan: tile job index out of range");
                }
            }
        for (const Op& op : ops_)
            if (op.is_gemm || op.is_chain)
                for (size_t k = op.job0; k < op.job0 + op.njobs; ++k) check_job(jobs_[k], op.v, k);
    }
    djobs_.alloc(std::max<size_t>(jobs_.size(), 1));
    dtiles_.alloc(std::max<size_t>(tiles_.size(), 1));
    if (sync_words_) {
        sync_.alloc(sync_words_);
        err_.alloc(1);
        NSLX_HIP(hipMemsetAsync(sync_.p, 0, sync_words_ * sizeof(unsigned), s));
        NSLX_HIP(hipMemsetAsync(err_.p, 0, sizeof(int), s));
    }
    if (!jobs_.empty())
        NSLX_HIP(hipMemcpyAsync(djobs_.p, jobs_.data(), jobs_.size() * sizeof(BJob), hipMemcpyHostToDevice, s));
    if (!tiles_.empty())
        NSLX_HIP(hipMemcpyAsync(dtiles_.p, tiles_.data(), tiles_.size() * sizeof(int4), hipMemcpyHostToDevice, s));
    NSLX_HIP(hipStreamSynchronize(s));
    uploaded_ = true;
}

void BPlan::run(hipStream_t s) const {
    NSLX_REQUIRE(uploaded_, "bgemm plan run before upload");
    for (const Op& op : ops_) {
        if (op.is_gemm && op.fused) op.fused(s, djobs_.p + op.job0, dtiles_.p + op.tile0, op.ntiles, op.v);
        else if (op.is_gemm) launch_bgemm(op.v, djobs_.p + op.job0, dtiles_.p + op.tile0, op.ntiles, s, op.flops, op.bytes);
        else if (op.is_chain)
            launch_chain(op.v, djobs_.p + op.job0, op.steps, op.nchain, op.rows, op.cols, sync_.p, err_.p, s, op.flops,
                         op.bytes);
        else op.fn(s);
    }
}

}  // namespace nslx
