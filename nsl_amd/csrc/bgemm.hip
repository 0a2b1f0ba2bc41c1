// Batched complex NT GEMM on CDNA4 matrix cores (see BJob in tss.hpp), the
// engine of the time-slice solver: chunk products, the forward/back sweeps
// of the reduced block system, the Gauss-Jordan trailing updates and (where
// the persistent recurrence launch of chain.hip does not fit) the in-chunk
// recurrences are all launches of this kernel over job lists.
//
// A workgroup (4 waves, 2 x 2) owns one BM x BN tile of one job; the K loop
// stages BK columns of A (16 for 64 x 64 tiles, 32 for 32 x 32; re / im
// planes, phase applied by the loader) and of B (re [/ im] planes) in LDS,
// the next chunk's global loads in flight during the current chunk's MFMAs,
// and feeds v_mfma_f64_16x16x4f64 (FP64) or
// v_mfma_f32_16x16x4f32 (FP32).  Complex A x real B is two real products
// (re, im share B); complex x complex is four (rr, ii, ri, ir), combined in
// the epilogue.
#include <cmath>
#include <cstdlib>
#include <functional>
#include <type_traits>

#include "prof.hpp"
#include "tss.hpp"

namespace nslx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <class S> struct Acc;
template <> struct Acc<double> { using T = d4; };
template <> struct Acc<float> { using T = f4; };

template <class S>
struct C2 { S re, im; };

template <class S>
__device__ inline C2<S> ldc(const void* base, long long idx) {
    if constexpr (sizeof(S) == 8) {
        const double2 v = reinterpret_cast<const double2*>(base)[idx];
        return {v.x, v.y};
    } else {
        const float2 v = reinterpret_cast<const float2*>(base)[idx];
        return {v.x, v.y};
    }
}
template <class S>
__device__ inline void stc(void* base, long long idx, C2<S> v) {
    if constexpr (sizeof(S) == 8) reinterpret_cast<double2*>(base)[idx] = make_double2(v.re, v.im);
    else reinterpret_cast<float2*>(base)[idx] = make_float2(v.re, v.im);
}

template <class S>
__device__ inline typename Acc<S>::T mfma(S a, S b, typename Acc<S>::T c) {
    if constexpr (sizeof(S) == 8) return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int PAD = 1;

// K chunk staged per step: 32 for the 32 x 32 tile (the skinny reduced-sweep
// and recurrence launches, where the K loop is long), 16 for 64 x 64 (LDS)
template <int BM> struct KChunk { static constexpr int BK = BM == 32 ? 32 : 16; };

template <class S, bool CB, int PH, int BM, int BN>
__global__ void __launch_bounds__(256) k_bgemm(const BJob* __restrict__ jobs, const int4* __restrict__ tiles) {
    constexpr int BK = KChunk<BM>::BK;
    constexpr int FM = BM / 32, FN = BN / 32;          // 16 x 16 blocks per wave (2 x 2 waves)
    constexpr int NPB = CB ? 2 : 1;
    constexpr int NA = BM * BK / 256, NB = BN * BK / 256;   // staged elements per thread
    __shared__ S sA[2][BM][BK + PAD];
    __shared__ S sB[NPB][BN][BK + PAD];
    using AT = typename Acc<S>::T;
    const int4 tl = tiles[blockIdx.x];
    const BJob& J = jobs[tl.x];
    const int rows = J.rows, cols = J.cols, K = J.K;
    const int r0 = tl.y * BM, c0 = tl.z * BN;
    if (J.zf) {
        // the tile's rows of B all zero: Y = Add = Y (uniform over the block)
        bool nz = false;
        for (int t = c0 / 32; t * 32 < min(c0 + BN, cols); ++t) nz = nz || J.zf[t] != 0;
        if (!nz) return;
    }
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int r16 = lane & 15, q = lane >> 4;

    AT acc[CB ? 4 : 2][FM][FN];
    #pragma unroll
    for (int a = 0; a < (CB ? 4 : 2); ++a)
        #pragma unroll
        for (int i = 0; i < FM; ++i)
            #pragma unroll
            for (int j = 0; j < FN; ++j) acc[a][i][j] = AT{0, 0, 0, 0};

    // staging coordinates of this thread, fixed over the K loop: its element
    // e is tile row (B: tile column) rb + e * RS at K offset kk of a chunk, so
    // the row addresses and phase rows are computed once, not per chunk
    constexpr int RS = 256 / BK;
    const int kk = tid % BK, rb = tid / BK;
    const size_t ce = 2 * sizeof(S);
    const char* A1base = static_cast<const char*>(J.A1 ? J.A1 : J.A0);
    const char* pa0[NA];
    const char* pa1[NA];
    const char* pph[NA];
    bool ra[NA];
    #pragma unroll
    for (int e = 0; e < NA; ++e) {
        const int r = r0 + rb + e * RS;
        ra[e] = r < rows;
        const long long rr = ra[e] ? r : 0;
        pa0[e] = static_cast<const char*>(J.A0) + rr * J.lda0 * ce;
        pa1[e] = A1base + rr * J.lda1 * ce;
        if constexpr (PH == 1) pph[e] = static_cast<const char*>(J.ph) + (rr / J.phdiv) * J.ldph * ce;
        else pph[e] = nullptr;
    }
    const char* pb[NB];
    bool cv[NB];
    #pragma unroll
    for (int e = 0; e < NB; ++e) {
        const int c = c0 + rb + e * RS;
        cv[e] = c < cols;
        pb[e] = static_cast<const char*>(J.B) + (long long)(cv[e] ? c : 0) * J.ldb * (CB ? ce : sizeof(S));
    }
    const int kseg = J.kseg;

    // register stage of one K chunk: the next chunk's global loads are issued
    // before the current chunk's MFMAs and stored to LDS after them
    C2<S> ga[NA];
    C2<S> gb[NB];
    auto load = [&](int k0) {
        const int k = k0 + kk;
        const bool kin = k < K;
        #pragma unroll
        for (int e = 0; e < NA; ++e) {
            C2<S> z{S(0), S(0)};
            if (ra[e] && kin) {
                z = k < kseg ? ldc<S>(pa0[e], k) : ldc<S>(pa1[e], k - kseg);
                if constexpr (PH == 1) {
                    const C2<S> p = ldc<S>(pph[e], k);
                    const S re = p.re * z.re - p.im * z.im;
                    const S im = p.re * z.im + p.im * z.re;
                    z.re = re;
                    z.im = im;
                }
            }
            ga[e] = z;
        }
        #pragma unroll
        for (int e = 0; e < NB; ++e) {
            C2<S> z{S(0), S(0)};
            if (cv[e] && kin) {
                if constexpr (CB) z = ldc<S>(pb[e], k);
                else z.re = reinterpret_cast<const S*>(pb[e])[k];
            }
            gb[e] = z;
        }
    };
    auto store = [&]() {
        #pragma unroll
        for (int e = 0; e < NA; ++e) {
            sA[0][rb + e * RS][kk] = ga[e].re;
            sA[1][rb + e * RS][kk] = ga[e].im;
        }
        #pragma unroll
        for (int e = 0; e < NB; ++e) {
            sB[0][rb + e * RS][kk] = gb[e].re;
            if constexpr (CB) sB[1][rb + e * RS][kk] = gb[e].im;
        }
    };
    // A 32 x 32 tile whose live rows fit its upper half (the reduced sweeps'
    // 8 right-hand sides per field) splits each chunk's K between the wave
    // rows instead: both wave rows compute rows 0-15, over ks halves, and the
    // halves are added through LDS before the epilogue.  The 16 x 16 x 4 FP64
    // MFMAs are what such a launch waits for; this halves them per wave.
    const bool ksplit = BM == 32 && rows - r0 <= BM / 2;
    auto compute_ks = [&](auto ks0c, auto ks1c, int wrow) {
        constexpr int KS0 = decltype(ks0c)::value, KS1 = decltype(ks1c)::value;
        #pragma unroll
        for (int ks = KS0; ks < KS1; ++ks) {
            const int k = 4 * ks + q;
            S ar[FM], ai[FM], br[FN], bi[FN];
            #pragma unroll
            for (int i = 0; i < FM; ++i) {
                const int row = wrow + 16 * i + r16;
                ar[i] = sA[0][row][k];
                ai[i] = sA[1][row][k];
            }
            #pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int col = wn * (BN / 2) + 16 * j + r16;
                br[j] = sB[0][col][k];
                if constexpr (CB) bi[j] = sB[1][col][k];
            }
            #pragma unroll
            for (int i = 0; i < FM; ++i)
                #pragma unroll
                for (int j = 0; j < FN; ++j) {
                    acc[0][i][j] = mfma<S>(ar[i], br[j], acc[0][i][j]);
                    acc[1][i][j] = mfma<S>(ai[i], br[j], acc[1][i][j]);
                    if constexpr (CB) {
                        acc[2][i][j] = mfma<S>(ai[i], bi[j], acc[2][i][j]);
                        acc[3][i][j] = mfma<S>(ar[i], bi[j], acc[3][i][j]);
                    }
                }
        }
    };
    using std::integral_constant;
    auto compute = [&]() {
        if (ksplit) {
            if (wm == 0) compute_ks(integral_constant<int, 0>{}, integral_constant<int, BK / 8>{}, 0);
            else compute_ks(integral_constant<int, BK / 8>{}, integral_constant<int, BK / 4>{}, 0);
        } else {
            compute_ks(integral_constant<int, 0>{}, integral_constant<int, BK / 4>{}, wm * (BM / 2));
        }
    };
    // the Add operand (the Gauss-Jordan updates' trailing matrix, the sweeps'
    // accumulators) is read before the K loop, so its latency overlaps the
    // MFMAs: read in the epilogue, each load waited behind the previous
    // element's store (Add may be Y, so the compiler keeps that order), which
    // for the K = 16 updates was most of the launch
    C2<S> addv[FM][FN][4];
    const bool has_add = J.Add != nullptr;
    if (has_add) {
        #pragma unroll
        for (int i = 0; i < FM; ++i)
            #pragma unroll
            for (int j = 0; j < FN; ++j)
                #pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rr = (sizeof(S) == 8) ? 4 * v + q : 4 * q + v;
                    const int r = r0 + wm * (BM / 2) + 16 * i + rr;
                    const int c = c0 + wn * (BN / 2) + 16 * j + r16;
                    addv[i][j][v] = (r < rows && c < cols) ? ldc<S>(J.Add, (long long)r * J.ldadd + c)
                                                           : C2<S>{S(0), S(0)};
                }
    }
    load(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
        store();
        __syncthreads();
        if (k0 + BK < K) load(k0 + BK);
        compute();
        __syncthreads();
    }

    if (ksplit) {
        // the second wave row's K halves to the first (LDS free since the K
        // loop's last barrier); the second wave row is then done
        constexpr int NACC = (CB ? 4 : 2) * FM * FN * 4;
        static_assert(BM != 32 || 2 * NACC * 64 <= 2 * BM * (BK + PAD), "partials fit sA");
        S* red = &sA[0][0][0];
        if (wm == 1) {
            #pragma unroll
            for (int a = 0; a < (CB ? 4 : 2); ++a)
                #pragma unroll
                for (int i = 0; i < FM; ++i)
                    #pragma unroll
                    for (int j = 0; j < FN; ++j)
                        #pragma unroll
                        for (int v = 0; v < 4; ++v)
                            red[(wn * NACC + ((a * FM + i) * FN + j) * 4 + v) * 64 + lane] = acc[a][i][j][v];
        }
        __syncthreads();
        if (wm == 1) return;
        #pragma unroll
        for (int a = 0; a < (CB ? 4 : 2); ++a)
            #pragma unroll
            for (int i = 0; i < FM; ++i)
                #pragma unroll
                for (int j = 0; j < FN; ++j)
                    #pragma unroll
                    for (int v = 0; v < 4; ++v)
                        acc[a][i][j][v] += red[(wn * NACC + ((a * FM + i) * FN + j) * 4 + v) * 64 + lane];
    }
    // epilogue: register v of lane (q, r16) holds row 4v + q (FP64) or
    // 4q + v (FP32) of its 16 x 16 block, column r16
    const S sg = S(J.sgn);
    #pragma unroll
    for (int i = 0; i < FM; ++i)
        #pragma unroll
        for (int j = 0; j < FN; ++j)
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rr = (sizeof(S) == 8) ? 4 * v + q : 4 * q + v;
                const int r = r0 + wm * (BM / 2) + 16 * i + rr;
                const int c = c0 + wn * (BN / 2) + 16 * j + r16;
                if (r >= rows || c >= cols) continue;
                S re, im;
                if constexpr (CB) {
                    re = S(acc[0][i][j][v]) - S(acc[2][i][j][v]);
                    im = S(acc[1][i][j][v]) + S(acc[3][i][j][v]);
                } else {
                    re = S(acc[0][i][j][v]);
                    im = S(acc[1][i][j][v]);
                }
                if constexpr (PH == 2) {
                    const C2<S> p = ldc<S>(J.ph, (long long)(r / J.phdiv) * J.ldph + c);
                    const S tr = p.re * re + p.im * im;   // conj(p) * (re + i im)
                    const S ti = p.re * im - p.im * re;
                    re = tr;
                    im = ti;
                }
                C2<S> out{sg * re, sg * im};
                if (has_add) {
                    out.re = out.re + addv[i][j][v].re;
                    out.im = out.im + addv[i][j][v].im;
                }
                stc<S>(J.Y, (long long)r * J.ldy + c, out);
            }
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
    NSLX_HIP(hipMemsetAsync(err_.p, 0, sizeof(int), s));
    NSLX_HIP(hipMemsetAsync(sync_.p, 0, sync_words_ * sizeof(unsigned), s));
    NSLX_HIP(hipStreamSynchronize(s));
    return true;
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
        if (op.is_gemm) launch_bgemm(op.v, djobs_.p + op.job0, dtiles_.p + op.tile0, op.ntiles, s, op.flops, op.bytes);
        else if (op.is_chain)
            launch_chain(op.v, djobs_.p + op.job0, op.steps, op.nchain, op.rows, op.cols, sync_.p, err_.p, s, op.flops,
                         op.bytes);
        else op.fn(s);
    }
}

}  // namespace nslx
