// The tile of the batched complex NT GEMM (k_bgemm in bgemm.hip), as a
// device function shared with the fused Gauss-Jordan launch (tss.hip).
//
// A workgroup (4 waves, 2 x 2) owns one BM x BN tile of one job; the K loop
// stages BK columns of A (16 for 64 x 64 tiles, 32 for 32 x 32; re / im
// planes, phase applied by the loader) and of B (re [/ im] planes) in LDS,
// the next chunk's global loads in flight during the current chunk's MFMAs,
// and feeds v_mfma_f64_16x16x4f64 (FP64) or
// v_mfma_f32_16x16x4f32 (FP32).  Complex A x real B is two real products
// (re, im share B); complex x complex is four (rr, ii, ri, ir), combined in
// the epilogue.
#pragma once

#include <type_traits>

#include "tss.hpp"

namespace nslx {
namespace bg {

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

// One BM x BN tile (tiles[bid]) of a batched launch, 256 threads.  Called by
// k_bgemm and by the fused Gauss-Jordan launch (tss.hip), whose other
// workgroups run a pivot panel beside the update tiles.
template <class S, bool CB, int PH, int BM, int BN>
__device__ __attribute__((always_inline)) inline void bgemm_tile(const BJob* __restrict__ jobs, const int4* __restrict__ tiles, int bid) {
    constexpr int BK = KChunk<BM>::BK;
    constexpr int FM = BM / 32, FN = BN / 32;          // 16 x 16 blocks per wave (2 x 2 waves)
    constexpr int NPB = CB ? 2 : 1;
    constexpr int NA = BM * BK / 256, NB = BN * BK / 256;   // staged elements per thread
    __shared__ S sA[2][BM][BK + PAD];
    __shared__ S sB[NPB][BN][BK + PAD];
    using AT = typename Acc<S>::T;
    const int4 tl = tiles[bid];
    const BJob& J = jobs[tl.x];
    const int rows = J.rows, cols = J.cols, K = J.K;
    const int r0 = tl.y * BM, c0 = tl.z * BN;
    if (J.zf) {
        // the tile's rows of B all zero: Y = Add = Y (uniform over the block)
        bool nz = false;
        const int e = min(c0 + BN, cols) + J.zfoff;
        for (int t = (c0 + J.zfoff) / 32; t * 32 < e; ++t) nz = nz || J.zf[t] != 0;
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

}  // namespace bg
}  // namespace nslx
