// Chunk products of the time-slice solver by row blocks that own whole rows.
//
// Row i of Tt_j is T_j e_i with T_j = A_{t0+c-1} ... A_{t0} and
// A_t = b_t E diag(P_{t-1}), so each row evolves on its own:
//     v <- b_t E (P_{t-1} . v),  v = e_i at the start,
// and a workgroup that owns 16 rows and ALL their columns needs nothing from
// any other workgroup for the whole chunk.  It keeps its 16 rows (phase
// applied, re / im planes) in LDS across all c steps, streams E from L2 each
// step and writes only the final T_j rows to memory: one launch instead of c,
// no HBM round trip of the product between steps, no padding of the column
// dimension beyond a multiple of 16.
//
// Per step the 4 waves split the 16-column blocks of the output (block
// w + 4b for wave w); each block is two real products (re, im share E) on
// v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32.  The K order inside an
// MFMA is permuted (lane quad q supplies k0 + 4q + ks in step ks, for
// A from LDS and for E from memory alike), so a lane's four k values of a
// 16-wide chunk are contiguous: one 32-byte (FP64) / 16-byte (FP32) read per
// operand per chunk instead of four scattered ones.  The next chunk's E
// values are loaded while the current chunk's MFMAs run.
//
// Same arithmetic as the per-step batched GEMM (k_bgemm, PH = 1): identical
// products, accumulated in a different order, so results agree to rounding.
#include <algorithm>
#include <cstdlib>

#include "prof.hpp"
#include "tss.hpp"

namespace nslx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

template <class S> struct CP;
template <> struct CP<double> {
    using Acc = d4;
    static __device__ inline Acc mfma(double a, double b, Acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // output register v of lane (q, r16): row 4v + q, column r16
    static __device__ inline int row(int q, int v) { return 4 * v + q; }
};
template <> struct CP<float> {
    using Acc = f4;
    static __device__ inline Acc mfma(float a, float b, Acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    // output register v of lane (q, r16): row 4q + v, column r16
    static __device__ inline int row(int q, int v) { return 4 * q + v; }
};

constexpr int CPR = 16;   // rows per workgroup
constexpr int CPT = 256;  // threads per workgroup (4 waves)
constexpr int LPAD = 4;   // LDS row padding (elements)

// four consecutive S starting at p (p aligned to 4 elements)
template <class S>
__device__ inline void ld4(const S* p, S out[4]) {
    if constexpr (sizeof(S) == 8) {
        const d2v a = *reinterpret_cast<const d2v*>(p);
        const d2v b = *reinterpret_cast<const d2v*>(p + 2);
        out[0] = a.x; out[1] = a.y; out[2] = b.x; out[3] = b.y;
    } else {
        const f4 a = *reinterpret_cast<const f4*>(p);
        out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = a.w;
    }
}

// grid (ceil(n / 16), nc, nf).  E: exp(s delta K), row-major, ld ldE
// (zero-padded to at least roundup(n, 16) rows and columns); P: phases
// [nf][nt][n] (complex S); Tt: [nf][nc][n][n] (complex S).  LDS: 2 planes
// x 16 rows x (64 MAXB + LPAD) elements; np = roundup(n, 16) <= 64 MAXB.
template <class S, int MAXB>
__global__ void __launch_bounds__(CPT) k_chunkprod(int n, int nt, int c, const S* __restrict__ E, long long ldE,
                                                   const S* __restrict__ P, S* __restrict__ Tt) {
    using Acc = typename CP<S>::Acc;
    constexpr int ld = 64 * MAXB + LPAD;
    __shared__ __align__(16) S rows_lds[2][CPR * ld];
    const int np = (n + 15) & ~15;
    S* sre = rows_lds[0];
    S* sim = rows_lds[1];
    const int r0 = blockIdx.x * CPR, j = blockIdx.y, f = blockIdx.z;
    const int nc = gridDim.y;
    const int t0 = j * c;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, r16 = lane & 15;
    const int nblk = np / 16;
    const S* Pf = P + 2 * (size_t)f * nt * n;

    // v = e_{r0 + r} with the first step's phase P_{t0 - 1} applied
    for (int e = tid; e < CPR * ld; e += CPT) {
        const int r = e / ld, x = e % ld;
        S vr = S(0), vi = S(0);
        if (x == r0 + r && x < n) {
            const int tp = (t0 - 1 + nt) % nt;
            vr = Pf[2 * ((size_t)tp * n + x)];
            vi = Pf[2 * ((size_t)tp * n + x) + 1];
        }
        sre[e] = vr;
        sim[e] = vi;
    }
    __syncthreads();

    // this lane's E row pointers: block b covers output columns y = 16 * (wave + 4b) + r16
    const S* erow[MAXB];
    bool bon[MAXB];
    #pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int blk = wave + 4 * b;
        bon[b] = blk < nblk;
        erow[b] = E + (size_t)(bon[b] ? 16 * blk + r16 : 0) * ldE + 4 * q;
    }
    const S* arow_re = sre + r16 * ld + 4 * q;
    const S* arow_im = sim + r16 * ld + 4 * q;

    for (int s = 0; s < c; ++s) {
        const int t = t0 + s;
        Acc are[MAXB], aim[MAXB];
        #pragma unroll
        for (int b = 0; b < MAXB; ++b) { are[b] = Acc{0, 0, 0, 0}; aim[b] = Acc{0, 0, 0, 0}; }
        if (s == 0) {
            // the first step maps e_i to column i of E times P_{t0-1}[i]: no product needed
            #pragma unroll
            for (int b = 0; b < MAXB; ++b) {
                if (!bon[b]) continue;
                const int y = 16 * (wave + 4 * b) + r16;
                #pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = r0 + CP<S>::row(q, v);
                    if (i < n) {
                        const S e = E[(size_t)y * ldE + i];
                        are[b][v] = sre[(i - r0) * ld + i] * e;
                        aim[b][v] = sim[(i - r0) * ld + i] * e;
                    }
                }
            }
        } else {
        S ec[MAXB][4], en[MAXB][4];
        #pragma unroll
        for (int b = 0; b < MAXB; ++b)
            if (bon[b]) ld4<S>(erow[b], ec[b]);
        for (int k0 = 0; k0 < np; k0 += 16) {
            if (k0 + 16 < np) {
                #pragma unroll
                for (int b = 0; b < MAXB; ++b)
                    if (bon[b]) ld4<S>(erow[b] + k0 + 16, en[b]);
            }
            S xr[4], xi[4];
            ld4<S>(arow_re + k0, xr);
            ld4<S>(arow_im + k0, xi);
            #pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                #pragma unroll
                for (int b = 0; b < MAXB; ++b)
                    if (bon[b]) {
                        are[b] = CP<S>::mfma(xr[ks], ec[b][ks], are[b]);
                        aim[b] = CP<S>::mfma(xi[ks], ec[b][ks], aim[b]);
                    }
            #pragma unroll
            for (int b = 0; b < MAXB; ++b)
                #pragma unroll
                for (int v = 0; v < 4; ++v) ec[b][v] = en[b][v];
        }
        }
        __syncthreads();   // every wave is done reading this step's rows
        const S sg = t == 0 ? S(-1) : S(1);
        const bool last = s == c - 1;
        #pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            if (!bon[b]) continue;
            const int y = 16 * (wave + 4 * b) + r16;
            S pr = S(0), pi = S(0);
            if (!last && y < n) {
                pr = Pf[2 * ((size_t)t * n + y)];
                pi = Pf[2 * ((size_t)t * n + y) + 1];
            }
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rr = CP<S>::row(q, v);
                const S vr = sg * S(are[b][v]), vi = sg * S(aim[b][v]);
                if (last) {
                    const int r = r0 + rr;
                    if (r < n && y < n) {
                        S* o = Tt + 2 * ((((size_t)f * nc + j) * n + r) * n + y);
                        o[0] = vr;
                        o[1] = vi;
                    }
                } else {
                    // the next step's A: P_t . v (zero beyond column n)
                    sre[rr * ld + y] = pr * vr - pi * vi;
                    sim[rr * ld + y] = pr * vi + pi * vr;
                }
            }
        }
        __syncthreads();
    }
}

// FP64 chunk products by 8-row blocks (grid (ceil(n / 8), nc, nf), 2
// workgroups per CU): the re and im planes of the block's 8 rows are stacked
// as the 16 rows of the MFMA's A operand, so one product serves both planes
// and the grid has twice the blocks of k_chunkprod (1152 at n = 288, 2 chunks,
// 16 fields: the last round of workgroups is a smaller share of the launch).
// In the f64 16x16x4 output layout lane (q, r16) holds rows 4v + q: v = 0, 1
// are the real parts of rows q and 4 + q, v = 2, 3 their imaginary parts, so
// the phase of the next step is applied without a lane exchange.  Column
// blocks are balanced over the 4 waves: wave w takes blocks w + 4b for
// b < nblk / 4, and the nblk % 4 remaining blocks are split by K into
// quarters, wave w taking quarter w of each; the quarters' partial sums meet
// in LDS (added in wave order).
// NF = nblk / 4 and RB = nblk % 4 are template arguments, so every loop
// over blocks is static.
constexpr int R8 = 8;
template <int NF, int RB>
__global__ void __launch_bounds__(CPT, 2) k_chunkprod8(int n, int nt, int c, const double* __restrict__ E,
                                                        long long ldE, const double* __restrict__ P,
                                                        double* __restrict__ Tt) {
    using Acc = d4;
    constexpr int ld = 64 * NF + 16 * RB + LPAD;
    constexpr int RBS = RB > 0 ? RB : 1;
    __shared__ __align__(16) double rows_lds[2][R8 * ld];
    __shared__ __align__(16) Acc red[4][RBS][64];
    const int np = (n + 15) & ~15;
    double* sre = rows_lds[0];
    double* sim = rows_lds[1];
    const int r0 = blockIdx.x * R8, j = blockIdx.y, f = blockIdx.z;
    const int nc = gridDim.y;
    const int t0 = j * c;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, r16 = lane & 15;
    constexpr int full = NF, rem = RB;
    const int nk = np / 16;                                         // 16-wide K chunks
    const int klo = wave * nk / 4, khi = (wave + 1) * nk / 4;       // this wave's quarter
    const double* Pf = P + 2 * (size_t)f * nt * n;

    // v = e_{r0 + r} with the first step's phase P_{t0 - 1} applied
    for (int e = tid; e < R8 * ld; e += CPT) {
        const int r = e / ld, x = e % ld;
        double vr = 0.0, vi = 0.0;
        if (x == r0 + r && x < n) {
            const int tp = (t0 - 1 + nt) % nt;
            vr = Pf[2 * ((size_t)tp * n + x)];
            vi = Pf[2 * ((size_t)tp * n + x) + 1];
        }
        sre[e] = vr;
        sim[e] = vi;
    }
    __syncthreads();

    // E rows of this lane: full block b covers columns 16 (wave + 4b) + r16
    // (64 b rows past e0), remaining block r covers 16 (4 full + r) + r16
    // (16 r rows past e1)
    const double* e0 = E + (size_t)(16 * wave + r16) * ldE + 4 * q;
    const double* e1 = E + (size_t)(16 * 4 * full + r16) * ldE + 4 * q;
    const long long sb = 64 * ldE, sr = 16 * ldE;
    // A operand: MFMA row r16 = real part of row r16 (< 8) or imaginary part of row r16 - 8
    const double* arow = (r16 < R8 ? sre + r16 * ld : sim + (r16 - R8) * ld) + 4 * q;

    for (int s = 0; s < c; ++s) {
        const int t = t0 + s;
        const bool last = s == c - 1;
        // this step's phases P_t of the lane's output columns, loaded now so
        // that the epilogue does not wait for them
        double pph[NF + 1][2];
        #pragma unroll
        for (int b = 0; b <= NF; ++b) {
            const int y = b < NF ? 16 * (wave + 4 * b) + r16 : 16 * (4 * full + wave) + r16;
            const bool on = !last && y < n && (b < NF || wave < rem);
            pph[b][0] = on ? Pf[2 * ((size_t)t * n + y)] : 0.0;
            pph[b][1] = on ? Pf[2 * ((size_t)t * n + y) + 1] : 0.0;
        }
        Acc acc[NF > 0 ? NF : 1], accr[RBS];
        #pragma unroll
        for (int b = 0; b < NF; ++b) acc[b] = Acc{0, 0, 0, 0};
        #pragma unroll
        for (int r = 0; r < RBS; ++r) accr[r] = Acc{0, 0, 0, 0};
        if (s == 0) {
            // the first step maps e_i to column i of E times P_{t0-1}[i]: no
            // product needed (a remaining block's value goes to wave 0's part)
            auto first = [&](int y, Acc& a) {
                #pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rr = (4 * v + q) & 7;
                    const int i = r0 + rr;
                    if (i < n) a[v] = (v < 2 ? sre : sim)[rr * ld + i] * E[(size_t)y * ldE + i];
                }
            };
            #pragma unroll
            for (int b = 0; b < NF; ++b)
                if (b < full) first(16 * (wave + 4 * b) + r16, acc[b]);
            if (wave == 0) {
                #pragma unroll
                for (int r = 0; r < RB; ++r)
                    if (r < rem) first(16 * (4 * full + r) + r16, accr[r]);
            }
        } else {
            // full blocks over all of K; the next chunk's E and A are loaded
            // while the current chunk's MFMAs run (two register sets used in
            // turn: no copies between chunks, so the loads of one set are
            // never waited on by the products of the other)
            constexpr int NFS = NF > 0 ? NF : 1;
            // the remaining blocks' first chunk is loaded before the full
            // blocks' products, so its wait is hidden behind them
            const bool quarter = RB > 0 && klo < khi;
            double ra[RBS][4], rb2[RBS][4], qa[4], qb[4];
            auto loadr = [&](double (&e)[RBS][4], double (&x)[4], int kc) {
                #pragma unroll
                for (int r = 0; r < RB; ++r) ld4<double>(e1 + r * sr + 16 * kc, e[r]);
                ld4<double>(arow + 16 * kc, x);
            };
            auto prodr = [&](const double (&e)[RBS][4], const double (&x)[4]) {
                #pragma unroll
                for (int ks = 0; ks < 4; ++ks)
                    #pragma unroll
                    for (int r = 0; r < RB; ++r) accr[r] = CP<double>::mfma(x[ks], e[r][ks], accr[r]);
            };
            if (quarter) {
                #pragma unroll
                for (int r = 0; r < RB; ++r) ld4<double>(e1 + r * sr + 16 * klo, ra[r]);
            }
            double ea[NFS][4], eb[NFS][4], xa[4], xb[4];
            auto load = [&](double (&e)[NFS][4], double (&x)[4], int kc) {
                #pragma unroll
                for (int b = 0; b < NF; ++b) ld4<double>(e0 + b * sb + 16 * kc, e[b]);
                ld4<double>(arow + 16 * kc, x);
            };
            auto prod = [&](const double (&e)[NFS][4], const double (&x)[4]) {
                #pragma unroll
                for (int ks = 0; ks < 4; ++ks)
                    #pragma unroll
                    for (int b = 0; b < NF; ++b) acc[b] = CP<double>::mfma(x[ks], e[b][ks], acc[b]);
            };
            if (NF > 0) {
                // chunk pairs with unconditional loads (the last pair loads
                // chunk nk - 1 again), so the wait counts are exact on the
                // loop's single path
                load(ea, xa, 0);
                #pragma unroll 1
                for (int pp = 0; pp < nk / 2; ++pp) {
                    load(eb, xb, 2 * pp + 1);
                    prod(ea, xa);
                    load(ea, xa, min(2 * pp + 2, nk - 1));
                    prod(eb, xb);
                }
                if (nk & 1) prod(ea, xa);
            }
            // remaining blocks over this wave's quarter of K, in the same
            // alternation of two register sets
            if (quarter) {
                const int nq = khi - klo;
                ld4<double>(arow + 16 * klo, qa);
                #pragma unroll 1
                for (int pp = 0; pp < nq / 2; ++pp) {
                    loadr(rb2, qb, klo + 2 * pp + 1);
                    prodr(ra, qa);
                    loadr(ra, qa, min(klo + 2 * pp + 2, khi - 1));
                    prodr(rb2, qb);
                }
                if (nq & 1) prodr(ra, qa);
            }
        }
        // the remaining blocks' quarter sums; every wave is done reading this
        // step's rows after the barrier
        if (rem > 0) {
            #pragma unroll
            for (int r = 0; r < RB; ++r)
                if (r < rem) red[wave][r][lane] = accr[r];
        }
        __syncthreads();
        Acc accw = Acc{0, 0, 0, 0};
        if (wave < rem) {
            #pragma unroll
            for (int w2 = 0; w2 < 4; ++w2) accw += red[w2][wave][lane];
        }
        const double sg = t == 0 ? -1.0 : 1.0;
        auto out = [&](int y, const Acc& a, const double (&ph)[2]) {
            const double pr = ph[0], pi = ph[1];
            #pragma unroll
            for (int v = 0; v < 2; ++v) {
                const int rr = 4 * v + q;
                const double vr = sg * a[v], vi = sg * a[v + 2];
                if (last) {
                    const int r = r0 + rr;
                    if (r < n && y < n) {
                        double* o = Tt + 2 * ((((size_t)f * nc + j) * n + r) * n + y);
                        o[0] = vr;
                        o[1] = vi;
                    }
                } else {
                    // the next step's A: P_t . v (zero beyond column n)
                    sre[rr * ld + y] = pr * vr - pi * vi;
                    sim[rr * ld + y] = pr * vi + pi * vr;
                }
            }
        };
        #pragma unroll
        for (int b = 0; b < NF; ++b) out(16 * (wave + 4 * b) + r16, acc[b], pph[b]);
        if (wave < rem) out(16 * (4 * full + wave) + r16, accw, pph[NF]);
        __syncthreads();
    }
}

template <int NF, int RB>
void launch_k8(int n, int nt, int c, int nc, int nf, const void* E, long long ldE, const void* P, void* Tt,
               hipStream_t s) {
    const dim3 grid((n + R8 - 1) / R8, nc, nf);
    launch_timed(k_chunkprod8<NF, RB>, grid, dim3(CPT), 0, s, n, nt, c, static_cast<const double*>(E), ldE,
                 static_cast<const double*>(P), static_cast<double*>(Tt));
}

template <int NF>
void launch_k8r(int rb, int n, int nt, int c, int nc, int nf, const void* E, long long ldE, const void* P, void* Tt,
                hipStream_t s) {
    switch (rb) {
        case 0: launch_k8<NF, 0>(n, nt, c, nc, nf, E, ldE, P, Tt, s); break;
        case 1: launch_k8<NF, 1>(n, nt, c, nc, nf, E, ldE, P, Tt, s); break;
        case 2: launch_k8<NF, 2>(n, nt, c, nc, nf, E, ldE, P, Tt, s); break;
        default: launch_k8<NF, 3>(n, nt, c, nc, nf, E, ldE, P, Tt, s); break;
    }
}

template <class S, int MAXB>
void launch_k(int n, int nt, int c, int nc, int nf, const void* E, long long ldE, const void* P, void* Tt,
              hipStream_t s) {
    const dim3 grid((n + CPR - 1) / CPR, nc, nf);
    launch_timed(k_chunkprod<S, MAXB>, grid, dim3(CPT), 0, s, n, nt, c, static_cast<const S*>(E),
                 ldE, static_cast<const S*>(P), static_cast<S*>(Tt));
}

template <class S>
void launch_s(int n, int nt, int c, int nc, int nf, const void* E, long long ldE, const void* P, void* Tt,
              hipStream_t s) {
    if constexpr (sizeof(S) == 8) {
        static const bool rows8 = [] {
            const char* env = std::getenv("NSLX_CHUNKPROD8");
            return !(env && std::atoi(env) == 0);
        }();
        // the 8-row kernel where its registers fit two workgroups per CU
        // without spilling: up to 18 column blocks (n <= 288)
        const int nblk = (n + 15) / 16, full = nblk / 4, rb = nblk % 4;   // full column blocks per wave
        if (rows8 && nblk <= 18) {
            switch (full) {
                case 0: launch_k8r<0>(rb, n, nt, c, nc, nf, E, ldE, P, Tt, s); break;
                case 1: launch_k8r<1>(rb, n, nt, c, nc, nf, E, ldE, P, Tt, s); break;
                case 2: launch_k8r<2>(rb, n, nt, c, nc, nf, E, ldE, P, Tt, s); break;
                case 3: launch_k8r<3>(rb, n, nt, c, nc, nf, E, ldE, P, Tt, s); break;
                default:
                    if (rb == 0) launch_k8<4, 0>(n, nt, c, nc, nf, E, ldE, P, Tt, s);
                    else if (rb == 1) launch_k8<4, 1>(n, nt, c, nc, nf, E, ldE, P, Tt, s);
                    else launch_k8<4, 2>(n, nt, c, nc, nf, E, ldE, P, Tt, s);
                    break;
            }
            return;
        }
    }
    const int need = ((n + 15) / 16 + 3) / 4;   // column blocks per wave
    if (need <= 1) launch_k<S, 1>(n, nt, c, nc, nf, E, ldE, P, Tt, s);
    else if (need <= 2) launch_k<S, 2>(n, nt, c, nc, nf, E, ldE, P, Tt, s);
    else launch_k<S, 5>(n, nt, c, nc, nf, E, ldE, P, Tt, s);
}

}  // namespace

// At most 5 column blocks per wave (20 blocks, roundup(n, 16) <= 320): the
// accumulators and the prefetched E values of more would spill.  Beyond that
// the per-step batched GEMM builds the products.  NSLX_CHUNKPROD=0 forces it.
bool chunkprod_supported(int prec, int n) {
    const char* env = std::getenv("NSLX_CHUNKPROD");
    if (n < 1 || (env && std::atoi(env) == 0)) return false;
    if (prec != NSLX_F64 && prec != NSLX_F32) return false;
    return (n + 15) / 16 <= 20;
}

void launch_chunkprod(int prec, int n, int nt, int c, int nc, int nf, const void* E, long long ldE, const void* P,
                      void* Tt, hipStream_t s) {
    NSLX_REQUIRE(chunkprod_supported(prec, n), "chunk products: lattice too large for the row-block kernel");
    NSLX_REQUIRE(ldE >= ((n + 15) & ~15), "chunk products: E must be padded to a multiple of 16 columns");
    NSLX_REQUIRE(c >= 1 && nc >= 1 && nf >= 1 && nc * c == nt, "chunk products: bad chunking");
    ProfScope prof(prof_on() ? std::string(prec == NSLX_F64 ? "chunkprod_f64" : "chunkprod_f32") : std::string(), s);
    const double ce = prec == NSLX_F64 ? 16.0 : 8.0;
    // algorithmic work: c complex x real products of n x n by n x n per chain;
    // bytes: the products written once plus E and the phase rows read once
    prof.work(4.0 * n * (double)n * n * c * nc * nf,
              ce * (double)n * n * nc * nf + ce / 2 * (double)n * n + ce * (double)nt * n * nf);
    if (prec == NSLX_F64) launch_s<double>(n, nt, c, nc, nf, E, ldE, P, Tt, s);
    else launch_s<float>(n, nt, c, nc, nf, E, ldE, P, Tt, s);
    NSLX_LAUNCH_CHECK();
}

}  // namespace nslx
