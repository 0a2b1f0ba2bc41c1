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
