// Fused hop GEMM kernels for gfx950 (CDNA4).  See hop.hpp for the algebra.
//
// Work decomposition: a workgroup owns a BM x BN output tile that lies in ONE
// right-hand side b (rows t0..t0+BM of that RHS's nt time slices, sites
// y0..y0+BN).  Per-tile partial sums therefore belong to exactly one RHS and
// are reduced in a fixed order, so every reduction is deterministic and
// independent of what else is in the batch.  The nct column tiles of one row
// block are placed on the same XCD (blocks are dispatched round-robin over the
// 8 XCDs), so the field rows they all read are shared through one L2.
//
// Inner product on matrix cores: the K loop runs over x in chunks of BK,
// staging the field rows (phase-multiplied, split into re/im planes) and the
// E rows in LDS.  Within a chunk each lane feeds the MFMA a contiguous run of
// k values (k permuted consistently for A and B, which leaves the sum
// unchanged), so LDS reads are wide and conflict-light.
//   F64:    v_mfma_f64_16x16x4_f64, re and im as two real GEMMs sharing E.
//   F32:    v_mfma_f32_16x16x4_f32.
//   BF16X3: v_mfma_f32_16x16x32_bf16 on hi/lo splits, hi*hi + hi*lo + lo*hi.
#include "hop.hpp"
#include "prof.hpp"

namespace nslx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

template <int WM_, int WN_, int FM_, int FN_, int BK_>
struct Tile {
    static constexpr int WM = WM_, WN = WN_, FM = FM_, FN = FN_, BK = BK_;
    static constexpr int BM = WM * FM * 16, BN = WN * FN * 16, NT = 64 * WM * WN;
};

struct PF64 {
    using S = double;      // field scalar
    using L = double;      // LDS / MFMA operand scalar
    using Acc = d4;
    static constexpr int KSTEP = 4, PAD = 2, NPA = 2, NPB = 1, VEC = 2;
};
struct PF32 {
    using S = float;
    using L = float;
    using Acc = f4;
    static constexpr int KSTEP = 4, PAD = 4, NPA = 2, NPB = 1, VEC = 4;
};
struct PB3 {
    using S = float;
    using L = uint16_t;    // bf16 bits
    using Acc = f4;
    static constexpr int KSTEP = 32, PAD = 8, NPA = 4, NPB = 2, VEC = 4;
};

__device__ inline uint16_t f2bf(float x) {   // round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

template <class S>
struct CS { S re, im; };

template <class S>
__device__ inline CS<S> ldc(const void* base, size_t idx) {
    if constexpr (sizeof(S) == 8) {
        const double2 v = reinterpret_cast<const double2*>(base)[idx];
        return {v.x, v.y};
    } else {
        const float2 v = reinterpret_cast<const float2*>(base)[idx];
        return {v.x, v.y};
    }
}
template <class S>
__device__ inline void stc(void* base, size_t idx, CS<S> v) {
    if constexpr (sizeof(S) == 8) {
        reinterpret_cast<double2*>(base)[idx] = make_double2(v.re, v.im);
    } else {
        reinterpret_cast<float2*>(base)[idx] = make_float2(v.re, v.im);
    }
}

template <int NT>
__device__ inline double block_sum(double v, double* red) {
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    #pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += red[i];
    __syncthreads();
    return s;
}

template <int NT>
__device__ inline int block_sum_int(int v, int* red) {
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    int s = 0;
    #pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += red[i];
    __syncthreads();
    return s;
}

// Last-block finalisation of the fused CG scalars (fixed summation order).
template <int NT>
__device__ void cg_finalize(const HopArgs& a, int mode, int nparts, double* red, int* redi) {
    const CgDev& c = a.cg;
    if (!c.global) {
        int active = 0;
        for (int b = threadIdx.x; b < a.nb; b += NT) {
            if (mode != MODE_CGI && c.done[b] >= 0) continue;
            double s = 0.0;
            for (int k = 0; k < nparts; ++k) s += c.part[(size_t)b * nparts + k];
            if (mode == MODE_CGA) {
                c.alpha[b] = (s > 0.0) ? c.rr[b] / s : 0.0;
            } else if (mode == MODE_CGB) {
                const double r0 = c.rr[b];
                c.beta[b] = (r0 > 0.0) ? s / r0 : 0.0;
                c.rr[b] = s;
                if (s <= c.tol2[b]) c.done[b] = a.iter; else ++active;
            } else {
                c.rr[b] = s;
                c.beta[b] = 0.0;
                if (s <= c.tol2[b]) c.done[b] = 0; else { c.done[b] = -1; ++active; }
            }
        }
        if (mode != MODE_CGA) {
            const int tot = block_sum_int<NT>(active, redi);
            if (threadIdx.x == 0) { c.status[0] = tot; c.status[1] = a.iter; }
        }
    } else {
        double s = 0.0;
        for (int b = threadIdx.x; b < a.nb; b += NT)
            for (int k = 0; k < nparts; ++k) s += c.part[(size_t)b * nparts + k];
        const double S = block_sum<NT>(s, red);
        if (threadIdx.x == 0) {
            if (mode == MODE_CGA) {
                c.alpha[0] = (S > 0.0) ? c.rr[0] / S : 0.0;
            } else if (mode == MODE_CGB) {
                const double r0 = c.rr[0];
                c.beta[0] = (r0 > 0.0) ? S / r0 : 0.0;
                c.rr[0] = S;
                const bool conv = S <= c.tol2[0];
                if (conv) c.done[0] = a.iter;
                c.status[0] = conv ? 0 : a.nb;
                c.status[1] = a.iter;
            } else {
                c.rr[0] = S;
                c.beta[0] = 0.0;
                const bool conv = S <= c.tol2[0];
                c.done[0] = conv ? 0 : -1;
                c.status[0] = conv ? 0 : a.nb;
                c.status[1] = a.iter;
            }
        }
    }
}

template <class TL, class PR, int OP, int MODE>
__global__ void __launch_bounds__(TL::NT) hop_kernel(HopArgs a, int nrt, int nct, int swz, int ldE, size_t planeE) {
    using S = typename PR::S;
    using L = typename PR::L;
    using Acc = typename PR::Acc;
    constexpr int BM = TL::BM, BN = TL::BN, BK = TL::BK, NT = TL::NT;
    constexpr int FM = TL::FM, FN = TL::FN;
    constexpr int LDK = BK + PR::PAD;
    constexpr int RUN = BK / 4;                       // k values per lane per chunk
    static_assert(BK % 4 == 0 && RUN % (PR::KSTEP / 4) == 0, "BK");

    __shared__ __attribute__((aligned(16))) L sA[PR::NPA][BM][LDK];
    __shared__ __attribute__((aligned(16))) L sB[PR::NPB][BN][LDK];
    __shared__ double red[NT / 64];
    __shared__ int redi[NT / 64];
    __shared__ int amLast;

    // ---- tile decode ----
    const int G = a.nb * nrt;
    int g, ct;
    if (swz) {
        const int xcd = blockIdx.x & 7, s = blockIdx.x >> 3;
        ct = s % nct;
        g = (s / nct) * 8 + xcd;
    } else {
        ct = blockIdx.x % nct;
        g = blockIdx.x / nct;
    }
    const bool vblk = g < G;
    const int b = vblk ? g / nrt : 0;
    const int rt = vblk ? g % nrt : 0;
    const int nt = a.nt, nx = a.nx;
    const int t0 = rt * BM, y0 = ct * BN;
    const int f = b / a.rpf;
    const int nparts = nrt * nct;

    bool skip = !vblk;
    double alpha = 0.0, beta = 0.0;
    if (MODE == MODE_CGA || MODE == MODE_CGB) {
        const int slot = a.cg.global ? 0 : b;
        if (vblk && a.cg.done[slot] >= 0) skip = true;
        if (MODE == MODE_CGA) beta = a.cg.beta[slot];
        else alpha = a.cg.alpha[slot];
    }

    double part = 0.0;
    if (!skip) {
        const int tid = threadIdx.x;
        const int lane = tid & 63, wave = tid >> 6;
        const int wm = wave / TL::WN, wn = wave % TL::WN;
        const int r16 = lane & 15, q = lane >> 4;

        Acc accR[FM][FN], accI[FM][FN];
        #pragma unroll
        for (int i = 0; i < FM; ++i)
            #pragma unroll
            for (int j = 0; j < FN; ++j) { accR[i][j] = Acc{0, 0, 0, 0}; accI[i][j] = Acc{0, 0, 0, 0}; }

        const S* Pf = reinterpret_cast<const S*>(a.P) + (size_t)f * nt * nx * 2;
        const size_t bbase = (size_t)b * nt * nx;

        // A loader geometry
        constexpr int VEC = PR::VEC;
        constexpr int TPR = BK / VEC;
        constexpr int RPP = NT / TPR;
        constexpr int APASS = (BM + RPP - 1) / RPP;
        // B loader geometry: 16-byte vectors of L
        constexpr int VB = 16 / sizeof(L);
        constexpr int TPRB = BK / VB;
        constexpr int RPPB = NT / TPRB;
        constexpr int BPASS = (BN + RPPB - 1) / RPPB;

        for (int k0 = 0; k0 < nx; k0 += BK) {
            // ---- stage A: rows of the (shifted, phased) input field ----
            #pragma unroll
            for (int ps = 0; ps < APASS; ++ps) {
                const int i = ps * RPP + tid / TPR;
                const int kk = (tid % TPR) * VEC;
                if (i < BM) {
                    const int t = t0 + i;
                    const bool vrow = t < nt;
                    const int tin = (OP == HOP_M) ? (t == 0 ? nt - 1 : t - 1) : (t + 1 == nt ? 0 : t + 1);
                    const size_t rowoff = (bbase + (size_t)tin * nx);
                    CS<S> v[VEC];
                    #pragma unroll
                    for (int e = 0; e < VEC; ++e) {
                        const int x = k0 + kk + e;
                        CS<S> z{S(0), S(0)};
                        if (vrow && x < nx) {
                            z = ldc<S>(a.X, rowoff + x);
                            if (MODE == MODE_CGA) {
                                const CS<S> po = ldc<S>(a.Pold, rowoff + x);
                                z.re = z.re + S(beta) * po.re;
                                z.im = z.im + S(beta) * po.im;
                                if (ct == 0) stc<S>(a.Pnew, rowoff + x, z);
                            }
                            if (OP == HOP_M) {
                                const CS<S> ph = ldc<S>(Pf, (size_t)tin * nx + x);
                                const S re = ph.re * z.re - ph.im * z.im;
                                const S im = ph.re * z.im + ph.im * z.re;
                                z.re = re; z.im = im;
                            }
                        }
                        v[e] = z;
                    }
                    if constexpr (PR::NPA == 2) {
                        #pragma unroll
                        for (int e = 0; e < VEC; ++e) {
                            sA[0][i][kk + e] = v[e].re;
                            sA[1][i][kk + e] = v[e].im;
                        }
                    } else {
                        #pragma unroll
                        for (int e = 0; e < VEC; ++e) {
                            const uint16_t rh = f2bf(v[e].re), ih = f2bf(v[e].im);
                            sA[0][i][kk + e] = rh;
                            sA[1][i][kk + e] = f2bf(v[e].re - bf2f(rh));
                            sA[2][i][kk + e] = ih;
                            sA[3][i][kk + e] = f2bf(v[e].im - bf2f(ih));
                        }
                    }
                }
            }
            // ---- stage B: rows y of E (zero padded to ldE columns / rows) ----
            #pragma unroll
            for (int ps = 0; ps < BPASS; ++ps) {
                const int j = ps * RPPB + tid / TPRB;
                const int kk = (tid % TPRB) * VB;
                if (j < BN) {
                    const size_t eoff = (size_t)(y0 + j) * ldE + k0 + kk;
                    #pragma unroll
                    for (int pl = 0; pl < PR::NPB; ++pl) {
                        const uint4 w = *reinterpret_cast<const uint4*>(
                            reinterpret_cast<const L*>(a.E) + (size_t)pl * planeE + eoff);
                        *reinterpret_cast<uint4*>(&sB[pl][j][kk]) = w;
                    }
                }
            }
            __syncthreads();

            // ---- MFMA over the chunk ----
            if constexpr (PR::KSTEP == 4) {
                #pragma unroll
                for (int s = 0; s < RUN; ++s) {
                    const int k = q * RUN + s;
                    L ar[FM], ai[FM], bv[FN];
                    #pragma unroll
                    for (int i = 0; i < FM; ++i) {
                        const int row = (wm * FM + i) * 16 + r16;
                        ar[i] = sA[0][row][k];
                        ai[i] = sA[1][row][k];
                    }
                    #pragma unroll
                    for (int j = 0; j < FN; ++j) bv[j] = sB[0][(wn * FN + j) * 16 + r16][k];
                    #pragma unroll
                    for (int i = 0; i < FM; ++i)
                        #pragma unroll
                        for (int j = 0; j < FN; ++j) {
                            if constexpr (sizeof(L) == 8) {
                                accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bv[j], accR[i][j], 0, 0, 0);
                                accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bv[j], accI[i][j], 0, 0, 0);
                            } else {
                                accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], bv[j], accR[i][j], 0, 0, 0);
                                accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[i], bv[j], accI[i][j], 0, 0, 0);
                            }
                        }
                }
            } else {
                #pragma unroll
                for (int s = 0; s < RUN / 8; ++s) {
                    const int k = q * RUN + 8 * s;
                    bf8 arh[FM], arl[FM], aih[FM], ail[FM], bh[FN], bl[FN];
                    #pragma unroll
                    for (int i = 0; i < FM; ++i) {
                        const int row = (wm * FM + i) * 16 + r16;
                        arh[i] = *reinterpret_cast<const bf8*>(&sA[0][row][k]);
                        arl[i] = *reinterpret_cast<const bf8*>(&sA[1][row][k]);
                        aih[i] = *reinterpret_cast<const bf8*>(&sA[2][row][k]);
                        ail[i] = *reinterpret_cast<const bf8*>(&sA[3][row][k]);
                    }
                    #pragma unroll
                    for (int j = 0; j < FN; ++j) {
                        const int col = (wn * FN + j) * 16 + r16;
                        bh[j] = *reinterpret_cast<const bf8*>(&sB[0][col][k]);
                        bl[j] = *reinterpret_cast<const bf8*>(&sB[1][col][k]);
                    }
                    #pragma unroll
                    for (int i = 0; i < FM; ++i)
                        #pragma unroll
                        for (int j = 0; j < FN; ++j) {
                            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arl[i], bh[j], accR[i][j], 0, 0, 0);
                            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh[i], bl[j], accR[i][j], 0, 0, 0);
                            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh[i], bh[j], accR[i][j], 0, 0, 0);
                            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ail[i], bh[j], accI[i][j], 0, 0, 0);
                            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih[i], bl[j], accI[i][j], 0, 0, 0);
                            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih[i], bh[j], accI[i][j], 0, 0, 0);
                        }
                }
            }
            __syncthreads();
        }

        // ---- epilogue: shift sign, identity, conj(P), fused recurrences ----
        #pragma unroll
        for (int i = 0; i < FM; ++i) {
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int t = t0 + (wm * FM + i) * 16 + 4 * q + v;
                if (t >= nt) continue;
                const size_t rowoff = bbase + (size_t)t * nx;
                #pragma unroll
                for (int j = 0; j < FN; ++j) {
                    const int y = y0 + (wn * FN + j) * 16 + r16;
                    if (y >= nx) continue;
                    const size_t idx = rowoff + y;
                    S fr = S(accR[i][j][v]), fi = S(accI[i][j][v]);
                    if (OP == HOP_M) {
                        if (t == 0) { fr = -fr; fi = -fi; }
                    } else {
                        const CS<S> ph = ldc<S>(Pf, (size_t)t * nx + y);   // conj(P_t) * acc
                        const S re = ph.re * fr + ph.im * fi;
                        const S im = ph.re * fi - ph.im * fr;
                        fr = re; fi = im;
                        if (t == nt - 1) { fr = -fr; fi = -fi; }
                    }
                    CS<S> id;
                    if (MODE == MODE_CGA) {
                        const CS<S> rv = ldc<S>(a.X, idx);
                        const CS<S> po = ldc<S>(a.Pold, idx);
                        id.re = rv.re + S(beta) * po.re;
                        id.im = rv.im + S(beta) * po.im;
                    } else {
                        id = ldc<S>(a.X, idx);
                    }
                    const CS<S> out{id.re - fr, id.im - fi};
                    if (MODE == MODE_PLAIN) {
                        stc<S>(a.Y, idx, out);
                    } else if (MODE == MODE_CGA) {
                        stc<S>(a.Y, idx, out);
                        part += double(out.re) * out.re + double(out.im) * out.im;
                    } else if (MODE == MODE_CGB) {
                        const CS<S> pv = ldc<S>(a.Pold, idx);
                        CS<S> xv = ldc<S>(a.Xs, idx);
                        CS<S> rv = ldc<S>(a.R, idx);
                        xv.re += S(alpha) * pv.re; xv.im += S(alpha) * pv.im;
                        rv.re -= S(alpha) * out.re; rv.im -= S(alpha) * out.im;
                        stc<S>(a.Xs, idx, xv);
                        stc<S>(a.R, idx, rv);
                        part += double(rv.re) * rv.re + double(rv.im) * rv.im;
                    } else {  // CGI
                        const CS<S> bv = ldc<S>(a.Bv, idx);
                        const CS<S> rv{bv.re - out.re, bv.im - out.im};
                        stc<S>(a.R, idx, rv);
                        stc<S>(a.Pnew, idx, rv);
                        part += double(rv.re) * rv.re + double(rv.im) * rv.im;
                    }
                }
            }
        }
    }

    if (MODE == MODE_PLAIN) return;

    // ---- partial sum of this tile, then last-block finalisation ----
    const double tot = block_sum<NT>(part, red);
    if (threadIdx.x == 0 && vblk) a.cg.part[(size_t)b * nparts + rt * nct + ct] = tot;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned prev = atomicAdd(a.cg.counter, 1u);
        amLast = (prev == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (amLast) {
        __threadfence();
        cg_finalize<NT>(a, MODE, nparts, red, redi);
        if (threadIdx.x == 0) *a.cg.counter = 0u;
    }
}

// ---------------------------------------------------------------- tile menus
using T64_a = Tile<2, 2, 2, 3, 16>;   // 64 x 96
using T64_b = Tile<2, 2, 1, 3, 16>;   // 32 x 96
using T64_c = Tile<1, 2, 1, 1, 16>;   // 16 x 32 (small lattices)
using T64_d = Tile<2, 2, 2, 2, 16>;   // 64 x 64
using T32_a = Tile<2, 2, 2, 3, 32>;
using T32_b = Tile<2, 2, 1, 3, 32>;
using T32_c = Tile<1, 2, 1, 1, 32>;
using T32_d = Tile<2, 2, 2, 2, 32>;
using TB3_a = Tile<2, 2, 2, 3, 64>;
using TB3_b = Tile<2, 2, 1, 3, 64>;
using TB3_c = Tile<1, 2, 1, 1, 64>;
using TB3_d = Tile<2, 2, 2, 2, 64>;

template <class TL>
TileInfo info(int id) { return {id, TL::BM, TL::BN, TL::BK, TL::NT}; }

template <class TL, class PR>
void launch_t(int op, int mode, const HopArgs& a, hipStream_t s) {
    const int nrt = (a.nt + TL::BM - 1) / TL::BM;
    const int nct = (a.nx + TL::BN - 1) / TL::BN;
    const int G = a.nb * nrt;
    const int swz = G >= 8 ? 1 : 0;
    const int grid = swz ? ((G + 7) / 8) * 8 * nct : G * nct;
    const int ldE = e_ld(a.nx);
    const size_t planeE = (size_t)e_rows(a.nx) * ldE;
    dim3 gr(grid), bl(TL::NT);
#define NSLX_HOP_CASE(OPV, MV)                                                                  \
    if (op == OPV && mode == MV) {                                                              \
        hipLaunchKernelGGL((hop_kernel<TL, PR, OPV, MV>), gr, bl, 0, s, a, nrt, nct, swz, ldE, planeE); \
        return;                                                                                 \
    }
    NSLX_HOP_CASE(HOP_M, MODE_PLAIN)
    NSLX_HOP_CASE(HOP_MD, MODE_PLAIN)
    NSLX_HOP_CASE(HOP_M, MODE_CGA)
    NSLX_HOP_CASE(HOP_MD, MODE_CGA)
    NSLX_HOP_CASE(HOP_M, MODE_CGB)
    NSLX_HOP_CASE(HOP_MD, MODE_CGB)
    NSLX_HOP_CASE(HOP_M, MODE_CGI)
    NSLX_HOP_CASE(HOP_MD, MODE_CGI)
#undef NSLX_HOP_CASE
    throw Error(NSLX_E_ARG, "hop: bad op/mode");
}

template <class PR, class T0, class T1, class T2, class T3>
void dispatch4(int tile, int op, int mode, const HopArgs& a, hipStream_t s) {
    switch (tile) {
        case 0: launch_t<T0, PR>(op, mode, a, s); break;
        case 1: launch_t<T1, PR>(op, mode, a, s); break;
        case 2: launch_t<T2, PR>(op, mode, a, s); break;
        case 3: launch_t<T3, PR>(op, mode, a, s); break;
        default: throw Error(NSLX_E_ARG, "hop: bad tile id");
    }
}

}  // namespace

const std::vector<TileInfo>& tile_menu(int prec) {
    static const std::vector<TileInfo> m64 = {info<T64_a>(0), info<T64_b>(1), info<T64_c>(2), info<T64_d>(3)};
    static const std::vector<TileInfo> m32 = {info<T32_a>(0), info<T32_b>(1), info<T32_c>(2), info<T32_d>(3)};
    static const std::vector<TileInfo> mb3 = {info<TB3_a>(0), info<TB3_b>(1), info<TB3_c>(2), info<TB3_d>(3)};
    if (prec == NSLX_F64) return m64;
    if (prec == NSLX_F32) return m32;
    return mb3;
}

int hop_nparts(int prec, int tile, int nx, int nt) {
    const TileInfo& ti = tile_menu(prec).at(tile);
    return ((nt + ti.bm - 1) / ti.bm) * ((nx + ti.bn - 1) / ti.bn);
}

int default_tile(int prec, int nx, int nt, int64_t nb) {
    (void)prec;
    if (nx <= 32) return 2;
    const int64_t rows = nb * nt;
    // enough workgroups to cover 256 CUs at least twice with the big tile?
    const int64_t tiles_a = ((rows + 63) / 64) * ((nx + 95) / 96);
    if (tiles_a >= 512) return 0;
    return 1;
}

namespace {
__global__ void k_pack_E(int nx, int ld, int rows, const double* E, double* E64, float* E32, uint16_t* Ebf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * ld) return;
    const int y = (int)(i / ld), x = (int)(i % ld);
    const double v = (y < nx && x < nx) ? E[(size_t)y * nx + x] : 0.0;
    if (E64) E64[i] = v;
    if (E32) E32[i] = (float)v;
    if (Ebf) {
        const float f = (float)v;
        const uint16_t hi = f2bf(f);
        Ebf[i] = hi;
        Ebf[(size_t)rows * ld + i] = f2bf(f - bf2f(hi));
    }
}
}  // namespace

void pack_E(int nx, const double* E, void* E64, void* E32, void* Ebf, hipStream_t s) {
    const int ld = e_ld(nx), rows = e_rows(nx);
    const int64_t n = (int64_t)rows * ld;
    hipLaunchKernelGGL(k_pack_E, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nx, ld, rows, E, (double*)E64,
                       (float*)E32, (uint16_t*)Ebf);
    NSLX_LAUNCH_CHECK();
}

void launch_hop(int prec, int op, int mode, const HopArgs& a, int tile, hipStream_t s) {
    NSLX_REQUIRE(a.nx > 0 && a.nt > 0 && a.nb > 0, "hop: empty shape");
    ProfScope prof(op == HOP_M ? "hop_M" : "hop_Mdag", prec, mode, s);
    if (prec == NSLX_F64) dispatch4<PF64, T64_a, T64_b, T64_c, T64_d>(tile, op, mode, a, s);
    else if (prec == NSLX_F32) dispatch4<PF32, T32_a, T32_b, T32_c, T32_d>(tile, op, mode, a, s);
    else if (prec == NSLX_BF16X3) dispatch4<PB3, TB3_a, TB3_b, TB3_c, TB3_d>(tile, op, mode, a, s);
    else throw Error(NSLX_E_ARG, "hop: bad precision");
    NSLX_LAUNCH_CHECK();
}

}  // namespace nslx
