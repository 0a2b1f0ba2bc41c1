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

#include <algorithm>

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
    static constexpr bool ROW4V = true;
};
struct PF32 {
    using S = float;
    using L = float;
    using Acc = f4;
    static constexpr int KSTEP = 4, PAD = 4, NPA = 2, NPB = 1, VEC = 4;
    static constexpr bool ROW4V = false;
};
struct PB3 {
    using S = float;
    using L = uint16_t;    // bf16 bits
    using Acc = f4;
    static constexpr int KSTEP = 32, PAD = 8, NPA = 4, NPB = 2, VEC = 4;
    static constexpr bool ROW4V = false;
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

// Pointers of one tile job (the field block of one problem).
struct JobPtrs {
    const void* E; const void* P;
    const void* X; void* Y; const void* Pold; void* Pnew; void* Xs; void* R; const void* Bv;
};

template <class TL, class PR>
struct Smem {
    typename PR::L A[PR::NPA][TL::BM][TL::BK + PR::PAD];
    typename PR::L B[PR::NPB][TL::BN][TL::BK + PR::PAD];
};

// One BM x BN output tile of rhs b (rows t0 = rt*BM.., sites y0 = ct*BN..).
// Called by the whole workgroup; returns this thread's share of the tile's
// reduction (|u|^2 or |r|^2 in the CG modes).
template <class TL, class PR, int OP, int MODE>
__device__ double hop_tile(const JobPtrs& J, int nx, int nt, int rpf, int ldE, size_t planeE, int b, int rt, int ct,
                           double alpha, double beta, Smem<TL, PR>& sm) {
    using S = typename PR::S;
    using L = typename PR::L;
    using Acc = typename PR::Acc;
    constexpr int BM = TL::BM, BN = TL::BN, BK = TL::BK, NT = TL::NT;
    constexpr int FM = TL::FM, FN = TL::FN;
    constexpr int RUN = BK / 4;                       // k values per lane per chunk
    static_assert(BK % 4 == 0 && RUN % (PR::KSTEP / 4) == 0, "BK");
    auto& sA = sm.A;
    auto& sB = sm.B;
    const int t0 = rt * BM, y0 = ct * BN;
    const int f = b / rpf;
    double part = 0.0;
        const int tid = threadIdx.x;
        const int lane = tid & 63, wave = tid >> 6;
        const int wm = wave / TL::WN, wn = wave % TL::WN;
        const int r16 = lane & 15, q = lane >> 4;

        Acc accR[FM][FN], accI[FM][FN];
        #pragma unroll
        for (int i = 0; i < FM; ++i)
            #pragma unroll
            for (int j = 0; j < FN; ++j) { accR[i][j] = Acc{0, 0, 0, 0}; accI[i][j] = Acc{0, 0, 0, 0}; }

        const S* Pf = reinterpret_cast<const S*>(J.P) + (size_t)f * nt * nx * 2;
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
                            z = ldc<S>(J.X, rowoff + x);
                            if (MODE == MODE_CGA) {
                                const CS<S> po = ldc<S>(J.Pold, rowoff + x);
                                z.re = z.re + S(beta) * po.re;
                                z.im = z.im + S(beta) * po.im;
                                if (ct == 0) stc<S>(J.Pnew, rowoff + x, z);
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
                            reinterpret_cast<const L*>(J.E) + (size_t)pl * planeE + eoff);
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
                // accumulator register v of lane (q, r16) holds row 4q+v of the
                // 16x16 block for the f32/bf16 MFMAs, row 4v+q for f64 (its
                // output groups have size 1)
                const int t = t0 + (wm * FM + i) * 16 + (PR::ROW4V ? 4 * v + q : 4 * q + v);
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
                        const CS<S> rv = ldc<S>(J.X, idx);
                        const CS<S> po = ldc<S>(J.Pold, idx);
                        id.re = rv.re + S(beta) * po.re;
                        id.im = rv.im + S(beta) * po.im;
                    } else {
                        id = ldc<S>(J.X, idx);
                    }
                    const CS<S> out{id.re - fr, id.im - fi};
                    if (MODE == MODE_PLAIN) {
                        stc<S>(J.Y, idx, out);
                    } else if (MODE == MODE_CGA) {
                        stc<S>(J.Y, idx, out);
                        part += double(out.re) * out.re + double(out.im) * out.im;
                    } else if (MODE == MODE_CGB) {
                        const CS<S> pv = ldc<S>(J.Pold, idx);
                        CS<S> xv = ldc<S>(J.Xs, idx);
                        CS<S> rv = ldc<S>(J.R, idx);
                        xv.re += S(alpha) * pv.re; xv.im += S(alpha) * pv.im;
                        rv.re -= S(alpha) * out.re; rv.im -= S(alpha) * out.im;
                        stc<S>(J.Xs, idx, xv);
                        stc<S>(J.R, idx, rv);
                        part += double(rv.re) * rv.re + double(rv.im) * rv.im;
                    } else {  // CGI
                        const CS<S> bv = ldc<S>(J.Bv, idx);
                        const CS<S> rv{bv.re - out.re, bv.im - out.im};
                        stc<S>(J.R, idx, rv);
                        stc<S>(J.Pnew, idx, rv);
                        part += double(rv.re) * rv.re + double(rv.im) * rv.im;
                    }
                }
            }
        }
    
    return part;
}

// Sum of n consecutive tile partials (fixed order).
__device__ inline double sum_parts(const double* p, size_t off, size_t n) {
    double s = 0.0;
    for (size_t k = 0; k < n; ++k) s += p[off + k];
    return s;
}

// Block-wide fixed-order sum of n partials (all threads get the result).
template <int NT>
__device__ inline double block_sum_parts(const double* p, size_t n, double* red) {
    double s = 0.0;
    for (size_t k = threadIdx.x; k < n; k += NT) s += p[k];
    return block_sum<NT>(s, red);
}

// CG decisions shared by the single and the grouped kernels.  Returns true if
// the tile's right-hand side is skipped; sets beta (CGA) or alpha (CGB).
// rr1 = |r_{i-1}|^2, rr2 = |r_{i-2}|^2, uu = |u_i|^2 of the slot.
template <int MODE>
__device__ inline bool cg_decide(const CgDev& c, int slot, int iter, bool designated, double rr1, double rr2,
                                 double uu, double& alpha, double& beta) {
    if (c.done[slot] >= 0) return true;
    if (MODE == MODE_CGA) {
        if (c.ru) {
            if (designated) c.want[slot] = (rr1 <= c.delta2 * c.rr_ref[slot] || rr1 <= c.tol2[slot]) ? 1 : 0;
        } else if (rr1 <= c.tol2[slot]) {
            if (designated) c.done[slot] = iter - 1;
            return true;
        }
        beta = (iter >= 2 && rr2 > 0.0) ? rr1 / rr2 : 0.0;
    } else if (MODE == MODE_CGB) {
        alpha = (uu > 0.0) ? rr1 / uu : 0.0;
    }
    return false;
}

template <class TL, class PR, int OP, int MODE>
__global__ void __launch_bounds__(TL::NT) hop_kernel(HopArgs a, int nrt, int nct, int swz, int ldE, size_t planeE) {
    constexpr int NT = TL::NT;
    __shared__ __attribute__((aligned(16))) Smem<TL, PR> sm;
    __shared__ double red[NT / 64];

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
    if (!vblk) return;
    const int b = g / nrt;
    const int rt = g % nrt;
    const int nparts = nrt * nct;

    bool skip = false;
    double alpha = 0.0, beta = 0.0;
    if (MODE == MODE_CGA || MODE == MODE_CGB) {
        const CgDev& c = a.cg;
        const int it = a.iter;
        double rr1, rr2 = 0.0, uu = 0.0;
        if (c.global) {
            const size_t n = (size_t)a.nb * nparts;
            rr1 = block_sum_parts<NT>(c.rpart[(it - 1) & 1], n, red);
            if (MODE == MODE_CGA && it >= 2) rr2 = block_sum_parts<NT>(c.rpart[(it - 2) & 1], n, red);
            if (MODE == MODE_CGB) uu = block_sum_parts<NT>(c.upart, n, red);
        } else {
            const size_t off = (size_t)b * nparts;
            rr1 = sum_parts(c.rpart[(it - 1) & 1], off, nparts);
            if (MODE == MODE_CGA && it >= 2) rr2 = sum_parts(c.rpart[(it - 2) & 1], off, nparts);
            if (MODE == MODE_CGB) uu = sum_parts(c.upart, off, nparts);
        }
        const int slot = c.global ? 0 : b;
        const bool designated = rt == 0 && ct == 0 && (!c.global || b == 0);
        skip = cg_decide<MODE>(c, slot, it, designated, rr1, rr2, uu, alpha, beta);
    }
    double part = 0.0;
    if (!skip) {
        const JobPtrs J{a.E, a.P, a.X, a.Y, a.Pold, a.Pnew, a.Xs, a.R, a.Bv};
        part = hop_tile<TL, PR, OP, MODE>(J, a.nx, a.nt, a.rpf, ldE, planeE, b, rt, ct, alpha, beta, sm);
    }
    if (MODE == MODE_PLAIN || skip) return;
    // ---- this tile's partial sum, reduced by the next launch ----
    const double tot = block_sum<NT>(part, red);
    if (threadIdx.x == 0) {
        double* dst = MODE == MODE_CGA ? a.cg.upart : a.cg.rpart[MODE == MODE_CGB ? (a.iter & 1) : 0];
        dst[(size_t)b * nparts + rt * nct + ct] = tot;
    }
}

// Persistent grouped launch: workgroups stride over a device tile list that
// spans many problems (each with its own nx, nt, E, phases and vectors).
// Per-RHS decisions only (no batch-global mode).
template <class TL, class PR, int OP, int MODE>
__global__ void __launch_bounds__(TL::NT) hop_grouped(GroupArgs g) {
    constexpr int NT = TL::NT;
    __shared__ __attribute__((aligned(16))) Smem<TL, PR> sm;
    __shared__ double red[NT / 64];
    for (int t = blockIdx.x; t < g.ntiles; t += gridDim.x) {
        const int4 T = g.tiles[t];
        const HopProblem& P = g.probs[T.x];
        const int b = T.y, rt = T.z, ct = T.w;
        const int slot = P.rhs0 + b;
        const int nparts = P.nrt * P.nct;
        const size_t off = (size_t)P.part0 + (size_t)b * nparts;
        bool skip = false;
        double alpha = 0.0, beta = 0.0;
        if (MODE == MODE_CGA || MODE == MODE_CGB) {
            const CgDev& c = g.cg;
            const double rr1 = sum_parts(c.rpart[(g.iter - 1) & 1], off, nparts);
            const double rr2 = (MODE == MODE_CGA && g.iter >= 2) ? sum_parts(c.rpart[(g.iter - 2) & 1], off, nparts) : 0.0;
            const double uu = MODE == MODE_CGB ? sum_parts(c.upart, off, nparts) : 0.0;
            skip = cg_decide<MODE>(c, slot, g.iter, rt == 0 && ct == 0, rr1, rr2, uu, alpha, beta);
        }
        double part = 0.0;
        if (!skip) {
            auto v = [&](int r) -> void* { return r >= 0 ? P.v[r] : nullptr; };
            const JobPtrs J{P.E, P.P, v(g.role[0]), v(g.role[1]), v(g.role[2]), v(g.role[3]),
                            v(g.role[4]), v(g.role[5]), v(g.role[6])};
            part = hop_tile<TL, PR, OP, MODE>(J, P.nx, P.nt, P.rpf, P.ldE, (size_t)P.planeE, b, rt, ct, alpha,
                                              beta, sm);
        }
        if (MODE != MODE_PLAIN && !skip) {
            const double tot = block_sum<NT>(part, red);
            if (threadIdx.x == 0) {
                double* dst = MODE == MODE_CGA ? g.cg.upart : g.cg.rpart[MODE == MODE_CGB ? (g.iter & 1) : 0];
                dst[off + rt * P.nct + ct] = tot;
            }
        } else {
            __syncthreads();   // LDS reuse by the next tile
        }
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

template <class TL, class PR>
void launch_g(int op, int mode, const GroupArgs& g, int grid, hipStream_t s) {
    dim3 gr(grid), bl(TL::NT);
#define NSLX_HOPG_CASE(OPV, MV)                                                    \
    if (op == OPV && mode == MV) {                                                 \
        hipLaunchKernelGGL((hop_grouped<TL, PR, OPV, MV>), gr, bl, 0, s, g);        \
        return;                                                                    \
    }
    NSLX_HOPG_CASE(HOP_M, MODE_PLAIN)
    NSLX_HOPG_CASE(HOP_MD, MODE_PLAIN)
    NSLX_HOPG_CASE(HOP_M, MODE_CGA)
    NSLX_HOPG_CASE(HOP_MD, MODE_CGA)
    NSLX_HOPG_CASE(HOP_M, MODE_CGB)
    NSLX_HOPG_CASE(HOP_MD, MODE_CGB)
    NSLX_HOPG_CASE(HOP_M, MODE_CGI)
    NSLX_HOPG_CASE(HOP_MD, MODE_CGI)
#undef NSLX_HOPG_CASE
    throw Error(NSLX_E_ARG, "hop: bad op/mode");
}

template <class PR, class T0, class T1, class T2, class T3>
void dispatch4g(int tile, int op, int mode, const GroupArgs& g, int grid, hipStream_t s) {
    switch (tile) {
        case 0: launch_g<T0, PR>(op, mode, g, grid, s); break;
        case 1: launch_g<T1, PR>(op, mode, g, grid, s); break;
        case 2: launch_g<T2, PR>(op, mode, g, grid, s); break;
        case 3: launch_g<T3, PR>(op, mode, g, grid, s); break;
        default: throw Error(NSLX_E_ARG, "hop: bad tile id");
    }
}

}  // namespace

void launch_hop_grouped(int prec, int op, int mode, int tile, const GroupArgs& g, int grid, hipStream_t s) {
    NSLX_REQUIRE(g.nprob > 0 && g.ntiles > 0 && grid > 0, "hop: empty grouped launch");
    ProfScope prof(op == HOP_M ? "grp_M" : "grp_Mdag", prec, mode, s);
    if (prec == NSLX_F64) dispatch4g<PF64, T64_a, T64_b, T64_c, T64_d>(tile, op, mode, g, grid, s);
    else if (prec == NSLX_F32) dispatch4g<PF32, T32_a, T32_b, T32_c, T32_d>(tile, op, mode, g, grid, s);
    else if (prec == NSLX_BF16X3) dispatch4g<PB3, TB3_a, TB3_b, TB3_c, TB3_d>(tile, op, mode, g, grid, s);
    else throw Error(NSLX_E_ARG, "hop: bad precision");
    NSLX_LAUNCH_CHECK();
}

std::vector<int4> build_tile_list(std::vector<HopProblem>& probs, int prec, int tile) {
    const TileInfo& ti = tile_menu(prec).at(tile);
    std::vector<int> order(probs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return probs[a].nx > probs[b].nx; });
    // row groups in cost order, each group's nct column tiles
    std::vector<std::vector<int4>> groups;
    for (int p : order) {
        HopProblem& P = probs[p];
        P.nrt = (P.nt + ti.bm - 1) / ti.bm;
        P.nct = (P.nx + ti.bn - 1) / ti.bn;
        P.ldE = e_ld(P.nx);
        P.planeE = (long long)e_rows(P.nx) * P.ldE;
        for (int b = 0; b < P.nb; ++b)
            for (int rt = 0; rt < P.nrt; ++rt) {
                std::vector<int4> gtiles;
                for (int ct = 0; ct < P.nct; ++ct) gtiles.push_back(make_int4(p, b, rt, ct));
                groups.push_back(std::move(gtiles));
            }
    }
    // interleave 8 groups at a time: entry 8*(j*maxct + ct) + i for group i of a block of 8
    std::vector<int4> out;
    for (size_t g0 = 0; g0 < groups.size(); g0 += 8) {
        size_t maxct = 0;
        for (size_t i = g0; i < std::min(groups.size(), g0 + 8); ++i) maxct = std::max(maxct, groups[i].size());
        for (size_t ct = 0; ct < maxct; ++ct)
            for (size_t i = g0; i < std::min(groups.size(), g0 + 8); ++i)
                if (ct < groups[i].size()) out.push_back(groups[i][ct]);
    }
    return out;
}

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
