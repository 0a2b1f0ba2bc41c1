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
#include <type_traits>

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

// 16 bytes of E in a native vector type: held in HIP's union-based uint4,
// the E stage of the M^dagger and HubbardDiag instantiations stayed in
// private memory (64-96 bytes of scratch per lane, one store and one load per
// K chunk); as a clang vector it is promoted to registers in every variant.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

struct PF64 {
    using S = double;      // field scalar
    using L = double;      // LDS / MFMA operand scalar
    using Acc = d4;
    static constexpr int KSTEP = 4, PAD = 2, NPA = 2, NPB = 1, VEC = 2, NBUF = 2;
    static constexpr bool ROW4V = true;
    // E chunk straight into LDS (global_load_lds_dwordx4, see load_b_direct):
    // off.  It removes the E register stage, but the unpadded 128-byte rows it
    // needs raise the LDS bank conflicts of the B reads from 2.7 to 3.8-4.4
    // cycles per LDS instruction and the fused CG kernels run ~7 % slower.
    static constexpr bool DLDS = false;
};
struct PF32 {
    using S = float;
    using L = float;
    using Acc = f4;
    static constexpr int KSTEP = 4, PAD = 4, NPA = 2, NPB = 1, VEC = 4, NBUF = 2;
    static constexpr bool ROW4V = false;
    static constexpr bool DLDS = false;   // as PF64
};
struct PB3 {
    using S = float;
    using L = uint16_t;    // bf16 bits
    using Acc = f4;
    // BK = 32 keeps two LDS buffers of the four A and two B bf16 planes at
    // 51 KB per workgroup (3 per CU); one k-step of the 16x16x32 MFMA per chunk
    static constexpr int KSTEP = 32, PAD = 8, NPA = 4, NPB = 2, VEC = 4, NBUF = 2;
    static constexpr bool ROW4V = false;
    static constexpr bool DLDS = false;   // 64-byte B rows; direct loads need 128
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

// 64 zero bytes: the source of every load that lies outside the field (a run
// is at most 64 bytes).  Not const: a const array would live in the constant
// address space, and a select between it and a global pointer would turn the
// loads into flat loads (whose waits cannot be counted per load).  Never
// written.
__device__ __attribute__((aligned(16))) double g_zero_run[8] = {0, 0, 0, 0, 0, 0, 0, 0};

// One complex element, or zero when !ok, without a branch on the data.
template <class S>
__device__ inline CS<S> ldc_or0(const void* base, size_t idx, bool ok) {
    if constexpr (sizeof(S) == 8) {
        const double2* p = ok ? reinterpret_cast<const double2*>(base) + idx
                              : reinterpret_cast<const double2*>(g_zero_run);
        const double2 v = *p;
        return {v.x, v.y};
    } else {
        const float2* p = ok ? reinterpret_cast<const float2*>(base) + idx
                             : reinterpret_cast<const float2*>(g_zero_run);
        const float2 v = *p;
        return {v.x, v.y};
    }
}

// LDS of a classic tile: NBUF copies of the A (field, split planes) and B (E)
// chunks; with two the next chunk is stored while the current one feeds the
// MFMAs (one barrier per chunk).  With DLDS the B rows (BK * sizeof(L) =
// 128 bytes) are unpadded and their 16-byte groups XOR-swizzled by row & 7
// (b_col below), the layout global_load_lds writes 8 rows at a time.
template <class TL, class PR>
struct Smem {
    typename PR::L A[PR::NBUF][PR::NPA][TL::BM][TL::BK + PR::PAD];
    typename PR::L B[PR::NBUF][PR::NPB][TL::BN][TL::BK + (PR::DLDS ? 0 : PR::PAD)];
};

// Column of element k of B row `row` in the LDS layout.
template <class PR>
__device__ inline int b_col(int row, int k) {
    if constexpr (PR::DLDS) {
        constexpr int VB = 16 / sizeof(typename PR::L);
        return (((k / VB) ^ (row & 7)) * VB) + (k % VB);
    } else {
        return k;
    }
}

// One BM x BN output tile of rhs b (rows t0 = rt*BM.., sites y0 = ct*BN..).
// Called by the whole workgroup; returns this thread's share of the tile's
// reduction (|u|^2 or |r|^2 in the CG modes).
template <class TL, class PR, int OP, int MODE, int DISC>
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
        // passes over the (plane, row) pairs of the B chunk
        constexpr int BPASS = (PR::NPB * BN + RPPB - 1) / RPPB;

        static_assert(APASS * RPP == BM && BPASS * RPPB == PR::NPB * BN, "loader passes must tile BM and BN exactly");
        // direct B loads: one wave instruction fills 8 rows of 128 bytes
        constexpr int NWAVE = NT / 64;
        constexpr int BSEG = BN / 8;
        constexpr int BPW = BSEG / NWAVE;
        static_assert(!PR::DLDS || (BK * sizeof(L) == 128 && BSEG % NWAVE == 0), "direct B load geometry");
        // lane's 16-byte group of E for B row seg*8 + lane/8 (swizzle of b_col)
        auto load_b_direct = [&](int k0, int buf) {
            #pragma unroll
            for (int p = 0; p < BPW; ++p) {
                const int seg = p * NWAVE + wave;
                const int row = seg * 8 + (lane >> 3);
                const int kg = (lane & 7) ^ (lane >> 3);
                #pragma unroll
                for (int pl = 0; pl < PR::NPB; ++pl) {
                    const L* src = reinterpret_cast<const L*>(J.E) + (size_t)pl * planeE + (size_t)(y0 + row) * ldE + k0 +
                                   kg * VB;
                    __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)src,
                                                     (void __attribute__((address_space(3)))*)&sB[buf][pl][seg * 8][0],
                                                     16, 0, 0);
                }
            }
        };

        // ---- staging of one K chunk: global loads into registers (branch-
        // free: out-of-field elements read g_zero_run), then LDS stores ----
        struct Stage {
            CS<S> x[APASS][VEC];
            CS<S> po[APASS][VEC];   // CGA: previous search direction
            CS<S> ph[APASS][VEC];   // M: phase
            u32x4 e[PR::DLDS ? 1 : BPASS];
        };
        static_assert(!PR::DLDS || PR::NBUF == 2, "direct E loads need the second LDS buffer");
        // global loads of chunk k0 (with DLDS the E rows go straight to LDS buffer buf)
        auto load = [&](Stage& g, int k0, int buf) {
            #pragma unroll
            for (int ps = 0; ps < APASS; ++ps) {
                const int i = ps * RPP + tid / TPR;
                const int kk = (tid % TPR) * VEC;
                const int t = t0 + i;
                const bool vrow = t < nt;
                // EXP: the GEMM reads the neighbouring slice; DIAG: the same slice
                const int tin = (DISC == DISC_DIAG) ? t
                              : (OP == HOP_M) ? (t == 0 ? nt - 1 : t - 1) : (t + 1 >= nt ? 0 : t + 1);
                const size_t rowoff = bbase + (size_t)tin * nx;
                #pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    const int x = k0 + kk + e;
                    const bool ok = vrow && x < nx;
                    g.x[ps][e] = ldc_or0<S>(J.X, rowoff + x, ok);
                    if (MODE == MODE_CGA) g.po[ps][e] = ldc_or0<S>(J.Pold, rowoff + x, ok);
                    if (OP == HOP_M && DISC == DISC_EXP) g.ph[ps][e] = ldc_or0<S>(Pf, (size_t)tin * nx + x, ok);
                }
            }
            // rows y of E, zero padded to ldE columns / e_rows rows: always in bounds
            if constexpr (PR::DLDS) {
                load_b_direct(k0, buf);
            } else {
                #pragma unroll
                for (int ps = 0; ps < BPASS; ++ps) {
                    const int jp = ps * RPPB + tid / TPRB;     // plane * BN + row
                    const int pl = jp / BN, j = jp % BN;
                    const int kk = (tid % TPRB) * VB;
                    const size_t eoff = (size_t)pl * planeE + (size_t)(y0 + j) * ldE + k0 + kk;
                    g.e[ps] = *reinterpret_cast<const u32x4*>(reinterpret_cast<const L*>(J.E) + eoff);
                }
            }
        };
        auto store = [&](const Stage& g, int buf) {
            #pragma unroll
            for (int ps = 0; ps < APASS; ++ps) {
                const int i = ps * RPP + tid / TPR;
                const int kk = (tid % TPR) * VEC;
                #pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    CS<S> z = g.x[ps][e];
                    if (MODE == MODE_CGA) {
                        z.re = z.re + S(beta) * g.po[ps][e].re;
                        z.im = z.im + S(beta) * g.po[ps][e].im;
                    }
                    if (OP == HOP_M && DISC == DISC_EXP) {
                        const CS<S> ph = g.ph[ps][e];
                        const S re = ph.re * z.re - ph.im * z.im;
                        const S im = ph.re * z.im + ph.im * z.re;
                        z.re = re; z.im = im;
                    }
                    if constexpr (PR::NPA == 2) {
                        sA[buf][0][i][kk + e] = z.re;
                        sA[buf][1][i][kk + e] = z.im;
                    } else {
                        const uint16_t rh = f2bf(z.re), ih = f2bf(z.im);
                        sA[buf][0][i][kk + e] = rh;
                        sA[buf][1][i][kk + e] = f2bf(z.re - bf2f(rh));
                        sA[buf][2][i][kk + e] = ih;
                        sA[buf][3][i][kk + e] = f2bf(z.im - bf2f(ih));
                    }
                }
            }
            if constexpr (!PR::DLDS) {
                #pragma unroll
                for (int ps = 0; ps < BPASS; ++ps) {
                    const int jp = ps * RPPB + tid / TPRB;
                    const int pl = jp / BN, j = jp % BN;
                    const int kk = (tid % TPRB) * VB;
                    *reinterpret_cast<u32x4*>(&sB[buf][pl][j][kk]) = g.e[ps];
                }
            }
        };

        // ---- MFMA over the chunk in LDS buffer cur ----
        auto mfma = [&](const int cur) {
            if constexpr (PR::KSTEP == 4) {
                #pragma unroll
                for (int s = 0; s < RUN; ++s) {
                    const int k = q * RUN + s;
                    L ar[FM], ai[FM], bv[FN];
                    #pragma unroll
                    for (int i = 0; i < FM; ++i) {
                        const int row = (wm * FM + i) * 16 + r16;
                        ar[i] = sA[cur][0][row][k];
                        ai[i] = sA[cur][1][row][k];
                    }
                    #pragma unroll
                    for (int j = 0; j < FN; ++j) {
                        const int col = (wn * FN + j) * 16 + r16;
                        bv[j] = sB[cur][0][col][b_col<PR>(col, k)];
                    }
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
                        arh[i] = *reinterpret_cast<const bf8*>(&sA[cur][0][row][k]);
                        arl[i] = *reinterpret_cast<const bf8*>(&sA[cur][1][row][k]);
                        aih[i] = *reinterpret_cast<const bf8*>(&sA[cur][2][row][k]);
                        ail[i] = *reinterpret_cast<const bf8*>(&sA[cur][3][row][k]);
                    }
                    #pragma unroll
                    for (int j = 0; j < FN; ++j) {
                        const int col = (wn * FN + j) * 16 + r16;
                        bh[j] = *reinterpret_cast<const bf8*>(&sB[cur][0][col][b_col<PR>(col, k)]);
                        bl[j] = *reinterpret_cast<const bf8*>(&sB[cur][1][col][b_col<PR>(col, k)]);
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
        };

        const int nch = (nx + BK - 1) / BK;
        Stage stg;
        load(stg, 0, 0);
        store(stg, 0);
        if constexpr (PR::DLDS) __builtin_amdgcn_s_waitcnt(0);   // the direct E loads have landed
        __syncthreads();
        for (int c = 0; c + 1 < nch; ++c) {
            // the next chunk's global loads are in flight during this chunk's
            // MFMAs: the scheduling barrier keeps global loads above the MFMAs
            // (VALU, SALU and LDS reads may cross; mask 0x146; for the F64
            // 32 x 96 tile the loads go between the first MFMAs instead); the
            // LDS writes of the next chunk are placed among the last MFMAs (F64
            // 32 x 96 tile, BF16x3) or after them (mask 0x146 again).  On the
            // other F64 tiles the fixed pattern was slower (16 x 32: batching
            // leg -3 %, serving p99 +3 %).  With DLDS the E rows of the next
            // chunk go to the other buffer, which every wave has finished
            // reading (barrier at the end of the previous iteration).
            const int nxt = PR::NBUF == 2 ? ((c + 1) & 1) : 0;
            load(stg, (c + 1) * BK, nxt);
            constexpr bool F64B = std::is_same<PR, PF64>::value && TL::FM == 1 && TL::FN == 3;
            if constexpr (!F64B) __builtin_amdgcn_sched_barrier(0x146);
            mfma(PR::NBUF == 2 ? (c & 1) : 0);
            if (PR::NBUF == 1) __syncthreads();   // all waves done reading the only buffer
            store(stg, nxt);
            if constexpr (F64B && PR::NBUF == 2) {
                // F64, 32 x 96 tile (24 MFMAs per chunk): the 7 global loads of the next
                // chunk one per MFMA at the start, its 5 LDS writes (2 A and 3
                // E rows of 16 bytes per lane) one per two MFMAs at the end
                // rather than after them: CG kernels 4-5 % faster.
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 7, 0);
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
            } else if constexpr (std::is_same<PR, PB3>::value && PR::NBUF == 2) {
                // BF16x3 (18 MFMAs, 4 A and 3 E writes per chunk): plain M 3 %
                // faster.  The same for F32 was 1 % slower: not used there.
                __builtin_amdgcn_sched_group_barrier(0x008, 10, 0);
                __builtin_amdgcn_sched_group_barrier(0x200, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x200, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
            } else {
                __builtin_amdgcn_sched_barrier(0x146);
            }
            if constexpr (PR::DLDS) __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
        mfma(PR::NBUF == 2 ? ((nch - 1) & 1) : 0);
        __syncthreads();   // the LDS is reused by the next tile of a grouped launch

        // ---- epilogue: shift sign, identity, conj(P), fused recurrences ----
        // BF16x3: per accumulator row (i, v), first every load of the row's
        // FN elements (branch-free, out-of-range elements read g_zero_run),
        // then the arithmetic and the stores; element by element each load
        // waits for the previous element's stores (the compiler cannot move a
        // load above a store to a buffer it may alias): plain BF16x3 applies
        // 5 % faster, the phi = 0 mixed solve 4 %.  F64 and F32 keep the
        // element-by-element form: the batched loads took the F64 32 x 96 CG
        // kernels past the register budget of 3 waves per SIMD (any
        // rearrangement there measured slower), and the F32 grouped kernels
        // of the batching leg ran 2 % slower with them.
        if constexpr (std::is_same<PR, PB3>::value) {
            #pragma unroll
            for (int i = 0; i < FM; ++i) {
                #pragma unroll
                for (int v = 0; v < 4; ++v) {
                    // accumulator register v of lane (q, r16) holds row 4q+v of the
                    // 16x16 block for the f32/bf16 MFMAs, row 4v+q for f64 (its
                    // output groups have size 1)
                    const int t = t0 + (wm * FM + i) * 16 + (PR::ROW4V ? 4 * v + q : 4 * q + v);
                    const bool tok = t < nt;
                    const int tc = tok ? t : 0;
                    const size_t rowoff = bbase + (size_t)tc * nx;
                    // DIAG: the neighbouring slice of the hop input and its phase
                    const int tn = (OP == HOP_M) ? (tc == 0 ? nt - 1 : tc - 1) : (tc + 1 >= nt ? 0 : tc + 1);
                    CS<S> lx[FN], lp[FN], lxs[FN], lr[FN], lph[FN], ln[FN], lpn[FN];
                    #pragma unroll
                    for (int j = 0; j < FN; ++j) {
                        const int y = y0 + (wn * FN + j) * 16 + r16;
                        const bool ok = tok && y < nx;
                        const size_t idx = rowoff + y;
                        lx[j] = ldc_or0<S>(J.X, idx, ok);
                        if (MODE == MODE_CGA || MODE == MODE_CGB) lp[j] = ldc_or0<S>(J.Pold, idx, ok);
                        if (MODE == MODE_CGB) {
                            lxs[j] = ldc_or0<S>(J.Xs, idx, ok);
                            lr[j] = ldc_or0<S>(J.R, idx, ok);
                        }
                        if (MODE == MODE_CGI) lr[j] = ldc_or0<S>(J.Bv, idx, ok);
                        if (DISC == DISC_DIAG) {
                            const size_t nidx = bbase + (size_t)tn * nx + y;
                            ln[j] = ldc_or0<S>(J.X, nidx, ok);
                            if (MODE == MODE_CGA) lpn[j] = ldc_or0<S>(J.Pold, nidx, ok);
                            lph[j] = ldc_or0<S>(Pf, (size_t)(OP == HOP_M ? tn : tc) * nx + y, ok);
                        } else if (OP != HOP_M) {
                            lph[j] = ldc_or0<S>(Pf, (size_t)tc * nx + y, ok);   // conj(P_t) * acc
                        }
                    }
                    if (!tok) continue;
                    #pragma unroll
                    for (int j = 0; j < FN; ++j) {
                        const int y = y0 + (wn * FN + j) * 16 + r16;
                        if (y >= nx) continue;
                        const size_t idx = rowoff + y;
                        S fr = S(accR[i][j][v]), fi = S(accI[i][j][v]);
                        if (DISC == DISC_DIAG) {
                            // + b_t P_{t-1} psi_{t-1} (M) or + b'_t conj(P_t) psi_{t+1} (M^dag),
                            // psi = the hop input (r + beta p in CGA)
                            CS<S> xn = ln[j];
                            if (MODE == MODE_CGA) {
                                xn.re = xn.re + S(beta) * lpn[j].re;
                                xn.im = xn.im + S(beta) * lpn[j].im;
                            }
                            const CS<S> ph = lph[j];
                            S dr, di;
                            if (OP == HOP_M) {
                                dr = ph.re * xn.re - ph.im * xn.im;
                                di = ph.re * xn.im + ph.im * xn.re;
                            } else {
                                dr = ph.re * xn.re + ph.im * xn.im;
                                di = ph.re * xn.im - ph.im * xn.re;
                            }
                            if (OP == HOP_M ? t == 0 : t == nt - 1) { dr = -dr; di = -di; }
                            fr = fr + dr;
                            fi = fi + di;
                        } else if (OP == HOP_M) {
                            if (t == 0) { fr = -fr; fi = -fi; }
                        } else {
                            const CS<S> ph = lph[j];
                            const S re = ph.re * fr + ph.im * fi;
                            const S im = ph.re * fi - ph.im * fr;
                            fr = re; fi = im;
                            if (t == nt - 1) { fr = -fr; fi = -fi; }
                        }
                        CS<S> id;
                        if (MODE == MODE_CGA) {
                            id.re = lx[j].re + S(beta) * lp[j].re;
                            id.im = lx[j].im + S(beta) * lp[j].im;
                        } else {
                            id = lx[j];
                        }
                        const CS<S> out{id.re - fr, id.im - fi};
                        if (MODE == MODE_PLAIN) {
                            stc<S>(J.Y, idx, out);
                        } else if (MODE == MODE_CGA) {
                            stc<S>(J.Y, idx, out);
                            stc<S>(J.Pnew, idx, id);   // p_new = r + beta p_old, each element once
                            part += double(out.re) * out.re + double(out.im) * out.im;
                        } else if (MODE == MODE_CGB) {
                            CS<S> xv = lxs[j], rv = lr[j];
                            xv.re += S(alpha) * lp[j].re; xv.im += S(alpha) * lp[j].im;
                            rv.re -= S(alpha) * out.re; rv.im -= S(alpha) * out.im;
                            stc<S>(J.Xs, idx, xv);
                            stc<S>(J.R, idx, rv);
                            part += double(rv.re) * rv.re + double(rv.im) * rv.im;
                        } else {  // CGI
                            const CS<S> bv = lr[j];
                            const CS<S> rv{bv.re - out.re, bv.im - out.im};
                            stc<S>(J.R, idx, rv);
                            stc<S>(J.Pnew, idx, rv);
                            part += double(rv.re) * rv.re + double(rv.im) * rv.im;
                        }
                    }
                }
            }
        } else {
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
                        if (DISC == DISC_DIAG) {
                            // + b_t P_{t-1} psi_{t-1} (M) or + b'_t conj(P_t) psi_{t+1} (M^dag),
                            // psi = the hop input (r + beta p in CGA)
                            const int tn = (OP == HOP_M) ? (t == 0 ? nt - 1 : t - 1) : (t + 1 >= nt ? 0 : t + 1);
                            const size_t nidx = bbase + (size_t)tn * nx + y;
                            CS<S> xn = ldc<S>(J.X, nidx);
                            if (MODE == MODE_CGA) {
                                const CS<S> pn = ldc<S>(J.Pold, nidx);
                                xn.re = xn.re + S(beta) * pn.re;
                                xn.im = xn.im + S(beta) * pn.im;
                            }
                            const CS<S> ph = ldc<S>(Pf, (size_t)(OP == HOP_M ? tn : t) * nx + y);
                            S dr, di;
                            if (OP == HOP_M) {
                                dr = ph.re * xn.re - ph.im * xn.im;
                                di = ph.re * xn.im + ph.im * xn.re;
                            } else {
                                dr = ph.re * xn.re + ph.im * xn.im;
                                di = ph.re * xn.im - ph.im * xn.re;
                            }
                            if (OP == HOP_M ? t == 0 : t == nt - 1) { dr = -dr; di = -di; }
                            fr = fr + dr;
                            fi = fi + di;
                        } else if (OP == HOP_M) {
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
                            stc<S>(J.Pnew, idx, id);   // p_new = r + beta p_old, each element once
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

template <class TL, class PR, int OP, int MODE, int DISC>
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
        part = hop_tile<TL, PR, OP, MODE, DISC>(J, a.nx, a.nt, a.rpf, ldE, planeE, b, rt, ct, alpha, beta, sm);
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
template <class TL, class PR, int OP, int MODE, int DISC>
__global__ void __launch_bounds__(TL::NT) hop_grouped(GroupArgs g) {
    constexpr int NT = TL::NT;
    __shared__ __attribute__((aligned(16))) Smem<TL, PR> sm;
    __shared__ double red[NT / 64];
    const int it = g.iter_dev ? *g.iter_dev + g.iter : g.iter;
    const int ntiles = g.ntiles_dev ? *g.ntiles_dev : g.ntiles;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
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
            const double rr1 = sum_parts(c.rpart[(it - 1) & 1], off, nparts);
            const double rr2 = (MODE == MODE_CGA && it >= 2) ? sum_parts(c.rpart[(it - 2) & 1], off, nparts) : 0.0;
            const double uu = MODE == MODE_CGB ? sum_parts(c.upart, off, nparts) : 0.0;
            skip = cg_decide<MODE>(c, slot, it, rt == 0 && ct == 0, rr1, rr2, uu, alpha, beta);
        }
        double part = 0.0;
        if (!skip) {
            auto v = [&](int r) -> void* { return r >= 0 ? P.v[r] : nullptr; };
            const JobPtrs J{P.E, P.P, v(g.role[0]), v(g.role[1]), v(g.role[2]), v(g.role[3]),
                            v(g.role[4]), v(g.role[5]), v(g.role[6])};
            part = hop_tile<TL, PR, OP, MODE, DISC>(J, P.nx, P.nt, P.rpf, P.ldE, (size_t)P.planeE, b, rt, ct, alpha,
                                          beta, sm);
        }
        if (MODE != MODE_PLAIN && !skip) {
            const double tot = block_sum<NT>(part, red);
            if (threadIdx.x == 0) {
                double* dst = MODE == MODE_CGA ? g.cg.upart : g.cg.rpart[MODE == MODE_CGB ? (it & 1) : 0];
                dst[off + rt * P.nct + ct] = tot;
            }
        } else {
            __syncthreads();   // LDS reuse by the next tile
        }
    }
}

// ------------------------------------------------ slice-resident kernels
// A workgroup keeps the column slice E[y0 .. y0+BN)[0 .. kp) in LDS for its
// whole life (loaded once, one barrier) and each of its W waves computes
// 16-row output blocks with the full K loop on its own: the wave streams its
// 16 input rows (time-shifted, phase-multiplied) straight from global memory
// into MFMA operand registers, one K chunk ahead of the MFMAs, with no LDS
// staging of the field and no barrier inside the K loop.  Lane (r16, q) feeds
// row r16 and the contiguous k run [k0 + q*VK, k0 + (q+1)*VK) of each chunk
// (the same k permutation for the E operand, read from LDS as 16-byte
// vectors).  The LDS row stride is padded so that the 16-byte reads of eight
// consecutive lanes cover all 32 banks once.
template <class PR>
struct SliceP;
template <>
struct SliceP<PF64> { static constexpr int KC = 16, VK = 4, LPAD = 2; };
template <>
struct SliceP<PF32> { static constexpr int KC = 16, VK = 4, LPAD = 4; };
template <>
struct SliceP<PB3> { static constexpr int KC = 32, VK = 8, LPAD = 8; };

template <class S, int VK>
struct Raw {
    CS<S> x[VK];    // input row chunk
    CS<S> ph[VK];   // phase chunk (M)
    CS<S> po[VK];   // previous search direction (CGA)
};


// VK complex (64 bytes) at element `off` of base, 16-byte aligned.  Branch-
// and select-free on the loaded data, so that the compiler can keep several
// chunks of loads in flight across the K loop: a run that is not valid (row
// past nt or chunk past kp) reads g_zero_run instead.
template <class S, int VK>
__device__ inline void load_run(const void* base, size_t off, bool ok, CS<S>* out) {
    static_assert(VK * 2 * sizeof(S) <= sizeof(g_zero_run), "zero run size");
    if constexpr (sizeof(S) == 8) {
        const double2* p = ok ? reinterpret_cast<const double2*>(base) + off
                              : reinterpret_cast<const double2*>(g_zero_run);
        #pragma unroll
        for (int e = 0; e < VK; ++e) {
            const double2 v = p[e];
            out[e] = CS<S>{v.x, v.y};
        }
    } else {
        const float4* p = ok ? reinterpret_cast<const float4*>(reinterpret_cast<const float2*>(base) + off)
                             : reinterpret_cast<const float4*>(g_zero_run);
        #pragma unroll
        for (int e = 0; e < VK / 2; ++e) {
            const float4 v = p[e];
            out[2 * e] = CS<S>{v.x, v.y};
            out[2 * e + 1] = CS<S>{v.z, v.w};
        }
    }
}

typedef unsigned short us8 __attribute__((ext_vector_type(8)));

// One 16 x BN output block (rows rt*16.., columns y0..) of rhs b by one wave.
// Returns the lane's share of the block's reduction (CG modes).
template <class PR, int FN, int OP, int MODE>
__device__ double slice_task(const JobPtrs& J, const typename PR::L* __restrict__ sE, int ld, int kp, int nx, int nt,
                             int rpf, int b, int rt, int s, double alpha, double beta) {
    using S = typename PR::S;
    using Acc = typename PR::Acc;
    constexpr int KC = SliceP<PR>::KC, VK = SliceP<PR>::VK, BN = FN * 16;
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, q = lane >> 4;
    const int f = b / rpf;
    const S* Pf = reinterpret_cast<const S*>(J.P) + (size_t)f * nt * nx * 2;
    const size_t bbase = (size_t)b * nt * nx;
    const int ta = rt * 16 + r16;
    const bool vrow = ta < nt;
    const int tin = (OP == HOP_M) ? (ta == 0 ? nt - 1 : ta - 1) : (ta + 1 >= nt ? 0 : ta + 1);
    const size_t xoff = bbase + (size_t)tin * nx;
    const size_t poff = (size_t)tin * nx;
    const int y0 = s * BN;

    Acc accR[FN], accI[FN];
    #pragma unroll
    for (int j = 0; j < FN; ++j) { accR[j] = Acc{0, 0, 0, 0}; accI[j] = Acc{0, 0, 0, 0}; }

    // nx is a multiple of VK for these tiles, so a lane's run is entirely
    // inside the row (k < nx) or entirely past it; kp is a multiple of 3 KC
    // and the LDS slice is zero past nx
    const int nch = kp / KC;
    auto load = [&](Raw<S, VK>& r, int c) {
        const int k = c * KC + q * VK;
        const bool ok = vrow && k < nx;
        load_run<S, VK>(J.X, xoff + k, ok, r.x);
        if (OP == HOP_M) load_run<S, VK>(Pf, poff + k, ok, r.ph);
        if (MODE == MODE_CGA) load_run<S, VK>(J.Pold, xoff + k, ok, r.po);
    };
    auto compute = [&](const Raw<S, VK>& r, int c) {
        const int k = c * KC + q * VK;
        CS<S> z[VK];
        #pragma unroll
        for (int e = 0; e < VK; ++e) {
            z[e] = r.x[e];
            if (MODE == MODE_CGA) {
                z[e].re = z[e].re + S(beta) * r.po[e].re;
                z[e].im = z[e].im + S(beta) * r.po[e].im;
            }
            if (OP == HOP_M) {
                const S re = r.ph[e].re * z[e].re - r.ph[e].im * z[e].im;
                const S im = r.ph[e].re * z[e].im + r.ph[e].im * z[e].re;
                z[e].re = re; z[e].im = im;
            }
        }
        if constexpr (PR::KSTEP == 4) {
            #pragma unroll
            for (int j = 0; j < FN; ++j) {
                const S* bp = sE + (size_t)(j * 16 + r16) * ld + k;
                S bv[VK];
                if constexpr (sizeof(S) == 8) {
                    #pragma unroll
                    for (int e = 0; e < VK; e += 2) {
                        const double2 w = *reinterpret_cast<const double2*>(bp + e);
                        bv[e] = w.x; bv[e + 1] = w.y;
                    }
                } else {
                    const float4 w = *reinterpret_cast<const float4*>(bp);
                    bv[0] = w.x; bv[1] = w.y; bv[2] = w.z; bv[3] = w.w;
                }
                #pragma unroll
                for (int e = 0; e < VK; ++e) {
                    if constexpr (sizeof(S) == 8) {
                        accR[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(z[e].re, bv[e], accR[j], 0, 0, 0);
                        accI[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(z[e].im, bv[e], accI[j], 0, 0, 0);
                    } else {
                        accR[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(z[e].re, bv[e], accR[j], 0, 0, 0);
                        accI[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(z[e].im, bv[e], accI[j], 0, 0, 0);
                    }
                }
            }
        } else {
            us8 rh, rl, ih, il;
            #pragma unroll
            for (int e = 0; e < VK; ++e) {
                const uint16_t a = f2bf(z[e].re), c2 = f2bf(z[e].im);
                rh[e] = a; rl[e] = f2bf(z[e].re - bf2f(a));
                ih[e] = c2; il[e] = f2bf(z[e].im - bf2f(c2));
            }
            const bf8 arh = __builtin_bit_cast(bf8, rh), arl = __builtin_bit_cast(bf8, rl);
            const bf8 aih = __builtin_bit_cast(bf8, ih), ail = __builtin_bit_cast(bf8, il);
            #pragma unroll
            for (int j = 0; j < FN; ++j) {
                const uint16_t* bp = reinterpret_cast<const uint16_t*>(sE) + (size_t)(j * 16 + r16) * ld + k;
                const bf8 bh = *reinterpret_cast<const bf8*>(bp);
                const bf8 bl = *reinterpret_cast<const bf8*>(bp + (size_t)BN * ld);
                accR[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arl, bh, accR[j], 0, 0, 0);
                accR[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh, bl, accR[j], 0, 0, 0);
                accR[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh, bh, accR[j], 0, 0, 0);
                accI[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ail, bh, accI[j], 0, 0, 0);
                accI[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih, bl, accI[j], 0, 0, 0);
                accI[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih, bh, accI[j], 0, 0, 0);
            }
        }
    };

    // three chunk buffers in rotation: the loads of chunks c+1 and c+2 are in
    // flight while chunk c feeds the MFMAs; no branch inside the loop, so the
    // waits before each chunk only cover that chunk's loads (the two loads
    // past the end read the zero run)
    // (the scheduling barriers keep the compiler from sinking the prefetch
    // loads next to their uses)
    Raw<S, VK> r0, r1, r2;
    load(r0, 0);
    load(r1, 1);
    for (int c = 0; c < nch; c += 3) {
        load(r2, c + 2);
        __builtin_amdgcn_sched_barrier(0);
        compute(r0, c);
        __builtin_amdgcn_sched_barrier(0);
        load(r0, c + 3);
        __builtin_amdgcn_sched_barrier(0);
        compute(r1, c + 1);
        __builtin_amdgcn_sched_barrier(0);
        load(r1, c + 4);
        __builtin_amdgcn_sched_barrier(0);
        compute(r2, c + 2);
        __builtin_amdgcn_sched_barrier(0);
    }

    // epilogue (as in hop_tile): shift sign, identity, conj(P), fused recurrences
    double part = 0.0;
    #pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int t = rt * 16 + (PR::ROW4V ? 4 * v + q : 4 * q + v);
        if (t >= nt) continue;
        const size_t rowoff = bbase + (size_t)t * nx;
        #pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int y = y0 + j * 16 + r16;
            if (y >= nx) continue;
            const size_t idx = rowoff + y;
            S fr = S(accR[j][v]), fi = S(accI[j][v]);
            if (OP == HOP_M) {
                if (t == 0) { fr = -fr; fi = -fi; }
            } else {
                const CS<S> ph = ldc<S>(Pf, (size_t)t * nx + y);
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
                stc<S>(J.Pnew, idx, id);   // p_new = r + beta p_old, each element once
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
    return part;
}

// Grid: nct column slices x P row groups (XCD-swizzled like hop_kernel: the
// nct slices of a row group are 8 block ids apart).  Row group gi owns the
// row blocks [gi*R*W, (gi+1)*R*W); wave w takes blocks (gi*R + i)*W + w.
template <class PR, int FN, int W, int OP, int MODE>
__global__ void __launch_bounds__(64 * W) hop_slice(HopArgs a, int nrt, int nct, int P, int R, int swz, int kp,
                                                    int ldE, size_t planeE) {
    using L = typename PR::L;
    constexpr int NT = 64 * W, BN = FN * 16;
    extern __shared__ __attribute__((aligned(16))) unsigned char slice_lds[];
    __shared__ double red[W];
    L* sE = reinterpret_cast<L*>(slice_lds);
    const int ld = kp + SliceP<PR>::LPAD;

    int gi, s;
    if (swz) {
        const int xcd = blockIdx.x & 7, l = blockIdx.x >> 3;
        s = l % nct;
        gi = (l / nct) * 8 + xcd;
    } else {
        s = blockIdx.x % nct;
        gi = blockIdx.x / nct;
    }
    if (gi >= P) return;

    // the E slice, once (columns past the padded global row are zero)
    constexpr int VL = 16 / sizeof(L);
    const int vpr = kp / VL;
    const int y0 = s * BN;
    #pragma unroll
    for (int pl = 0; pl < PR::NPB; ++pl) {
        const L* src = reinterpret_cast<const L*>(a.E) + (size_t)pl * planeE;
        for (int i = threadIdx.x; i < BN * vpr; i += NT) {
            const int j = i / vpr, kv = (i - j * vpr) * VL;
            const u32x4 w = kv < ldE ? *reinterpret_cast<const u32x4*>(src + (size_t)(y0 + j) * ldE + kv)
                                     : u32x4{0u, 0u, 0u, 0u};
            *reinterpret_cast<u32x4*>(sE + (size_t)(pl * BN + j) * ld + kv) = w;
        }
    }
    const int nparts = nrt * nct;
    double g1 = 0.0, g2 = 0.0, gu = 0.0;
    if ((MODE == MODE_CGA || MODE == MODE_CGB) && a.cg.global) {
        const size_t n = (size_t)a.nb * nparts;
        g1 = block_sum_parts<NT>(a.cg.rpart[(a.iter - 1) & 1], n, red);
        if (MODE == MODE_CGA && a.iter >= 2) g2 = block_sum_parts<NT>(a.cg.rpart[(a.iter - 2) & 1], n, red);
        if (MODE == MODE_CGB) gu = block_sum_parts<NT>(a.cg.upart, n, red);
    }
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int G = a.nb * nrt;
    const JobPtrs J{a.E, a.P, a.X, a.Y, a.Pold, a.Pnew, a.Xs, a.R, a.Bv};
    for (int i = 0; i < R; ++i) {
        const int rb = (gi * R + i) * W + wave;
        if (rb >= G) break;
        const int b = rb / nrt, rt = rb - (rb / nrt) * nrt;
        double alpha = 0.0, beta = 0.0;
        if (MODE == MODE_CGA || MODE == MODE_CGB) {
            const CgDev& c = a.cg;
            const int it = a.iter;
            double rr1 = g1, rr2 = g2, uu = gu;
            if (!c.global) {
                const size_t off = (size_t)b * nparts;
                rr1 = sum_parts(c.rpart[(it - 1) & 1], off, nparts);
                rr2 = (MODE == MODE_CGA && it >= 2) ? sum_parts(c.rpart[(it - 2) & 1], off, nparts) : 0.0;
                uu = MODE == MODE_CGB ? sum_parts(c.upart, off, nparts) : 0.0;
            }
            const int slot = c.global ? 0 : b;
            const bool designated = rt == 0 && s == 0 && (!c.global || b == 0);
            if (cg_decide<MODE>(c, slot, it, designated, rr1, rr2, uu, alpha, beta)) continue;
        }
        double part = slice_task<PR, FN, OP, MODE>(J, sE, ld, kp, a.nx, a.nt, a.rpf, b, rt, s, alpha, beta);
        if (MODE == MODE_PLAIN) continue;
        #pragma unroll
        for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
        if (lane == 0) {
            double* dst = MODE == MODE_CGA ? a.cg.upart : a.cg.rpart[MODE == MODE_CGB ? (a.iter & 1) : 0];
            dst[(size_t)b * nparts + rt * nct + s] = part;
        }
    }
}

// K extent of the LDS slice: nx rounded up to three chunks (the K loop's unroll).
template <class PR>
int slice_kp(int nx) {
    constexpr int K3 = 3 * SliceP<PR>::KC;
    return (nx + K3 - 1) / K3 * K3;
}

template <class PR, int FN>
size_t slice_lds_bytes(int nx) {
    return (size_t)PR::NPB * FN * 16 * (slice_kp<PR>(nx) + SliceP<PR>::LPAD) * sizeof(typename PR::L);
}

template <class PR, int FN, int W, int OP, int MODE>
void launch_slice_k(const HopArgs& a, int R, hipStream_t s) {
    constexpr int BN = FN * 16;
    NSLX_REQUIRE(a.nx % SliceP<PR>::VK == 0, "hop: slice tiles need nx to be a multiple of the lane run");
    const int nrt = (a.nt + 15) / 16, nct = (a.nx + BN - 1) / BN;
    const int kp = slice_kp<PR>(a.nx);
    const size_t lds = slice_lds_bytes<PR, FN>(a.nx);
    NSLX_REQUIRE(lds + W * sizeof(double) <= 160 * 1024, "hop: lattice too large for this slice tile");
    const int G = a.nb * nrt;
    const int P = (G + R * W - 1) / (R * W);
    const int swz = P >= 8 ? 1 : 0;
    const int grid = swz ? ((P + 7) / 8) * 8 * nct : P * nct;
    const int ldE = e_ld(a.nx);
    const size_t planeE = (size_t)e_rows(a.nx) * ldE;
    auto kern = hop_slice<PR, FN, W, OP, MODE>;
    static const bool attr = [&] {   // once per kernel, thread-safe
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  150 * 1024);
        (void)hipGetLastError();
        return true;
    }();
    (void)attr;
    launch_timed(kern, dim3(grid), dim3(64 * W), lds, s, a, nrt, nct, P, R, swz, kp, ldE, planeE);
}

template <class PR, int FN, int W>
void launch_slice(int op, int mode, const HopArgs& a, int R, hipStream_t s) {
#define NSLX_SLICE_CASE(OPV, MV)                           \
    if (op == OPV && mode == MV) {                         \
        launch_slice_k<PR, FN, W, OPV, MV>(a, R, s);       \
        return;                                            \
    }
    NSLX_SLICE_CASE(HOP_M, MODE_PLAIN)
    NSLX_SLICE_CASE(HOP_MD, MODE_PLAIN)
    NSLX_SLICE_CASE(HOP_M, MODE_CGA)
    NSLX_SLICE_CASE(HOP_MD, MODE_CGA)
    NSLX_SLICE_CASE(HOP_M, MODE_CGB)
    NSLX_SLICE_CASE(HOP_MD, MODE_CGB)
    NSLX_SLICE_CASE(HOP_M, MODE_CGI)
    NSLX_SLICE_CASE(HOP_MD, MODE_CGI)
#undef NSLX_SLICE_CASE
    throw Error(NSLX_E_ARG, "hop: bad op/mode");
}

// Slice tile menus per precision, tile ids SLICE_BASE + index:
// (FN: slice width / 16, W: waves per workgroup, R: row blocks per wave).
constexpr int SLICE_BASE = 4;
struct SliceEntry { int fn, w, r; };
constexpr int SLICE_N = 4;
constexpr SliceEntry SLICE_MENU64[SLICE_N] = {{3, 8, 1}, {2, 4, 1}, {2, 8, 1}, {1, 8, 1}};
constexpr SliceEntry SLICE_MENU32[SLICE_N] = {{3, 8, 1}, {6, 8, 1}, {3, 4, 1}, {2, 8, 1}};
constexpr SliceEntry SLICE_MENUB3[SLICE_N] = {{6, 8, 1}, {3, 8, 1}, {3, 4, 1}, {2, 8, 1}};

const SliceEntry* slice_menu(int prec) {
    return prec == NSLX_F64 ? SLICE_MENU64 : prec == NSLX_F32 ? SLICE_MENU32 : SLICE_MENUB3;
}

template <class PR, int W, int FA, int FB, int FC>
void dispatch_slice_w(const SliceEntry& e, int op, int mode, const HopArgs& a, hipStream_t s) {
    if (e.fn == FA) launch_slice<PR, FA, W>(op, mode, a, e.r, s);
    else if (e.fn == FB) launch_slice<PR, FB, W>(op, mode, a, e.r, s);
    else if (e.fn == FC) launch_slice<PR, FC, W>(op, mode, a, e.r, s);
    else throw Error(NSLX_E_ARG, "hop: bad slice tile");
}

template <class PR, int FA, int FB, int FC>
void dispatch_slice(int id, int op, int mode, const HopArgs& a, hipStream_t s, const SliceEntry* menu) {
    const SliceEntry e = menu[id];
    if (e.w == 4) dispatch_slice_w<PR, 4, FA, FB, FC>(e, op, mode, a, s);
    else if (e.w == 8) dispatch_slice_w<PR, 8, FA, FB, FC>(e, op, mode, a, s);
    else throw Error(NSLX_E_ARG, "hop: bad slice tile");
}

template <class PR>
size_t slice_lds_fn(int fn, int nx) {
    switch (fn) {
        case 1: return slice_lds_bytes<PR, 1>(nx);
        case 2: return slice_lds_bytes<PR, 2>(nx);
        case 3: return slice_lds_bytes<PR, 3>(nx);
        default: return slice_lds_bytes<PR, 6>(nx);
    }
}

bool slice_fits(int prec, int id, int nx) {
    const SliceEntry e = slice_menu(prec)[id];
    size_t lds = 0;
    int vk = 0;
    if (prec == NSLX_F64) { lds = slice_lds_fn<PF64>(e.fn, nx); vk = SliceP<PF64>::VK; }
    else if (prec == NSLX_F32) { lds = slice_lds_fn<PF32>(e.fn, nx); vk = SliceP<PF32>::VK; }
    else { lds = slice_lds_fn<PB3>(e.fn, nx); vk = SliceP<PB3>::VK; }
    return nx % vk == 0 && lds + e.w * sizeof(double) <= 160 * 1024;
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
using TB3_a = Tile<2, 2, 2, 3, 32>;
using TB3_b = Tile<2, 2, 1, 3, 32>;
using TB3_c = Tile<1, 2, 1, 1, 32>;
using TB3_d = Tile<2, 2, 2, 2, 32>;

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
#define NSLX_HOP_CASE(OPV, MV)                                                                          \
    if (op == OPV && mode == MV) {                                                                      \
        if (a.disc == DISC_DIAG)                                                                        \
            launch_timed(hop_kernel<TL, PR, OPV, MV, DISC_DIAG>, gr, bl, 0, s, a, nrt, nct, swz, ldE, planeE); \
        else                                                                                            \
            launch_timed(hop_kernel<TL, PR, OPV, MV, DISC_EXP>, gr, bl, 0, s, a, nrt, nct, swz, ldE, planeE);  \
        return;                                                                                         \
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

// Grouped HubbardDiag launches exist for the CG pair the batching engine
// runs (M^dagger CGA, M CGB); HubbardExp for every op and mode.
template <class TL, class PR>
void launch_g(int op, int mode, int disc, const GroupArgs& g, int grid, hipStream_t s) {
    dim3 gr(grid), bl(TL::NT);
    if (disc == DISC_DIAG) {
        if (op == HOP_MD && mode == MODE_CGA) { launch_timed(hop_grouped<TL, PR, HOP_MD, MODE_CGA, DISC_DIAG>, gr, bl, 0, s, g); return; }
        if (op == HOP_M && mode == MODE_CGB) { launch_timed(hop_grouped<TL, PR, HOP_M, MODE_CGB, DISC_DIAG>, gr, bl, 0, s, g); return; }
        throw Error(NSLX_E_ARG, "hop: grouped HubbardDiag launches are M^dagger CGA and M CGB");
    }
#define NSLX_HOPG_CASE(OPV, MV)                                                    \
    if (op == OPV && mode == MV) {                                                 \
        launch_timed(hop_grouped<TL, PR, OPV, MV, DISC_EXP>, gr, bl, 0, s, g);      \
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

template <class TL, class PR>
GroupedKernel kernel_g(int op, int mode, int disc) {
    if (disc == DISC_DIAG) {
        if (op == HOP_MD && mode == MODE_CGA)
            return {reinterpret_cast<const void*>(&hop_grouped<TL, PR, HOP_MD, MODE_CGA, DISC_DIAG>), TL::NT};
        if (op == HOP_M && mode == MODE_CGB)
            return {reinterpret_cast<const void*>(&hop_grouped<TL, PR, HOP_M, MODE_CGB, DISC_DIAG>), TL::NT};
        throw Error(NSLX_E_ARG, "hop: grouped HubbardDiag launches are M^dagger CGA and M CGB");
    }
#define NSLX_HOPG_FN(OPV, MV)                                                                  \
    if (op == OPV && mode == MV) return {reinterpret_cast<const void*>(&hop_grouped<TL, PR, OPV, MV, DISC_EXP>), TL::NT};
    NSLX_HOPG_FN(HOP_M, MODE_PLAIN)
    NSLX_HOPG_FN(HOP_MD, MODE_PLAIN)
    NSLX_HOPG_FN(HOP_M, MODE_CGA)
    NSLX_HOPG_FN(HOP_MD, MODE_CGA)
    NSLX_HOPG_FN(HOP_M, MODE_CGB)
    NSLX_HOPG_FN(HOP_MD, MODE_CGB)
    NSLX_HOPG_FN(HOP_M, MODE_CGI)
    NSLX_HOPG_FN(HOP_MD, MODE_CGI)
#undef NSLX_HOPG_FN
    throw Error(NSLX_E_ARG, "hop: bad op/mode");
}

template <class PR, class T0, class T1, class T2, class T3>
GroupedKernel kernel4g(int tile, int op, int mode, int disc) {
    switch (tile) {
        case 0: return kernel_g<T0, PR>(op, mode, disc);
        case 1: return kernel_g<T1, PR>(op, mode, disc);
        case 2: return kernel_g<T2, PR>(op, mode, disc);
        case 3: return kernel_g<T3, PR>(op, mode, disc);
        default: throw Error(NSLX_E_ARG, "hop: bad tile id");
    }
}

template <class PR, class T0, class T1, class T2, class T3>
void dispatch4g(int tile, int op, int mode, int disc, const GroupArgs& g, int grid, hipStream_t s) {
    switch (tile) {
        case 0: launch_g<T0, PR>(op, mode, disc, g, grid, s); break;
        case 1: launch_g<T1, PR>(op, mode, disc, g, grid, s); break;
        case 2: launch_g<T2, PR>(op, mode, disc, g, grid, s); break;
        case 3: launch_g<T3, PR>(op, mode, disc, g, grid, s); break;
        default: throw Error(NSLX_E_ARG, "hop: bad tile id");
    }
}

}  // namespace

void launch_hop_grouped(int prec, int op, int mode, int tile, int disc, const GroupArgs& g, int grid, hipStream_t s) {
    NSLX_REQUIRE(g.nprob > 0 && g.ntiles > 0 && grid > 0, "hop: empty grouped launch");
    NSLX_REQUIRE(tile >= 0 && tile < SLICE_BASE, "hop: grouped launches take tiles 0-3");
    ProfScope prof(op == HOP_M ? "grp_M" : "grp_Mdag", prec, mode, s);
    if (prec == NSLX_F64) dispatch4g<PF64, T64_a, T64_b, T64_c, T64_d>(tile, op, mode, disc, g, grid, s);
    else if (prec == NSLX_F32) dispatch4g<PF32, T32_a, T32_b, T32_c, T32_d>(tile, op, mode, disc, g, grid, s);
    else if (prec == NSLX_BF16X3) dispatch4g<PB3, TB3_a, TB3_b, TB3_c, TB3_d>(tile, op, mode, disc, g, grid, s);
    else throw Error(NSLX_E_ARG, "hop: bad precision");
    NSLX_LAUNCH_CHECK();
}

GroupedKernel hop_grouped_kernel(int prec, int op, int mode, int tile, int disc) {
    NSLX_REQUIRE(tile >= 0 && tile < SLICE_BASE, "hop: grouped launches take tiles 0-3");
    if (prec == NSLX_F64) return kernel4g<PF64, T64_a, T64_b, T64_c, T64_d>(tile, op, mode, disc);
    if (prec == NSLX_F32) return kernel4g<PF32, T32_a, T32_b, T32_c, T32_d>(tile, op, mode, disc);
    if (prec == NSLX_BF16X3) return kernel4g<PB3, TB3_a, TB3_b, TB3_c, TB3_d>(tile, op, mode, disc);
    throw Error(NSLX_E_ARG, "hop: bad precision");
}

std::vector<int4> build_tile_list(std::vector<HopProblem>& probs, int prec, int tile) {
    NSLX_REQUIRE(tile >= 0 && tile < SLICE_BASE, "hop: grouped launches take tiles 0-3");
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
    auto build = [](std::vector<TileInfo> m, const SliceEntry* menu, int kc) {
        for (int i = 0; i < SLICE_N; ++i) m.push_back({SLICE_BASE + i, 16, menu[i].fn * 16, kc, 64 * menu[i].w});
        return m;
    };
    static const std::vector<TileInfo> m64 =
        build({info<T64_a>(0), info<T64_b>(1), info<T64_c>(2), info<T64_d>(3)}, SLICE_MENU64, SliceP<PF64>::KC);
    static const std::vector<TileInfo> m32 =
        build({info<T32_a>(0), info<T32_b>(1), info<T32_c>(2), info<T32_d>(3)}, SLICE_MENU32, SliceP<PF32>::KC);
    static const std::vector<TileInfo> mb3 =
        build({info<TB3_a>(0), info<TB3_b>(1), info<TB3_c>(2), info<TB3_d>(3)}, SLICE_MENUB3, SliceP<PB3>::KC);
    if (prec == NSLX_F64) return m64;
    if (prec == NSLX_F32) return m32;
    return mb3;
}

bool tile_fits(int prec, int tile, int nx) {
    if (tile < SLICE_BASE) return true;
    return slice_fits(prec, tile - SLICE_BASE, nx);
}

int hop_nparts(int prec, int tile, int nx, int nt) {
    const TileInfo& ti = tile_menu(prec).at(tile);
    return ((nt + ti.bm - 1) / ti.bm) * ((nx + ti.bn - 1) / ti.bn);
}

int default_tile(int prec, int nx, int nt, int64_t nb) {
    (void)prec;
    if (nx <= 32) return 2;
    // fewer than 512 workgroups with the 32 x 96 tile: the 16 x 32 one
    // (6x the workgroups, and no half-empty row tiles at Nt = 16)
    if (nb * ((nt + 31) / 32) * ((nx + 95) / 96) < 512) return 2;
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
    NSLX_REQUIRE(tile >= 0 && tile < SLICE_BASE + SLICE_N, "hop: bad tile id");
    ProfScope prof(op == HOP_M ? "hop_M" : "hop_Mdag", prec, mode, s);
    NSLX_REQUIRE(a.disc == DISC_EXP || tile < SLICE_BASE, "hop: HubbardDiag runs on the classic tiles 0-3");
    if (tile >= SLICE_BASE) {
        const int id = tile - SLICE_BASE;
        if (prec == NSLX_F64) dispatch_slice<PF64, 1, 2, 3>(id, op, mode, a, s, SLICE_MENU64);
        else if (prec == NSLX_F32) dispatch_slice<PF32, 2, 3, 6>(id, op, mode, a, s, SLICE_MENU32);
        else if (prec == NSLX_BF16X3) dispatch_slice<PB3, 2, 3, 6>(id, op, mode, a, s, SLICE_MENUB3);
        else throw Error(NSLX_E_ARG, "hop: bad precision");
    } else if (prec == NSLX_F64) dispatch4<PF64, T64_a, T64_b, T64_c, T64_d>(tile, op, mode, a, s);
    else if (prec == NSLX_F32) dispatch4<PF32, T32_a, T32_b, T32_c, T32_d>(tile, op, mode, a, s);
    else if (prec == NSLX_BF16X3) dispatch4<PB3, TB3_a, TB3_b, TB3_c, TB3_d>(tile, op, mode, a, s);
    else throw Error(NSLX_E_ARG, "hop: bad precision");
    NSLX_LAUNCH_CHECK();
}

}  // namespace nslx
