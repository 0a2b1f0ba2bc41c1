// Time-slice direct solve of HubbardExp's M and M^dag (algebra in tss.hpp).
//
// Setup per populated field set, in the solver's precision:
//   1. chunk products T_j (c launches of the batched GEMM, every field and
//      chunk at once): row i of Tt_j is T_j e_i, built as A_t applied to the
//      unit vectors slice by slice;
//   2. for k = 0 .. nc-2 the 2Nx-row matrix Q_k = [block rows k, k+1 of R |
//      identity] is formed and its first Nx columns are eliminated by blocked
//      Gauss-Jordan with partial pivoting (per 16-column block: pivot search
//      and LU of the sub-panel in one workgroup per field, row swaps + pivot
//      block inverse per column, rank-16 update of all other rows on MFMA);
//   3. the last diagonal block is inverted the same way, and conjugate
//      transposes of the stored blocks are taken for the R^H solve.
// Apply (M^{-1}): c-1 in-chunk recurrence launches, the reduced forward and
// back sweeps (one launch per block step, every field at once), c launches of
// the in-chunk recurrence from the reduced solution; M^{-dag} mirrors it
// with the backward recurrence, the chunks shifted by one slice, and R^H.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <type_traits>

#include "bgemm_tile.hpp"
#include "engine.hpp"
#include "prof.hpp"
#include "tss.hpp"
#include "vec.hpp"

namespace nslx {
namespace {

template <class S>
struct Z { S re, im; };

template <class S>
__device__ inline Z<S> ldz(const void* p, long long i) {
    if constexpr (sizeof(S) == 8) { const double2 v = reinterpret_cast<const double2*>(p)[i]; return {v.x, v.y}; }
    else { const float2 v = reinterpret_cast<const float2*>(p)[i]; return {v.x, v.y}; }
}
template <class S>
__device__ inline void st(void* p, long long i, Z<S> v) {
    if constexpr (sizeof(S) == 8) reinterpret_cast<double2*>(p)[i] = make_double2(v.re, v.im);
    else reinterpret_cast<float2*>(p)[i] = make_float2(v.re, v.im);
}
template <class S>
__device__ inline Z<S> zmul(Z<S> a, Z<S> b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
template <class S>
__device__ inline Z<S> zsub(Z<S> a, Z<S> b) { return {a.re - b.re, a.im - b.im}; }
template <class S>
__device__ inline Z<S> zinv(Z<S> a) {
    // 1/a with the larger component scaled out (no overflow for |a| near the range ends)
    if (fabs((double)a.re) >= fabs((double)a.im)) {
        const S r = a.im / a.re, d = a.re + a.im * r;
        return {S(1) / d, -r / d};
    }
    const S r = a.re / a.im, d = a.re * r + a.im;
    return {r / d, S(-1) / d};
}

// out[i][j] = (i == j) for an n x n complex matrix (ld n)
template <class S>
__global__ void k_ident(int n, void* out) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    st<S>(out, e, Z<S>{S(i == j ? 1 : 0), S(0)});
}

// Q_k (2n rows, ld W, width Wk = last ? 4n : 5n) of field blockIdx.z:
//   rows [0, n):  [D | 0 | S | I_2n top]   or, last:  [D | S | I_2n top]
//   rows [n, 2n): [-T_{k+1} | I | 0 | I_2n bottom]     [-T_{k+1} | I | I_2n bottom]
// D, S: from the previous step's bottom rows (prevQ + n rows, cols n.. and
// 2n..), or, for k = 0, D = I and S = -T_0.  Tt: row i of T_j is T_j e_i.
template <class S>
__global__ void k_build_q(int n, int k, int last, long long W, const void* Tt, long long strideT, const void* prevQ,
                          long long strideQ, void* Q) {
    const int f = blockIdx.z;
    const int Wk = last ? 4 * n : 5 * n;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2LL * n * Wk) return;
    const int row = (int)(e / Wk), col = (int)(e % Wk);
    const char* tb = reinterpret_cast<const char*>(Tt) + (size_t)f * strideT * sizeof(Z<S>);
    const long long nn = (long long)n * n;
    auto T = [&](int j, int x, int i) {   // T_j[x][i] = Tt_j[i][x]
        return ldz<S>(tb, (long long)j * nn + (long long)i * n + x);
    };
    const char* pq = prevQ ? reinterpret_cast<const char*>(prevQ) + (size_t)f * strideQ * sizeof(Z<S>) : nullptr;
    Z<S> v{S(0), S(0)};
    const int aug = last ? 2 * n : 3 * n;
    if (col >= aug) {
        v.re = S(col - aug == row ? 1 : 0);
    } else if (row < n) {
        const int blk = col / n, x = col % n;
        const bool dcol = blk == 0, scol = last ? blk == 1 : blk == 2;
        if (dcol) {
            if (k == 0) v.re = S(x == row ? 1 : 0);
            else v = ldz<S>(pq, (long long)(n + row) * W + n + x);
        } else if (scol) {
            if (k == 0) { const Z<S> t = T(0, row, x); v = {-t.re, -t.im}; }
            else v = ldz<S>(pq, (long long)(n + row) * W + 2 * n + x);
        }
    } else {
        const int r = row - n, blk = col / n, x = col % n;
        if (blk == 0) { const Z<S> t = T(k + 1, r, x); v = {-t.re, -t.im}; }
        else if (blk == 1) v.re = S(x == r ? 1 : 0);
    }
    st<S>(reinterpret_cast<char*>(Q) + (size_t)f * strideQ * sizeof(Z<S>), (long long)row * W + col, v);
}

// Qfin (n x 2n, ld 2n) = [D | I]: D from the last step's bottom rows (cols n..2n
// of prevQ) or, with one chunk, I - T_0.
template <class S>
__global__ void k_build_fin(int n, long long W, const void* prevQ, long long strideQ, const void* Tt, long long strideT,
                            void* Qf) {
    const int f = blockIdx.z;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2LL * n * n) return;
    const int row = (int)(e / (2 * n)), col = (int)(e % (2 * n));
    Z<S> v{S(0), S(0)};
    if (col >= n) {
        v.re = S(col - n == row ? 1 : 0);
    } else if (prevQ) {
        v = ldz<S>(reinterpret_cast<const char*>(prevQ) + (size_t)f * strideQ * sizeof(Z<S>), (long long)(n + row) * W + n + col);
    } else {
        const Z<S> t = ldz<S>(reinterpret_cast<const char*>(Tt) + (size_t)f * strideT * sizeof(Z<S>), (long long)col * n + row);
        v = {S(row == col ? 1 : 0) - t.re, -t.im};
    }
    st<S>(reinterpret_cast<char*>(Qf) + (size_t)f * 2LL * n * n * sizeof(Z<S>), e, v);
}

constexpr int GJB = 16;   // column block of the Gauss-Jordan elimination
// Per field and column block the pivot record holds 4 x GJB ints: piv[jj] =
// the row swapped with jb + jj, then the block's row map for k_gj_swap
// (gj_row_map).
constexpr int GJREC = 4 * GJB;

// Row map of a column block from its w pivots spiv (LDS), written by threads
// tid < GJB: map[m] = the row whose values end at pivot position jb + m
// (m < w); map[GJB + k], map[2 GJB + k] = destination and source row of a row
// that swap k moved below the pivot positions (-1: none).  The swaps only
// exchange a pivot position with a row at or below it, so a row that ends
// below the pivot positions, if moved, holds an original pivot-position row.
__device__ inline void gj_row_map(int tid, int jb, int w, const int* spiv, int* map) {
    if (tid >= GJB) return;
    auto origin = [&](int r) {
        int x = r;
        for (int k = w - 1; k >= 0; --k) {
            const int j = jb + k, pk = spiv[k];
            x = x == j ? pk : (x == pk ? j : x);
        }
        return x;
    };
    int src = -1, dst = -1, msrc = -1;
    if (tid < w) {
        src = origin(jb + tid);
        const int pk = spiv[tid];
        if (pk >= jb + w) {
            const int o = origin(pk);
            if (o != pk) { dst = pk; msrc = o; }
        }
    }
    map[tid] = src;
    map[GJB + tid] = dst;
    map[2 * GJB + tid] = msrc;
}

// One column block [jb, jb + w) of the Gauss-Jordan elimination of field
// blockIdx.x's matrix (rows x ld, row-major): LU with partial pivoting of the
// sub-panel rows [jb, rows) (the pivot search covers the rows not yet pivotal),
// then Binv = (L_B U_B)^{-1} of the w x w pivot block and C = the sub-panel's
// values before this block at every non-pivot row (rows above jb untouched,
// rows below = L21 U_B), zero at the pivot rows.  piv[jj] = row swapped with
// jb + jj, followed by the row map (gj_row_map).  Only the sub-panel columns
// of Q are modified.
template <class S>
__global__ void __launch_bounds__(256) k_gj_panel(void* Qb, long long ld, long long strideQ, int rows, int jb, int w,
                                                  int* pivb, void* Binvb, void* Cb, int* singular) {
    const int f = blockIdx.x;
    char* Q = reinterpret_cast<char*>(Qb) + (size_t)f * strideQ * sizeof(Z<S>);
    int* piv = pivb + f * GJREC;
    char* Binv = reinterpret_cast<char*>(Binvb) + (size_t)f * GJB * GJB * sizeof(Z<S>);
    char* C = reinterpret_cast<char*>(Cb) + (size_t)f * rows * GJB * sizeof(Z<S>);
    const int tid = threadIdx.x;
    __shared__ double rv[256];
    __shared__ int ri[256];
    __shared__ int s_p;
    __shared__ int s_pv[GJB];
    auto at = [&](int i, int j) { return (long long)i * ld + j; };
    for (int jj = 0; jj < w; ++jj) {
        const int j = jb + jj;
        // pivot: largest |Q[i][j]|^2 over i >= j, lowest index on ties
        double best = -1.0;
        int bi = j;
        for (int i = j + tid; i < rows; i += 256) {
            const Z<S> z = ldz<S>(Q, at(i, j));
            const double a = (double)z.re * z.re + (double)z.im * z.im;
            if (a > best) { best = a; bi = i; }
        }
        rv[tid] = best;
        ri[tid] = bi;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (tid < o) {
                const double a = rv[tid + o];
                const int ia = ri[tid + o];
                if (a > rv[tid] || (a == rv[tid] && ia < ri[tid])) { rv[tid] = a; ri[tid] = ia; }
            }
            __syncthreads();
        }
        if (tid == 0) {
            s_p = ri[0];
            s_pv[jj] = ri[0];
            piv[jj] = ri[0];
            if (!(rv[0] > 0.0)) atomicExch(singular, 1);
        }
        __syncthreads();
        const int p = s_p;
        if (p != j && tid < w) {
            const Z<S> a = ldz<S>(Q, at(j, jb + tid)), b = ldz<S>(Q, at(p, jb + tid));
            st<S>(Q, at(j, jb + tid), b);
            st<S>(Q, at(p, jb + tid), a);
        }
        __syncthreads();
        const Z<S> d = ldz<S>(Q, at(j, j));
        const Z<S> dinv = zinv<S>(d);
        for (int i = j + 1 + tid; i < rows; i += 256) {
            const Z<S> l = zmul<S>(ldz<S>(Q, at(i, j)), dinv);
            st<S>(Q, at(i, j), l);
            for (int cc = jj + 1; cc < w; ++cc) {
                const Z<S> u = ldz<S>(Q, at(j, jb + cc));
                st<S>(Q, at(i, jb + cc), zsub<S>(ldz<S>(Q, at(i, jb + cc)), zmul<S>(l, u)));
            }
        }
        __syncthreads();
    }
    // Binv column c (thread c < w): solve L_B y = e_c, then U_B x = y
    if (tid < w) {
        Z<S> y[GJB];
        for (int i = 0; i < w; ++i) {
            Z<S> s{S(i == tid ? 1 : 0), S(0)};
            for (int m = 0; m < i; ++m) s = zsub<S>(s, zmul<S>(ldz<S>(Q, at(jb + i, jb + m)), y[m]));
            y[i] = s;
        }
        for (int i = w - 1; i >= 0; --i) {
            Z<S> s = y[i];
            for (int m = i + 1; m < w; ++m) s = zsub<S>(s, zmul<S>(ldz<S>(Q, at(jb + i, jb + m)), y[m]));
            y[i] = zmul<S>(s, zinv<S>(ldz<S>(Q, at(jb + i, jb + i))));
        }
        for (int i = 0; i < GJB; ++i) st<S>(Binv, (long long)i * GJB + tid, i < w ? y[i] : Z<S>{S(0), S(0)});
    } else if (tid < GJB) {
        for (int i = 0; i < GJB; ++i) st<S>(Binv, (long long)i * GJB + tid, Z<S>{S(0), S(0)});
    }
    // C rows: above jb the current values, pivot rows 0, below L21 U_B
    for (int i = tid; i < rows; i += 256) {
        for (int cc = 0; cc < GJB; ++cc) {
            Z<S> v{S(0), S(0)};
            if (cc < w) {
                if (i < jb) {
                    v = ldz<S>(Q, at(i, jb + cc));
                } else if (i >= jb + w) {
                    for (int m = 0; m <= cc; ++m)
                        v = {v.re + (ldz<S>(Q, at(i, jb + m)).re * ldz<S>(Q, at(jb + m, jb + cc)).re -
                                     ldz<S>(Q, at(i, jb + m)).im * ldz<S>(Q, at(jb + m, jb + cc)).im),
                             v.im + (ldz<S>(Q, at(i, jb + m)).re * ldz<S>(Q, at(jb + m, jb + cc)).im +
                                     ldz<S>(Q, at(i, jb + m)).im * ldz<S>(Q, at(jb + m, jb + cc)).re)};
                }
            }
            st<S>(C, (long long)i * GJB + cc, v);
        }
    }
    gj_row_map(tid, jb, w, s_pv, piv + GJB);
}

// Pivot candidates of the register kernel as one double each: |a|^2 (capped
// at the largest finite value) with its low 11 mantissa bits replaced by
// 2047 - (row - jb), so that the largest key is the largest |a|^2, lowest
// row among values equal to 2^-41 relative; -1 = no candidate.  A plain
// maximum then reduces value and row together.
constexpr int GJ_KEYBITS = 11;
__device__ inline double gj_key(double v, int rel) {
    const long long b = __double_as_longlong(fmin(v, DBL_MAX));
    return __longlong_as_double((b & ~((1LL << GJ_KEYBITS) - 1)) | ((1 << GJ_KEYBITS) - 1 - rel));
}
__device__ inline int gj_key_row(double k) {
    return (1 << GJ_KEYBITS) - 1 - (int)(__double_as_longlong(k) & ((1LL << GJ_KEYBITS) - 1));
}
__device__ inline double gj_key_value(double k) {
    return __longlong_as_double(__double_as_longlong(k) & ~((1LL << GJ_KEYBITS) - 1));
}
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double readlane_f64(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}
// Maximum over the 64 lanes of the wave, uniform: within each 16-lane row by
// DPP (lane xor 1, xor 2, then the mirrors over 8 and 16 lanes), then the
// four rows' results.
__device__ inline double wave_max(double k) {
    k = fmax(k, dpp_f64<0xB1>(k));    // quad_perm [1, 0, 3, 2]
    k = fmax(k, dpp_f64<0x4E>(k));    // quad_perm [2, 3, 0, 1]
    k = fmax(k, dpp_f64<0x141>(k));   // row_half_mirror
    k = fmax(k, dpp_f64<0x140>(k));   // row_mirror
    return fmax(fmax(readlane_f64(k, 0), readlane_f64(k, 16)), fmax(readlane_f64(k, 32), readlane_f64(k, 48)));
}

// The same column block with the sub-panel rows [jb, rows) held in registers:
// NT threads, thread t owning rows jb + t + m NT (m < RPT), all 16 columns.
// The elimination only chooses the pivots: C and the pivot block B are the
// sub-panel's values before this block, taken in pivot order (A21 = L21 U_B
// and B = L_B U_B hold exactly, so no product of the factors is formed and
// both are one rounding closer than k_gj_panel's).  Every instruction of the
// column loop is issued by every wave, so the loop is kept short:
//   - one barrier per pivot column: each wave reduces its candidates' keys
//     (gj_key) and posts the maximum, its owner lane posts that row, and the
//     owner of row j posts row j; after the barrier every thread takes the
//     maximum of the posts, the winning wave's row as the pivot row, and
//     swaps and updates its own rows.  Posts and rows alternate between two
//     copies by column parity: a thread writes column jj + 2's copy only
//     after the barrier of column jj + 1, which every reader of column jj
//     has passed;
//   - column jj is register column 0 of step jj (the columns shift left after
//     each step), so the loop stays rolled with static register indices, and
//     only the live columns, in bands of 4, are updated and moved (all
//     columns with 512 threads of several rows, whose band copies spill);
//   - the multipliers use conj(p) / |p|^2 (they only steer the pivot search);
//   - no global memory operation inside the loop.
// C is left in the block's columns of Q itself (no later step reads them):
// every row already holds its values before the block, so only the pivot
// rows are zeroed and the rows the swaps moved below the pivot rows are
// rewritten.  Such a row came from a pivot position (a row moved up from
// below stays a pivot row), so its values are the original pivot-position
// rows, kept in LDS.  B's rows are read back from Q before those writes, and
// B's inverse is a Gauss-Jordan sweep of [B | I] in LDS.  Pivots are those of
// k_gj_panel except between values equal to 2^-41 relative; Binv and C agree
// with it to rounding.
template <class S, int RPT, int NT>
__device__ __attribute__((always_inline)) inline void gj_panel_body(int f, void* Qb, long long ld, long long strideQ, int rows, int jb, int w,
                                     int* pivb, void* Binvb, int* singular) {
    static_assert(NT % 64 == 0 && NT >= 64 && (2 * GJB * GJB) % NT == 0, "whole waves, [B | I] split evenly");
    static_assert(RPT * NT <= (1 << GJ_KEYBITS), "relative rows fit the key");
    char* Q = reinterpret_cast<char*>(Qb) + (size_t)f * strideQ * sizeof(Z<S>);
    int* piv = pivb + f * GJREC;
    char* Binv = reinterpret_cast<char*>(Binvb) + (size_t)f * GJB * GJB * sizeof(Z<S>);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int NW = NT / 64;
    __shared__ double wkey[2][NW];
    __shared__ Z<S> wrow[2][NW][GJB];        // each wave's candidate row
    __shared__ Z<S> rowj[2][GJB];
    __shared__ Z<S> orig[GJB][GJB];          // rows [jb, jb + w) before this block
    __shared__ int s_piv[GJB];
    __shared__ Z<S> aug[GJB][2 * GJB + 1];   // [B | I] -> [I | B^-1]
    // the sub-panel (loads clamped in bounds, zero outside it)
    Z<S> a[RPT][GJB];
    const int wl = w - 1;
    #pragma unroll
    for (int m = 0; m < RPT; ++m) {
        const int gi = jb + tid + m * NT;
        const long long rb = (long long)min(gi, rows - 1) * ld + jb;
        #pragma unroll
        for (int c = 0; c < GJB; ++c) a[m][c] = ldz<S>(Q, rb + min(c, wl));
        #pragma unroll
        for (int c = 0; c < GJB; ++c)
            if (gi >= rows || c >= w) a[m][c] = Z<S>{S(0), S(0)};
    }
    if (tid < GJB) {
        #pragma unroll
        for (int c = 0; c < GJB; ++c) orig[tid][c] = a[0][c];
    }
    bool sing = false;
    #pragma unroll 1
    for (int jj = 0; jj < w; ++jj) {
        const int par = jj & 1;
        const int j = jb + jj;
        double key[RPT];
        double best = -1.0;
        #pragma unroll
        for (int m = 0; m < RPT; ++m) {
            const int gi = jb + tid + m * NT;
            const double v = (double)a[m][0].re * a[m][0].re + (double)a[m][0].im * a[m][0].im;
            key[m] = (gi >= j && gi < rows && v == v) ? gj_key(v, gi - jb) : -1.0;
            best = fmax(best, key[m]);
        }
        best = wave_max(best);
        if (lane == 0) wkey[par][wv] = best;
        #pragma unroll
        for (int m = 0; m < RPT; ++m) {
            const int gi = jb + tid + m * NT;
            if (key[m] == best && best >= 0.0) {
                #pragma unroll
                for (int c = 0; c < GJB; ++c) wrow[par][wv][c] = a[m][c];
            }
            if (gi == j) {
                #pragma unroll
                for (int c = 0; c < GJB; ++c) rowj[par][c] = a[m][c];
            }
        }
        __syncthreads();
        double kmax = wkey[par][0];
        #pragma unroll
        for (int k = 1; k < NW; ++k) kmax = fmax(kmax, wkey[par][k]);
        // the pivot row: the winning wave's post, or row j itself when no
        // entry is comparable (all NaN: no swap, the singular flag is set)
        int p;
        const Z<S>* prow;
        if (kmax >= 0.0) {
            const int rel = gj_key_row(kmax);
            p = jb + rel;
            prow = wrow[par][(rel % NT) >> 6];
        } else {
            p = j;
            prow = rowj[par];
        }
        sing = sing || !(kmax >= 0.0 && gj_key_value(kmax) > 0.0);
        if (tid == 0) s_piv[jj] = p;
        #pragma unroll
        for (int m = 0; m < RPT; ++m) {
            const int gi = jb + tid + m * NT;
            if (gi == p && p != j) {
                #pragma unroll
                for (int c = 0; c < GJB; ++c) a[m][c] = rowj[par][c];
            }
            if (gi == j) {
                #pragma unroll
                for (int c = 0; c < GJB; ++c) a[m][c] = prow[c];
            }
        }
        // eliminate below the pivot row, then shift.  The live register
        // columns are [0, w - jj); columns beyond them are stale and never
        // read again.
        const Z<S> pv = prow[0];
        const double rn = 1.0 / ((double)pv.re * pv.re + (double)pv.im * pv.im);
        const Z<S> dinv{S(pv.re * rn), S(-pv.im * rn)};
        auto step = [&](auto ncol) {
            constexpr int NC = decltype(ncol)::value;   // columns 1 .. NC updated, 0 .. NC-1 moved
            Z<S> pr[NC + 1];
            #pragma unroll
            for (int c = 1; c <= NC; ++c) pr[c] = prow[c];
            #pragma unroll
            for (int m = 0; m < RPT; ++m) {
                const int gi = jb + tid + m * NT;
                if (gi >= rows) continue;
                if (gi > j) {
                    const Z<S> l = zmul<S>(a[m][0], dinv);
                    #pragma unroll
                    for (int c = 1; c <= NC; ++c) a[m][c] = zsub<S>(a[m][c], zmul<S>(l, pr[c]));
                }
                #pragma unroll
                for (int c = 0; c < NC; ++c) a[m][c] = a[m][c + 1];
            }
        };
        static_assert(GJB == 16, "bands of 4 columns");
        if constexpr (NT * RPT > 512 && NT > 256) {
            // 512 threads with several rows each: the band copies would spill
            step(std::integral_constant<int, 15>{});
        } else {
            // one uniform branch picks the band's straight-line code
            const int live = w - jj;
            if (live > 12) step(std::integral_constant<int, 15>{});
            else if (live > 8) step(std::integral_constant<int, 11>{});
            else if (live > 4) step(std::integral_constant<int, 7>{});
            else step(std::integral_constant<int, 3>{});
        }
    }
    __syncthreads();
    int pv_[GJB];
    #pragma unroll
    for (int k = 0; k < GJB; ++k) pv_[k] = k < w ? s_piv[k] : jb + k;
    if (tid < w) piv[tid] = s_piv[tid];
    gj_row_map(tid, jb, w, s_piv, piv + GJB);
    if (tid == 0 && sing) atomicExch(singular, 1);
    // row of the original sub-panel now at row r: the swaps undone from the last
    auto origin = [&](int r) {
        int x = r;
        #pragma unroll
        for (int k = GJB - 1; k >= 0; --k) {
            const int j = jb + k, pk = pv_[k];
            x = x == j ? pk : (x == pk ? j : x);
        }
        return x;
    };
    // [B | I]: B = the pivot rows' values before this block, padded with the
    // identity beyond w; element e = tid + q NT of the 16 x 32 augmented matrix
    constexpr int EPT = 2 * GJB * GJB / NT;
    #pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = tid + q * NT, gi_ = e / (2 * GJB), gc_ = e % (2 * GJB);
        Z<S> v{S(0), S(0)};
        if (gc_ < GJB) {
            if (gi_ < w && gc_ < w) v = ldz<S>(Q, (long long)origin(jb + gi_) * ld + jb + gc_);
            else v.re = S(gi_ == gc_ ? 1 : 0);
        } else {
            v.re = S(gc_ - GJB == gi_ ? 1 : 0);
        }
        aug[gi_][gc_] = v;
    }
    // every read of the sub-panel columns of Q (B above) is done after the
    // barrier: they become C in place, zero at the pivot rows, and a row moved
    // below the pivot rows holds an original pivot-position row (every other
    // row already holds its values before the block)
    __syncthreads();
    #pragma unroll
    for (int m = 0; m < RPT; ++m) {
        const int gi = jb + tid + m * NT;
        if (gi >= rows) continue;
        if (gi < jb + w) {
            #pragma unroll
            for (int c = 0; c < GJB; ++c)
                if (c < w) st<S>(Q, (long long)gi * ld + jb + c, Z<S>{S(0), S(0)});
        } else {
            const int o = origin(gi);
            if (o != gi) {
                #pragma unroll
                for (int c = 0; c < GJB; ++c)
                    if (c < w) st<S>(Q, (long long)gi * ld + jb + c, orig[o - jb][c]);
            }
        }
    }
    // Gauss-Jordan on [B | I] without further pivoting: the rows are in
    // partial-pivoting order, so the pivots are those of the LU above
    #pragma unroll 1
    for (int k = 0; k < GJB; ++k) {
        const Z<S> pk = aug[k][k];
        Z<S> fk[EPT], rk[EPT];
        #pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = tid + q * NT, gi_ = e / (2 * GJB), gc_ = e % (2 * GJB);
            fk[q] = aug[gi_][k];
            rk[q] = aug[k][gc_];
        }
        __syncthreads();
        const Z<S> pinv = zinv<S>(pk);
        #pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = tid + q * NT, gi_ = e / (2 * GJB), gc_ = e % (2 * GJB);
            const Z<S> rn = zmul<S>(rk[q], pinv);
            if (gi_ == k) aug[gi_][gc_] = rn;
            else aug[gi_][gc_] = zsub<S>(aug[gi_][gc_], zmul<S>(fk[q], rn));
        }
        __syncthreads();
    }
    #pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = tid + q * NT, gi_ = e / (2 * GJB), gc_ = e % (2 * GJB);
        if (gc_ >= GJB) {
            const int c = gc_ - GJB;
            st<S>(Binv, (long long)gi_ * GJB + c, (gi_ < w && c < w) ? aug[gi_][gc_] : Z<S>{S(0), S(0)});
        }
    }
}

template <class S, int RPT, int NT>
__global__ void __launch_bounds__(NT) k_gj_panel_reg(void* Qb, long long ld, long long strideQ, int rows, int jb,
                                                      int w, int* pivb, void* Binvb, int* singular) {
    gj_panel_body<S, RPT, NT>(blockIdx.x, Qb, ld, strideQ, rows, jb, w, pivb, Binvb, singular);
}

// One launch for block k's trailing update and block k + 1's pivot panel.
// Workgroups f < nf first apply block k's rank-wp update to block k + 1's
// columns [jb, jb + w) in the panel's rows [jb, rows) (multipliers: Q's
// columns [jbp, jbp + wp), pivot-row values: YT rows 0 .. w - 1, since block
// k's update starts at column jbp + wp = jb), then factor the panel
// (gj_panel_body, which reads those columns back after the barrier: the
// workgroup's own writes); the other workgroups are the 64 x 64 tiles of
// block k's update of the columns from jb + w on and of the panel columns'
// rows above jb (the panel reads no row above jb).  The parts write
// disjoint parts of Q, so block k + 1's panel runs beside the update
// instead of after it.
template <class S, int RPT>
__global__ void __launch_bounds__(256) k_gj_fused(const BJob* __restrict__ jobs, const int4* __restrict__ tiles,
                                                  int nf, void* Qb, long long ld, long long strideQ, int rows, int jb,
                                                  int w, int jbp, int wp, const void* YTb, int* pivb, void* Binvb,
                                                  int* singular) {
    if ((int)blockIdx.x >= nf) {
        bg::bgemm_tile<S, true, 0, 64, 64>(jobs, tiles, (int)blockIdx.x - nf);
        return;
    }
    const int f = blockIdx.x;
    char* Q = reinterpret_cast<char*>(Qb) + (size_t)f * strideQ * sizeof(Z<S>);
    const char* YT = reinterpret_cast<const char*>(YTb) + (size_t)f * ld * GJB * sizeof(Z<S>);
    // Rows in chunks of 64: the chunk's multipliers (64 x 16) staged in LDS,
    // the next chunk's loads in flight during this chunk's products.  Thread
    // (rg, c) updates columns jb + c and jb + c + 8 of the chunk's rows rg and
    // rg + 32, with those columns' pivot-row values in registers.  Loads are
    // unconditional (rows clamped; rows beyond are not stored).
    constexpr int PR = 64;
    __shared__ Z<S> sa[PR][GJB + 1];
    const int tid = threadIdx.x;
    const int c = tid % 8, rg = tid / 8;
    Z<S> y0[GJB], y1[GJB];
    #pragma unroll
    for (int m = 0; m < GJB; ++m) {
        y0[m] = (c < w && m < wp) ? ldz<S>(YT, (long long)c * GJB + m) : Z<S>{S(0), S(0)};
        y1[m] = (c + 8 < w && m < wp) ? ldz<S>(YT, (long long)(c + 8) * GJB + m) : Z<S>{S(0), S(0)};
    }
    const int wl = w - 1, wpl = wp - 1;
    Z<S> fa[4], fq[2][2];
    auto fetch = [&](int r0) {
        #pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u, r = min(r0 + e / GJB, rows - 1);
            fa[u] = ldz<S>(Q, (long long)r * ld + jbp + min(e % GJB, wpl));
        }
        #pragma unroll
        for (int i = 0; i < 2; ++i) {
            const long long rb = (long long)min(r0 + rg + 32 * i, rows - 1) * ld + jb;
            fq[i][0] = ldz<S>(Q, rb + min(c, wl));
            fq[i][1] = ldz<S>(Q, rb + min(c + 8, wl));
        }
    };
    fetch(jb);
    for (int r0 = jb; r0 < rows; r0 += PR) {
        #pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            sa[e / GJB][e % GJB] = fa[u];
        }
        Z<S> v[2][2];
        #pragma unroll
        for (int i = 0; i < 2; ++i) {
            v[i][0] = fq[i][0];
            v[i][1] = fq[i][1];
        }
        __syncthreads();
        fetch(r0 + PR);
        #pragma unroll
        for (int m = 0; m < GJB; ++m) {
            if (m < wp) {
                #pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const Z<S> a = sa[rg + 32 * i][m];
                    v[i][0] = zsub<S>(v[i][0], zmul<S>(a, y0[m]));
                    v[i][1] = zsub<S>(v[i][1], zmul<S>(a, y1[m]));
                }
            }
        }
        #pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = r0 + rg + 32 * i;
            if (r >= rows) continue;
            if (c < w) st<S>(Q, (long long)r * ld + jb + c, v[i][0]);
            if (c + 8 < w) st<S>(Q, (long long)r * ld + jb + c + 8, v[i][1]);
        }
        __syncthreads();
    }
    gj_panel_body<S, RPT, 256>(f, Qb, ld, strideQ, rows, jb, w, pivb, Binvb, singular);
}

// Columns [c0, width) of field blockIdx.y: the block's row swaps, then the
// pivot rows become Binv x (also written transposed to YT, ld 16, for the
// rank-w update of the other rows).  The swaps are applied through the row
// map (gj_row_map): every value a column needs is read first, then the moved
// rows and the pivot rows are written, one memory round trip per column.
// GJSQ threads share a column, each reading and writing GJB / GJSQ of its
// rows; the x values meet in LDS.  The row loads are issued before Binv's,
// and the block's barrier separates every read of a column from its writes.
// zf[f * ((ld + 31) / 32) + t] = whether YT has a nonzero in columns
// [c0 + 32 t, c0 + 32 t + 32) (read by the update, BJob::zf).
constexpr int GJS = 64;   // columns per block of k_gj_swap
constexpr int GJSQ = 4;   // threads per column
static_assert(GJS % 32 == 0, "zero flags per 32 columns");
template <class S>
__global__ void __launch_bounds__(GJS * GJSQ) k_gj_swap(void* Qb, long long ld, long long strideQ, int jb, int w,
                                                        int c0, int width, const int* pivb, const void* Binvb,
                                                        void* YTb, int* zf) {
    constexpr int MR = GJB / GJSQ;    // rows per thread
    const int f = blockIdx.y;
    const int tc = threadIdx.x % GJS, mq = threadIdx.x / GJS;
    const int c = c0 + blockIdx.x * GJS + tc;
    const bool live = c < width;
    char* Q = reinterpret_cast<char*>(Qb) + (size_t)f * strideQ * sizeof(Z<S>);
    const int* map = pivb + f * GJREC + GJB;
    const char* Binv = reinterpret_cast<const char*>(Binvb) + (size_t)f * GJB * GJB * sizeof(Z<S>);
    char* YT = reinterpret_cast<char*>(YTb) + (size_t)f * ld * GJB * sizeof(Z<S>);
    __shared__ Z<S> sb[GJB][GJB];
    __shared__ Z<S> sx[GJB][GJS];
    __shared__ int snz[GJS / 32];
    if (threadIdx.x < GJS / 32) snz[threadIdx.x] = 0;
    // all loads to registers first (rows, then Binv), so their latencies overlap
    constexpr int NB = GJB * GJB / (GJS * GJSQ);
    static_assert(NB * GJS * GJSQ == GJB * GJB, "Binv split evenly");
    Z<S> xv[MR], y[MR], bv[NB];
    #pragma unroll
    for (int i = 0; i < MR; ++i) {
        const int m = mq * MR + i;
        xv[i] = (live && m < w) ? ldz<S>(Q, (long long)map[m] * ld + c) : Z<S>{S(0), S(0)};
        y[i] = (live && m < w && map[GJB + m] >= 0) ? ldz<S>(Q, (long long)map[2 * GJB + m] * ld + c)
                                                    : Z<S>{S(0), S(0)};
    }
    #pragma unroll
    for (int i = 0; i < NB; ++i) bv[i] = ldz<S>(Binv, threadIdx.x + i * GJS * GJSQ);
    #pragma unroll
    for (int i = 0; i < MR; ++i) sx[mq * MR + i][tc] = xv[i];
    #pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int e = threadIdx.x + i * GJS * GJSQ;
        sb[e / GJB][e % GJB] = bv[i];
    }
    __syncthreads();
    bool nz = false;
    if (live) {
        #pragma unroll
        for (int i = 0; i < MR; ++i) {
            const int m = mq * MR + i;
            if (m < w && map[GJB + m] >= 0) st<S>(Q, (long long)map[GJB + m] * ld + c, y[i]);
        }
        // Binv is zero outside its w x w block and x beyond w, so the full
        // 16 x 16 product is exact; rows m >= w of YT are written as zeros
        #pragma unroll
        for (int i = 0; i < MR; ++i) {
            const int m = mq * MR + i;
            Z<S> s{S(0), S(0)};
            #pragma unroll
            for (int l = 0; l < GJB; ++l) {
                const Z<S> b = sb[m][l];
                const Z<S> x = sx[l][tc];
                s = {s.re + (b.re * x.re - b.im * x.im), s.im + (b.re * x.im + b.im * x.re)};
            }
            if (m < w) st<S>(Q, (long long)(jb + m) * ld + c, s);
            st<S>(YT, (long long)(c - c0) * GJB + m, s);
            nz = nz || s.re != S(0) || s.im != S(0);    // NaN counts as nonzero
        }
    }
    if (nz) snz[tc / 32] = 1;
    __syncthreads();
    if (threadIdx.x < GJS / 32) {
        const int t = blockIdx.x * (GJS / 32) + threadIdx.x;
        if (32 * t < width - c0) zf[(long long)f * ((ld + 31) / 32) + t] = snz[threadIdx.x];
    }
}

// out[i][j] = conj(in[j][i]), i < rows_out, j < cols_out; field blockIdx.z
template <class S>
__global__ void k_conj_t(const void* in, long long ldin, long long sin, void* out, long long ldout, long long sout,
                         int rows_out, int cols_out) {
    __shared__ Z<S> tile[32][33];
    const int f = blockIdx.z;
    const char* ib = reinterpret_cast<const char*>(in) + (size_t)f * sin * sizeof(Z<S>);
    char* ob = reinterpret_cast<char*>(out) + (size_t)f * sout * sizeof(Z<S>);
    const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 256 threads: 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int j = j0 + r, i = i0 + tx;                     // read in[j][i] along i
        if (j < cols_out && i < rows_out) tile[r][tx] = ldz<S>(ib, (long long)j * ldin + i);
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        if (i < rows_out && j < cols_out) {
            const Z<S> z = tile[tx][r];
            st<S>(ob, (long long)i * ldout + j, Z<S>{z.re, -z.im});
        }
    }
}

inline unsigned blocks(long long n) { return (unsigned)((n + 255) / 256); }

template <class S>
void conj_t(const void* in, long long ldin, long long sin, void* out, long long ldout, long long sout, int rows_out,
            int cols_out, int nf, hipStream_t s) {
    dim3 g((cols_out + 31) / 32, (rows_out + 31) / 32, nf);
    hipLaunchKernelGGL(k_conj_t<S>, g, dim3(256), 0, s, in, ldin, sin, out, ldout, sout, rows_out, cols_out);
    NSLX_LAUNCH_CHECK();
}

}  // namespace

int tss_chunk(int nt, double g, double bound, int cap) {
    int best = 1;
    for (int c = 1; c <= std::min(nt, cap); ++c)
        if (nt % c == 0 && (c == 1 || c * g <= bound)) best = c;
    return best;
}

const void* TimeSliceSolver::E() const { return fm_.Eptr(prec_, nullptr); }
const void* TimeSliceSolver::P() const { return fm_.Pptr(prec_); }

TimeSliceSolver::TimeSliceSolver(FermionMatrix& fm, int prec, int nc, hipStream_t s)
    : fm_(fm), prec_(prec), n_(fm.nx), nt_(fm.nt), nf_((int)fm.nfield) {
    NSLX_REQUIRE(prec == NSLX_F64 || prec == NSLX_F32, "time-slice solve: precision must be F64 or F32");
    NSLX_REQUIRE(fm.disc == DISC_EXP, "time-slice solve: HubbardExp only");
    fm.require_populated();
    esz_ = prec == NSLX_F64 ? 16 : 8;
    if (nc > 0) {
        NSLX_REQUIRE(nt_ % nc == 0, "time-slice solve: nchunks must divide Nt");
        nc_ = nc;
    } else {
        nc_ = nt_ / tss_chunk(nt_, fm.slice_growth());
    }
    c_ = nt_ / nc_;
    const size_t n = n_, nn = n * n;
    Tt_.alloc((size_t)nf_ * nc_ * nn * esz_ * 2);   // chunk products + ping-pong half
    if (nc_ >= 2) Q_.alloc((size_t)nf_ * (nc_ - 1) * (2 * n) * (5 * n) * esz_);
    Qfin_.alloc((size_t)nf_ * n * 2 * n * esz_);
    DinvH_.alloc((size_t)nf_ * nn * esz_);
    if (nc_ >= 2) {
        PhiH_.alloc((size_t)nf_ * (nc_ - 1) * 4 * nn * esz_);
        WHcat_.alloc((size_t)nf_ * n * (nc_ - 1) * n * esz_);
    }
    if (nc_ >= 3) VH_.alloc((size_t)nf_ * (nc_ - 2) * nn * esz_);
    ident_.alloc(nn * esz_);
    if (prec_ == NSLX_F64) hipLaunchKernelGGL(k_ident<double>, dim3(blocks(nn)), dim3(256), 0, s, n_, ident_.p);
    else hipLaunchKernelGGL(k_ident<float>, dim3(blocks(nn)), dim3(256), 0, s, n_, ident_.p);
    NSLX_LAUNCH_CHECK();
    setup(s);
}

size_t TimeSliceSolver::bytes() const {
    size_t b = Tt_.n + Q_.n + Qfin_.n + DinvH_.n + PhiH_.n + WHcat_.n + VH_.n + ident_.n;
    for (auto& kv : plans_) {
        const ApplyBufs& B = *kv.second;
        b += B.mid.n + B.atil.n + B.v[0].n + B.v[1].n + B.z.n + B.q.n + B.cp[0].n + B.cp[1].n + B.sb.n + B.in.n + B.out.n;
    }
    return b;
}

void TimeSliceSolver::gj(BPlan& p, unsigned char* Q, long long ldq, long long strideQ, int rows, int width) {
    const bool f64 = prec_ == NSLX_F64;
    int* piv = gj_piv_.p;
    int* sing = gj_sing_.p;
    int* zf = gj_zf_.p;
    unsigned char* binv = gj_binv_.p;
    unsigned char* C = gj_C_.p;
    unsigned char* YT = gj_YT_.p;
    const int nf = nf_;
    const size_t e = esz_;
    check_extent(Q, (long long)(nf - 1) * strideQ + (long long)(rows - 1) * ldq + width - 1, e, "Gauss-Jordan matrix");
    NSLX_REQUIRE(width <= ldq, "Gauss-Jordan: width exceeds the row stride");
    check_extent(C, (long long)nf * rows * GJB - 1, e, "Gauss-Jordan C");
    check_extent(YT, (long long)nf * ldq * GJB - 1, e, "Gauss-Jordan YT");
    check_extent(binv, (long long)nf * GJB * GJB - 1, e, "Gauss-Jordan Binv");
    check_extent(piv, (long long)nf * GJREC - 1, sizeof(int), "Gauss-Jordan pivot records");
    check_extent(zf, (long long)nf * ((ldq + 31) / 32) - 1, sizeof(int), "Gauss-Jordan zero flags");
    // tile of the rank-16 updates (K = 16, bound by reading and writing the
    // trailing matrix): NSLX_GJ_TILE 0 (64 x 64) / 1 (32 x 32), else by count
    static const int gj_tile = (int)env_double("NSLX_GJ_TILE", -1.0);
    // NSLX_GJ_FUSE=0: every panel and update in launches of their own (read
    // when a plan is built: tests compare both forms)
    const bool gj_fuse = env_double("NSLX_GJ_FUSE", 1.0) != 0.0;
    auto update_tile = [&](const std::vector<BJob>& jobs) {
        if (gj_tile == 0 || gj_tile == 1) return gj_tile;
        std::vector<int4> t;
        bgemm_tiles(jobs, 0, t);
        return t.size() >= 512 ? 0 : 1;
    };
    // a block's update is planned with the next block: in one launch with
    // that block's panel (k_gj_fused, 64 x 64 update tiles: hiding the panel
    // is worth more than the smaller tiles of a small update) where both run
    // register kernels of 256 threads, else on its own
    std::vector<BJob> pend;
    int pend_jb = 0, pend_w = 0;
    bool pend_reg = false;
    auto flush = [&]() {
        if (!pend.empty()) p.gemm(pend, BVariant{prec_, true, 0, 0}, update_tile(pend));
        pend.clear();
    };
    for (int jb = 0; jb < n_; jb += GJB) {
        const int w = std::min(GJB, n_ - jb);
        // register kernel shape for the live sub-panel rows: 256 threads with
        // up to 2 (F64) / 4 (F32) rows each, then 512 threads (register
        // budget), else the global-memory kernel (reg 0)
        const int live = rows - jb;
        int reg = 0;
        if (f64) reg = live <= 256 ? 1 : live <= 512 ? 2 : live <= 1024 ? 3 : 0;
        else reg = live <= 256 ? 1 : live <= 512 ? 2 : live <= 768 ? 3 : live <= 1024 ? 4 : live <= 1536 ? 5 : 0;
        // work of the pivot search and sub-panel LU (8 flops per complex
        // multiply-add): the register kernel reads the sub-panel once and
        // rewrites at most 2w of its rows (C in place); the global-memory one
        // also writes the sub-panel back and forms C = L21 U_B in its buffer
        const double pbytes = reg ? (double)nf * (rows + 2.0 * GJB) * w * e
                                  : (double)nf * rows * (2.0 * w + GJB) * e;
        double pflops = 0.0;
        for (int jj = 0; jj < w; ++jj) pflops += 8.0 * (rows - jb - jj) * (w - jj);
        pflops = (pflops + (reg ? 0.0 : 4.0 * rows * w * w)) * nf;
        // (the fused launch has 256 threads: FP64 sub-panels of up to 768
        // rows take 3 rows per thread there, with the update tile's spare
        // accumulation registers)
        const bool fusable = f64 ? (reg >= 1 && reg <= 2) || (reg == 3 && live <= 3 * 256) : (reg >= 1 && reg <= 4);
        if (gj_fuse && !pend.empty() && pend_reg && fusable && gj_tile != 1) {
            // block k's update without this panel's columns (those are
            // updated by the panel's workgroups)
            std::vector<BJob> rest = pend;
            for (BJob& j : rest) {
                j.Y = static_cast<char*>(j.Y) + (size_t)w * e;
                j.Add = j.Y;
                j.B = static_cast<const char*>(j.B) + (size_t)w * GJB * e;
                j.cols -= w;
                j.zfoff = w;
            }
            // and the panel columns' rows above the panel (block k's own
            // update job restricted to them)
            if (jb > 0)
                for (int f = 0; f < nf; ++f) {
                    BJob j = pend[f];
                    j.rows = jb;
                    j.cols = w;
                    rest.push_back(j);
                }
            const int jbp = pend_jb, wp = pend_w;
            double ufl = 0.0, uby = 0.0;
            bgemm_work(rest, BVariant{prec_, true, 0, 0}, &ufl, &uby);
            // the panel columns' update of the panel's rows: read (and write)
            // them, read the multipliers, 8 flops per complex multiply-add
            const double ffl = pflops + ufl + 8.0 * nf * (rows - jb) * w * wp;
            const double fby = pbytes + uby + (double)nf * (rows - jb) * (2.0 * w + wp) * e;
            p.gemm_with(rest, BVariant{prec_, true, 0, 0}, 0,
                        [=](hipStream_t s, const BJob* dj, const int4* dt, int ntl, const BVariant&) {
                ProfScope prof(prof_on() ? std::string(f64 ? "gj_fused_f64" : "gj_fused_f32") : std::string(), s);
                prof.work(ffl, fby);
                void* Qv = (void*)Q;
                const void* yt = (const void*)YT;
                void* bv = (void*)binv;
                const dim3 g((unsigned)(nf + ntl));
                if (f64) {
                    if (reg == 3)
                        launch_timed(k_gj_fused<double, 3>, g, dim3(256), 0, s, dj, dt, nf, Qv, ldq, strideQ, rows,
                                     jb, w, jbp, wp, yt, piv, bv, sing);
                    else if (reg == 1)
                        launch_timed(k_gj_fused<double, 1>, g, dim3(256), 0, s, dj, dt, nf, Qv, ldq, strideQ, rows,
                                     jb, w, jbp, wp, yt, piv, bv, sing);
                    else
                        launch_timed(k_gj_fused<double, 2>, g, dim3(256), 0, s, dj, dt, nf, Qv, ldq, strideQ, rows,
                                     jb, w, jbp, wp, yt, piv, bv, sing);
                } else {
                    if (reg == 1)
                        launch_timed(k_gj_fused<float, 1>, g, dim3(256), 0, s, dj, dt, nf, Qv, ldq, strideQ, rows, jb,
                                     w, jbp, wp, yt, piv, bv, sing);
                    else if (reg == 2)
                        launch_timed(k_gj_fused<float, 2>, g, dim3(256), 0, s, dj, dt, nf, Qv, ldq, strideQ, rows, jb,
                                     w, jbp, wp, yt, piv, bv, sing);
                    else if (reg == 3)
                        launch_timed(k_gj_fused<float, 3>, g, dim3(256), 0, s, dj, dt, nf, Qv, ldq, strideQ, rows, jb,
                                     w, jbp, wp, yt, piv, bv, sing);
                    else
                        launch_timed(k_gj_fused<float, 4>, g, dim3(256), 0, s, dj, dt, nf, Qv, ldq, strideQ, rows, jb,
                                     w, jbp, wp, yt, piv, bv, sing);
                }
                NSLX_LAUNCH_CHECK();
            });
            pend.clear();
        } else {
        flush();
        p.custom([=](hipStream_t s) {
            ProfScope prof(prof_on() ? std::string(f64 ? "gj_panel_f64" : "gj_panel_f32") : std::string(), s);
            prof.work(pflops, pbytes);
            void* Qv = (void*)Q;
            void* bv = (void*)binv;
            void* Cv = (void*)C;
            if (f64) {
                if (reg == 1) launch_timed(k_gj_panel_reg<double, 1, 256>, dim3(nf), dim3(256), 0, s, Qv, ldq, strideQ, rows, jb, w, piv, bv, sing);
                else if (reg == 2) launch_timed(k_gj_panel_reg<double, 2, 256>, dim3(nf), dim3(256), 0, s, Qv, ldq, strideQ, rows, jb, w, piv, bv, sing);
                else if (reg == 3) launch_timed(k_gj_panel_reg<double, 2, 512>, dim3(nf), dim3(512), 0, s, Qv, ldq, strideQ, rows, jb, w, piv, bv, sing);
                else launch_timed(k_gj_panel<double>, dim3(nf), dim3(256), 0, s, Qv, ldq, strideQ, rows, jb, w, piv, bv, Cv, sing);
            } else {
                if (reg == 1) launch_timed(k_gj_panel_reg<float, 1, 256>, dim3(nf), dim3(256), 0, s, Qv, ldq, strideQ, rows, jb, w, piv, bv, sing);
                else if (reg == 2) launch_timed(k_gj_panel_reg<float, 2, 256>, dim3(nf), dim3(256), 0, s, Qv, ldq, strideQ, rows, jb, w, piv, bv, sing);
                else if (reg == 3) launch_timed(k_gj_panel_reg<float, 3, 256>, dim3(nf), dim3(256), 0, s, Qv, ldq, strideQ, rows, jb, w, piv, bv, sing);
                else if (reg == 4) launch_timed(k_gj_panel_reg<float, 4, 256>, dim3(nf), dim3(256), 0, s, Qv, ldq, strideQ, rows, jb, w, piv, bv, sing);
                else if (reg == 5) launch_timed(k_gj_panel_reg<float, 3, 512>, dim3(nf), dim3(512), 0, s, Qv, ldq, strideQ, rows, jb, w, piv, bv, sing);
                else launch_timed(k_gj_panel<float>, dim3(nf), dim3(256), 0, s, Qv, ldq, strideQ, rows, jb, w, piv, bv, Cv, sing);
            }
            NSLX_LAUNCH_CHECK();
        });
        }
        const int c0 = jb + w;
        const int ncol = width - c0;
        // the pivot rows' sources read and the pivot rows written (w each),
        // YT written (16), the moved rows read and written (at most w each,
        // counted in full)
        const double sbytes = (double)nf * ncol * (4.0 * w + GJB) * e;
        const double sflops = (double)nf * ncol * 8.0 * GJB * GJB;
        p.custom([=](hipStream_t s) {
            ProfScope prof(prof_on() ? std::string(f64 ? "gj_swap_f64" : "gj_swap_f32") : std::string(), s);
            prof.work(sflops, sbytes);
            dim3 g((ncol + GJS - 1) / GJS, nf);
            if (f64)
                launch_timed(k_gj_swap<double>, g, dim3(GJS * GJSQ), 0, s, (void*)Q, ldq, strideQ, jb, w, c0,
                             width, (const int*)piv, (const void*)binv, (void*)YT, zf);
            else
                launch_timed(k_gj_swap<float>, g, dim3(GJS * GJSQ), 0, s, (void*)Q, ldq, strideQ, jb, w, c0,
                             width, (const int*)piv, (const void*)binv, (void*)YT, zf);
            NSLX_LAUNCH_CHECK();
        });
        std::vector<BJob> jobs(nf);
        for (int f = 0; f < nf; ++f) {
            BJob& j = jobs[f];
            // C: the block's columns of Q (register kernel) or the C buffer
            j.A0 = reg ? static_cast<const void*>(Q + ((size_t)f * strideQ + jb) * e)
                       : static_cast<const void*>(C + (size_t)f * rows * GJB * e);
            j.lda0 = reg ? ldq : GJB;
            j.rows = rows;
            j.cols = ncol;
            j.K = w;
            j.B = YT + (size_t)f * ldq * GJB * e;
            j.ldb = GJB;
            j.Y = Q + ((size_t)f * strideQ + c0) * e;
            j.ldy = ldq;
            j.Add = j.Y;
            j.ldadd = ldq;
            j.sgn = -1.0;
            // columns whose pivot-row values are all zero (the augmented
            // identity's columns of rows not yet pivots) are left untouched
            j.zf = zf + (size_t)f * ((ldq + 31) / 32);
        }
        pend = std::move(jobs);
        pend_jb = jb;
        pend_w = w;
        pend_reg = reg != 0;
    }
    flush();
}

void TimeSliceSolver::refactor(hipStream_t s) {
    fm_.require_populated();
    NSLX_REQUIRE(fm_.nfield == nf_, "time-slice solve: the field count changed");
    // the plans hold E and phase-table addresses: rebuilt if either moved
    if (E() != e_used_ || P() != p_used_) {
        plans_.clear();
        setup_plan_ = BPlan();
        setup_built_ = false;
    }
    setup(s);
}

void TimeSliceSolver::setup(hipStream_t s) {
    if (!setup_built_) build_setup(s);
    NSLX_HIP(hipMemsetAsync(gj_sing_.p, 0, sizeof(int), s));
    setup_plan_.run(s);
    // the singular flag goes to pinned memory without a wait: the caller
    // checks it at its next synchronisation (check_singular)
    sing_h_.ensure(1);
    NSLX_HIP(hipMemcpyAsync(sing_h_.p, gj_sing_.p, sizeof(int), hipMemcpyDeviceToHost, s));
    sing_pending_ = true;
    gen_ = fm_.pop_gen;
}

void TimeSliceSolver::check_singular(hipStream_t s, bool synced) {
    if (!sing_pending_) return;
    if (!synced) NSLX_HIP(hipStreamSynchronize(s));
    sing_pending_ = false;
    if (sing_h_.p[0]) throw Error(NSLX_E_STATE, "time-slice solve: the reduced block system is singular");
}

void TimeSliceSolver::build_setup(hipStream_t s) {
    const int n = n_, nf = nf_, nc = nc_, c = c_, nt = nt_;
    const long long nn = (long long)n * n;
    const bool f64 = prec_ == NSLX_F64;
    const void* Ep = fm_.Eptr(prec_, s);
    const long long ldE = e_ld(n);
    const void* Pp = P();
    e_used_ = Ep;
    p_used_ = Pp;
    BPlan& p = setup_plan_;
    p = BPlan();
    const long long W = 5LL * n;
    gj_piv_.alloc((size_t)nf * GJREC);
    gj_sing_.alloc(1);
    gj_zf_.alloc((size_t)nf * ((W + 31) / 32));
    gj_binv_.alloc((size_t)nf * GJB * GJB * esz_);
    gj_C_.alloc((size_t)nf * 2 * n * GJB * esz_);
    gj_YT_.alloc((size_t)nf * W * GJB * esz_);
    // 1. chunk products: one launch of row blocks that keep their rows for
    // all c steps, where they fit; else c batched GEMM launches, step q
    // writing buffer (c-1-q) & 1 (0 = Tt, 1 = ping-pong half)
    if (chunkprod_supported(prec_, n)) {
        unsigned char* Tt0 = Tt_.p;
        const int prec = prec_;
        p.custom([=](hipStream_t st) { launch_chunkprod(prec, n, nt, c, nc, nf, Ep, ldE, Pp, Tt0, st); });
    } else {
        unsigned char* buf[2] = {Tt_.p, Tt_.p + (size_t)nf * nc * nn * esz_};
        for (int q = 0; q < c; ++q) {
            unsigned char* dst = buf[(c - 1 - q) & 1];
            unsigned char* src = buf[(c - q) & 1];
            std::vector<BJob> jobs;
            for (int f = 0; f < nf; ++f)
                for (int j = 0; j < nc; ++j) {
                    const int t = j * c + q;
                    BJob b;
                    const size_t off = ((size_t)f * nc + j) * nn * esz_;
                    b.A0 = q == 0 ? (const void*)ident_.p : (const void*)(src + off);
                    b.lda0 = n;
                    b.rows = n;
                    b.cols = n;
                    b.K = n;
                    b.B = Ep;
                    b.ldb = ldE;
                    b.Y = dst + off;
                    b.ldy = n;
                    b.ph = static_cast<const char*>(Pp) + ((size_t)f * nt + (t - 1 + nt) % nt) * n * esz_;
                    b.ldph = 0;
                    b.sgn = t == 0 ? -1.0 : 1.0;
                    jobs.push_back(b);
                }
            p.gemm(jobs, BVariant{prec_, false, 1, 0});
        }
    }
    const long long strideQ = 2LL * n * W;                 // per (field, step) in Q_
    const long long strideT = (long long)nc * nn;          // per field in Tt
    const long long fstride = (long long)(nc - 1) * strideQ;   // between fields in Q_
    unsigned char* Tt = Tt_.p;
    // 2. block steps (Q_ holds [field][step][2n][5n])
    for (int k = 0; k + 1 < nc; ++k) {
        const bool last = k == nc - 2;
        unsigned char* Qk = Q_.p + (size_t)k * strideQ * esz_;
        const unsigned char* prev = k == 0 ? nullptr : Q_.p + (size_t)(k - 1) * strideQ * esz_;
        const int Wk = last ? 4 * n : 5 * n;
        check_extent(Qk, (long long)(nf - 1) * fstride + (2LL * n - 1) * W + Wk - 1, esz_, "Q_k");
        if (prev) check_extent(prev, (long long)(nf - 1) * fstride + (2LL * n - 1) * W + 3LL * n - 1, esz_, "Q_{k-1}");
        check_extent(Tt, (long long)(nf - 1) * strideT + (long long)nc * nn - 1, esz_, "chunk products");
        p.custom([=](hipStream_t st) {
            dim3 g(blocks(2LL * n * Wk), 1, nf);
            if (f64)
                hipLaunchKernelGGL(k_build_q<double>, g, dim3(256), 0, st, n, k, last ? 1 : 0, W, Tt, strideT, prev,
                                   fstride, Qk);
            else
                hipLaunchKernelGGL(k_build_q<float>, g, dim3(256), 0, st, n, k, last ? 1 : 0, W, Tt, strideT, prev, fstride,
                                   Qk);
            NSLX_LAUNCH_CHECK();
        });
        gj(p, Qk, W, fstride, 2 * n, Wk);
    }
    {
        const unsigned char* prev = nc >= 2 ? Q_.p + (size_t)(nc - 2) * strideQ * esz_ : nullptr;
        unsigned char* Qf = Qfin_.p;
        check_extent(Qf, (long long)nf * 2 * nn - 1, esz_, "Qfin");
        if (prev) check_extent(prev, (long long)(nf - 1) * fstride + (2LL * n - 1) * W + 2LL * n - 1, esz_, "Q_{nc-2}");
        p.custom([=](hipStream_t st) {
            dim3 g(blocks(2LL * nn), 1, nf);
            if (f64)
                hipLaunchKernelGGL(k_build_fin<double>, g, dim3(256), 0, st, n, W, prev, fstride, Tt, strideT, Qf);
            else
                hipLaunchKernelGGL(k_build_fin<float>, g, dim3(256), 0, st, n, W, prev, fstride, Tt, strideT, Qf);
            NSLX_LAUNCH_CHECK();
        });
        gj(p, Qfin_.p, 2 * n, 2LL * nn, n, 2 * n);
    }
    // 3. conjugate transposes for the R^H solve
    auto ct = [&](const void* in, long long ldin, long long sin, void* out, long long ldout, long long sout, int ro, int co) {
        check_extent(in, (long long)(nf - 1) * sin + (long long)(co - 1) * ldin + ro - 1, esz_, "conj-transpose input");
        check_extent(out, (long long)(nf - 1) * sout + (long long)(ro - 1) * ldout + co - 1, esz_, "conj-transpose output");
        p.custom([=](hipStream_t st) {
            if (f64) conj_t<double>(in, ldin, sin, out, ldout, sout, ro, co, nf, st);
            else conj_t<float>(in, ldin, sin, out, ldout, sout, ro, co, nf, st);
        });
    };
    ct(Qfin_.p + (size_t)n * esz_, 2 * n, 2LL * nn, DinvH_.p, n, nn, n, n);
    if (nc >= 2) {
        const long long wcat = (long long)(nc - 1) * n;
        for (int k = 0; k + 1 < nc; ++k) {
            const bool last = k == nc - 2;
            const unsigned char* Qk = Q_.p + (size_t)k * strideQ * esz_;
            const long long aug = last ? 2LL * n : 3LL * n;
            ct(Qk + (size_t)aug * esz_, W, fstride, PhiH_.p + (size_t)k * 4 * nn * esz_, 2 * n, (long long)(nc - 1) * 4 * nn,
               2 * n, 2 * n);
            if (last) {
                // WHcat block 0 = V_{nc-2}^H
                ct(Qk + (size_t)n * esz_, W, fstride, WHcat_.p, wcat, (long long)n * wcat, n, n);
            } else {
                // VH_k = V_k^H;  WHcat block k + 1 = W_k^H
                ct(Qk + (size_t)n * esz_, W, fstride, VH_.p + (size_t)k * nn * esz_, n, (long long)(nc - 2) * nn, n, n);
                ct(Qk + (size_t)2 * n * esz_, W, fstride, WHcat_.p + (size_t)(k + 1) * n * esz_, wcat, (long long)n * wcat,
                   n, n);
            }
        }
    }
    p.upload(s);
    setup_launches_ = p.launches();
    setup_built_ = true;
}

TimeSliceSolver::ApplyBufs& TimeSliceSolver::bufs(int64_t nb, hipStream_t s) {
    if (plans_chains_ != fm_.chains_ok) {   // rebuilt for the other launch form
        plans_.clear();
        plans_chains_ = fm_.chains_ok;
    }
    auto it = plans_.find(nb);
    if (it != plans_.end()) return *it->second;
    NSLX_REQUIRE(nb >= 1 && nb % nf_ == 0, "time-slice solve: nbatch must be a positive multiple of the fields");
    auto B = std::make_unique<ApplyBufs>();
    const size_t vec = (size_t)nb * nt_ * n_ * esz_, red = (size_t)nb * nc_ * n_ * esz_, one = (size_t)nb * n_ * esz_;
    B->in.alloc(vec); B->mid.alloc(vec); B->out.alloc(vec);
    B->atil.alloc(red); B->z.alloc(red); B->q.alloc(red);
    B->v[0].alloc(one); B->v[1].alloc(one); B->cp[0].alloc(one); B->cp[1].alloc(one); B->sb.alloc(one);
    build_minv(B->Minv, B->in.p, B->mid.p, *B, nb);
    build_mdinv(B->Mdinv, B->mid.p, B->out.p, *B, nb);
    B->Minv.upload(s);
    B->Mdinv.upload(s);
    auto& ref = *B;
    plans_.emplace(nb, std::move(B));
    return ref;
}

void TimeSliceSolver::build_minv(BPlan& p, const void* y, void* x, ApplyBufs& B, int64_t nb) {
    const int n = n_, nt = nt_, nc = nc_, c = c_;
    const int rpf = (int)(nb / nf_);
    const long long lv = (long long)nt * n;      // row stride of a field vector
    const long long lr = (long long)nc * n;      // row stride of a reduced vector
    const size_t e = esz_;
    const char* Y = static_cast<const char*>(y);
    char* X = static_cast<char*>(x);
    const void* Ep = E();
    const long long ldE = e_ld(n);
    const char* Pp = static_cast<const char*>(P());
    auto pass_job = [&](int t, const void* a, long long lda) {
        BJob b;
        b.A0 = a;
        b.lda0 = lda;
        b.rows = (int)nb;
        b.cols = n;
        b.K = n;
        b.B = Ep;
        b.ldb = ldE;
        b.Y = X + (size_t)t * n * e;
        b.ldy = lv;
        b.Add = Y + (size_t)t * n * e;
        b.ldadd = lv;
        b.ph = Pp + (size_t)((t - 1 + nt) % nt) * n * e;
        b.ldph = lv;
        b.phdiv = rpf;
        b.sgn = t == 0 ? -1.0 : 1.0;
        return b;
    };
    // pass 1: in-chunk recurrence from zero inflow
    {
        std::vector<std::vector<BJob>> steps;
        for (int q = 1; q < c; ++q) {
            std::vector<BJob> jobs;
            for (int j = 0; j < nc; ++j) {
                const int t = j * c + q;
                jobs.push_back(pass_job(t, (q == 1 ? Y : X) + (size_t)(t - 1) * n * e, lv));
            }
            steps.push_back(std::move(jobs));
        }
        p.chain(steps, BVariant{prec_, false, 1, 0}, fm_.chains_ok);
    }
    const char* wsrc = c == 1 ? Y : X;
    auto w_at = [&](int j, int f) { return wsrc + ((size_t)f * rpf * lv + (size_t)(j * c + c - 1) * n) * e; };
    char* Z_ = reinterpret_cast<char*>(B.z.p);
    char* At = reinterpret_cast<char*>(B.atil.p);
    auto zrow = [&](char* base, int f, int slot) { return base + ((size_t)f * rpf * lr + (size_t)slot * n) * e; };
    auto vrow = [&](DevBuf<unsigned char>& b, int f) { return reinterpret_cast<char*>(b.p) + (size_t)f * rpf * n * e; };
    const long long W = 5LL * n, strideQ = 2LL * n * W, fstride = (long long)(nc - 1) * strideQ;
    auto Qk = [&](int f, int k) { return reinterpret_cast<const char*>(Q_.p) + ((size_t)f * fstride + (size_t)k * strideQ) * e; };
    auto Qf = [&](int f) { return reinterpret_cast<const char*>(Qfin_.p) + (size_t)f * 2 * n * n * e; };
    auto red_job = [&](int f) {
        BJob b;
        b.rows = rpf;
        b.cols = n;
        return b;
    };
    if (nc == 1) {
        std::vector<BJob> jobs;
        for (int f = 0; f < nf_; ++f) {
            BJob b = red_job(f);
            b.A0 = w_at(0, f); b.lda0 = lv; b.K = n;
            b.B = Qf(f) + (size_t)n * e; b.ldb = 2 * n;
            b.Y = zrow(Z_, f, 0); b.ldy = lr;
            jobs.push_back(b);
        }
        p.gemm(jobs, BVariant{prec_, true, 0, 1});
    } else {
        for (int k = 0; k + 1 < nc; ++k) {
            const bool last = k == nc - 2;
            const long long aug = last ? 2LL * n : 3LL * n;
            std::vector<BJob> jobs;
            for (int f = 0; f < nf_; ++f)
                for (int part = 0; part < 2; ++part) {
                    BJob b = red_job(f);
                    if (k == 0) { b.A0 = w_at(0, f); b.lda0 = lv; }
                    else { b.A0 = vrow(B.v[k & 1], f); b.lda0 = n; }
                    b.A1 = w_at(k + 1, f); b.lda1 = lv; b.kseg = n; b.K = 2 * n;
                    b.B = Qk(f, k) + ((size_t)part * n * W + aug) * e; b.ldb = W;
                    if (part == 0) { b.Y = zrow(At, f, k); b.ldy = lr; }
                    else { b.Y = vrow(B.v[(k + 1) & 1], f); b.ldy = n; }
                    jobs.push_back(b);
                }
            p.gemm(jobs, BVariant{prec_, true, 0, 1});
        }
        {
            std::vector<BJob> jobs;
            for (int f = 0; f < nf_; ++f) {
                BJob b = red_job(f);
                b.A0 = vrow(B.v[(nc - 1) & 1], f); b.lda0 = n; b.K = n;
                b.B = Qf(f) + (size_t)n * e; b.ldb = 2 * n;
                b.Y = zrow(Z_, f, nc - 1); b.ldy = lr;
                jobs.push_back(b);
            }
            p.gemm(jobs, BVariant{prec_, true, 0, 1});
        }
        for (int k = nc - 2; k >= 0; --k) {
            std::vector<BJob> jobs;
            for (int f = 0; f < nf_; ++f) {
                BJob b = red_job(f);
                b.A0 = zrow(Z_, f, k + 1); b.lda0 = lr;
                b.A1 = zrow(Z_, f, nc - 1); b.lda1 = lr; b.kseg = n;
                b.K = k == nc - 2 ? n : 2 * n;
                b.B = Qk(f, k) + (size_t)n * e; b.ldb = W;
                b.Add = zrow(At, f, k); b.ldadd = lr;
                b.Y = zrow(Z_, f, k); b.ldy = lr;
                b.sgn = -1.0;
                jobs.push_back(b);
            }
            p.gemm(jobs, BVariant{prec_, true, 0, 1});
        }
    }
    // pass 2: in-chunk recurrence from the reduced solution
    {
        std::vector<std::vector<BJob>> steps;
        for (int q = 0; q < c; ++q) {
            std::vector<BJob> jobs;
            for (int j = 0; j < nc; ++j) {
                const int t = j * c + q;
                if (q == 0) jobs.push_back(pass_job(t, Z_ + (size_t)((j - 1 + nc) % nc) * n * e, lr));
                else jobs.push_back(pass_job(t, X + (size_t)(t - 1) * n * e, lv));
            }
            steps.push_back(std::move(jobs));
        }
        p.chain(steps, BVariant{prec_, false, 1, 0}, fm_.chains_ok);
    }
}

void TimeSliceSolver::build_mdinv(BPlan& p, const void* y, void* x, ApplyBufs& B, int64_t nb) {
    const int n = n_, nt = nt_, nc = nc_, c = c_;
    const int rpf = (int)(nb / nf_);
    const long long lv = (long long)nt * n, lr = (long long)nc * n;
    const size_t e = esz_;
    const char* Y = static_cast<const char*>(y);
    char* X = static_cast<char*>(x);
    const void* Ep = E();
    const long long ldE = e_ld(n);
    const char* Pp = static_cast<const char*>(P());
    auto first = [&](int j) { return (j * c - 1 + nt) % nt; };     // f_j
    auto lastix = [&](int j) { return (j * c + c - 2 + nt) % nt; };  // e_j
    auto pass_job = [&](int s, const void* a, long long lda) {
        BJob b;
        b.A0 = a;
        b.lda0 = lda;
        b.rows = (int)nb;
        b.cols = n;
        b.K = n;
        b.B = Ep;
        b.ldb = ldE;
        b.Y = X + (size_t)s * n * e;
        b.ldy = lv;
        b.Add = Y + (size_t)s * n * e;
        b.ldadd = lv;
        b.ph = Pp + (size_t)s * n * e;
        b.ldph = lv;
        b.phdiv = rpf;
        b.sgn = s == nt - 1 ? -1.0 : 1.0;
        return b;
    };
    {
        std::vector<std::vector<BJob>> steps;
        for (int m = 1; m < c; ++m) {
            std::vector<BJob> jobs;
            for (int j = 0; j < nc; ++j) {
                const int s = (lastix(j) - m + nt) % nt;
                jobs.push_back(pass_job(s, (m == 1 ? Y : X) + (size_t)((s + 1) % nt) * n * e, lv));
            }
            steps.push_back(std::move(jobs));
        }
        p.chain(steps, BVariant{prec_, false, 2, 0}, fm_.chains_ok);
    }
    const char* wsrc = c == 1 ? Y : X;
    // what_k = w'_{k+1}, at slot f_{(k+1) mod nc}
    auto wh_at = [&](int k, int f) { return wsrc + ((size_t)f * rpf * lv + (size_t)first((k + 1) % nc) * n) * e; };
    char* Z_ = reinterpret_cast<char*>(B.z.p);
    char* Qb = reinterpret_cast<char*>(B.q.p);
    auto zrow = [&](char* base, int f, int slot) { return base + ((size_t)f * rpf * lr + (size_t)slot * n) * e; };
    auto vrow = [&](DevBuf<unsigned char>& b, int f) { return reinterpret_cast<char*>(b.p) + (size_t)f * rpf * n * e; };
    const long long nn = (long long)n * n;
    auto DH = [&](int f) { return reinterpret_cast<const char*>(DinvH_.p) + (size_t)f * nn * e; };
    auto red_job = [&]() {
        BJob b;
        b.rows = rpf;
        b.cols = n;
        return b;
    };
    if (nc == 1) {
        std::vector<BJob> jobs;
        for (int f = 0; f < nf_; ++f) {
            BJob b = red_job();
            b.A0 = wh_at(0, f); b.lda0 = lv; b.K = n;
            b.B = DH(f); b.ldb = n;
            b.Y = zrow(Z_, f, 0); b.ldy = lr;
            jobs.push_back(b);
        }
        p.gemm(jobs, BVariant{prec_, true, 0, 1});
    } else {
        // q_0 = what_0 (a strided copy into the q array)
        {
            const char* src = wh_at(0, 0);
            char* dst = Qb;
            const size_t spitch = (size_t)lv * e, dpitch = (size_t)lr * e, width = (size_t)n * e;
            const size_t height = (size_t)nb;
            check_extent(src, (long long)((height - 1) * spitch + width) - 1, 1, "q_0 copy source");
            check_extent(dst, (long long)((height - 1) * dpitch + width) - 1, 1, "q_0 copy destination");
            p.custom([=](hipStream_t st) {
                NSLX_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToDevice, st));
            });
        }
        for (int k = 1; k + 1 < nc; ++k) {
            std::vector<BJob> jobs;
            for (int f = 0; f < nf_; ++f) {
                BJob b = red_job();
                b.A0 = zrow(Qb, f, k - 1); b.lda0 = lr; b.K = n;
                b.B = reinterpret_cast<const char*>(VH_.p) + ((size_t)f * (nc - 2) + (k - 1)) * nn * e; b.ldb = n;
                b.Add = wh_at(k, f); b.ldadd = lv;
                b.Y = zrow(Qb, f, k); b.ldy = lr;
                b.sgn = -1.0;
                jobs.push_back(b);
            }
            p.gemm(jobs, BVariant{prec_, true, 0, 1});
        }
        const long long wcat = (long long)(nc - 1) * n;
        {
            std::vector<BJob> jobs;
            for (int f = 0; f < nf_; ++f) {
                BJob b = red_job();
                b.A0 = zrow(Qb, f, nc - 2); b.lda0 = lr;
                b.A1 = zrow(Qb, f, 0); b.lda1 = lr; b.kseg = n;
                b.K = (int)wcat;
                b.B = reinterpret_cast<const char*>(WHcat_.p) + (size_t)f * n * wcat * e; b.ldb = wcat;
                b.Add = wh_at(nc - 1, f); b.ldadd = lv;
                b.Y = vrow(B.sb, f); b.ldy = n;
                b.sgn = -1.0;
                jobs.push_back(b);
            }
            p.gemm(jobs, BVariant{prec_, true, 0, 1});
        }
        {
            std::vector<BJob> jobs;
            for (int f = 0; f < nf_; ++f) {
                BJob b = red_job();
                b.A0 = vrow(B.sb, f); b.lda0 = n; b.K = n;
                b.B = DH(f); b.ldb = n;
                b.Y = zrow(Qb, f, nc - 1); b.ldy = lr;
                jobs.push_back(b);
            }
            p.gemm(jobs, BVariant{prec_, true, 0, 1});
        }
        for (int k = nc - 2; k >= 0; --k) {
            std::vector<BJob> jobs;
            for (int f = 0; f < nf_; ++f)
                for (int part = 0; part < 2; ++part) {
                    BJob b = red_job();
                    b.A0 = zrow(Qb, f, k); b.lda0 = lr;
                    if (k == nc - 2) { b.A1 = zrow(Qb, f, nc - 1); b.lda1 = lr; }
                    else { b.A1 = vrow(B.cp[(k + 1) & 1], f); b.lda1 = n; }
                    b.kseg = n;
                    b.K = 2 * n;
                    b.B = reinterpret_cast<const char*>(PhiH_.p) +
                          (((size_t)f * (nc - 1) + k) * 4 * nn + (size_t)part * n * 2 * n) * e;
                    b.ldb = 2 * n;
                    if (part == 0) {
                        if (k == 0) { b.Y = zrow(Z_, f, 0); b.ldy = lr; }
                        else { b.Y = vrow(B.cp[k & 1], f); b.ldy = n; }
                    } else {
                        b.Y = zrow(Z_, f, k + 1); b.ldy = lr;
                    }
                    jobs.push_back(b);
                }
            p.gemm(jobs, BVariant{prec_, true, 0, 1});
        }
    }
    // pass 2: backward recurrence from zhat_j = z'_{j+1}
    {
        std::vector<std::vector<BJob>> steps;
        for (int m = 0; m < c; ++m) {
            std::vector<BJob> jobs;
            for (int j = 0; j < nc; ++j) {
                const int s = (lastix(j) - m + nt) % nt;
                if (m == 0) jobs.push_back(pass_job(s, Z_ + (size_t)j * n * e, lr));
                else jobs.push_back(pass_job(s, X + (size_t)((s + 1) % nt) * n * e, lv));
            }
            steps.push_back(std::move(jobs));
        }
        p.chain(steps, BVariant{prec_, false, 2, 0}, fm_.chains_ok);
    }
}

void TimeSliceSolver::queue_pass_checks(hipStream_t s) {
    pass_plans_.clear();
    for (auto& kv : plans_) {
        ApplyBufs& B = *kv.second;
        for (BPlan* p : {&B.Minv, &B.Mdinv}) pass_plans_.push_back(p);
        for (auto& q : B.bound)
            for (BPlan* p : {&q->first, &q->second}) pass_plans_.push_back(p);
    }
    pass_plans_.erase(std::remove_if(pass_plans_.begin(), pass_plans_.end(),
                                     [](BPlan* p) { return p->chain_err() == nullptr; }),
                      pass_plans_.end());
    pass_h_.ensure(std::max<size_t>(pass_plans_.size(), 1));
    for (size_t i = 0; i < pass_plans_.size(); ++i)
        NSLX_HIP(hipMemcpyAsync(pass_h_.p + i, pass_plans_[i]->chain_err(), sizeof(int), hipMemcpyDeviceToHost, s));
}

void TimeSliceSolver::check_passes_synced(hipStream_t s) {
    bool failed = false;
    for (size_t i = 0; i < pass_plans_.size(); ++i)
        if (pass_h_.p[i]) {
            pass_plans_[i]->chain_reset(s);
            failed = true;
        }
    pass_plans_.clear();
    if (failed)
        throw Error(NSLX_E_INTERNAL,
                    "time-slice solve: a recurrence launch stopped waiting for its workgroups (results invalid; "
                    "set NSLX_CHAINS=0 for one launch per slice)");
}

void TimeSliceSolver::check_passes(hipStream_t s) {
    bool failed = false;
    for (auto& kv : plans_) {
        failed = kv.second->Minv.chain_failed(s) || failed;
        failed = kv.second->Mdinv.chain_failed(s) || failed;
        for (auto& q : kv.second->bound) {
            failed = q->first.chain_failed(s) || failed;
            failed = q->second.chain_failed(s) || failed;
        }
    }
    if (failed)
        throw Error(NSLX_E_INTERNAL,
                    "time-slice solve: a recurrence launch stopped waiting for its workgroups (results invalid; "
                    "set NSLX_CHAINS=0 for one launch per slice)");
}

void TimeSliceSolver::apply(int op, const void* y, void* x, int64_t nb, bool io_c128, hipStream_t s) {
    NSLX_REQUIRE(op == NSLX_OP_M || op == NSLX_OP_MDAGGER || op == NSLX_OP_MMDAGGER || op == NSLX_OP_MDAGGERM,
                 "time-slice solve: bad op");
    NSLX_REQUIRE(y && x && y != x, "time-slice solve: null or aliased buffers");
    NSLX_REQUIRE(gen_ == fm_.pop_gen, "time-slice solve: the fermion matrix was re-populated since the factorisation");
    ApplyBufs& B = bufs(nb, s);
    const int64_t n = nb * (int64_t)nt_ * n_;
    const size_t bytes = (size_t)n * esz_;
    const bool conv = io_c128 && prec_ == NSLX_F32;
    auto in = [&](void* dst) {
        if (conv) c128_to_c64(y, dst, n, s);
        else NSLX_HIP(hipMemcpyAsync(dst, y, bytes, hipMemcpyDeviceToDevice, s));
    };
    auto out = [&](const void* src) {
        if (conv) c64_to_c128(src, x, n, s);
        else NSLX_HIP(hipMemcpyAsync(x, src, bytes, hipMemcpyDeviceToDevice, s));
    };
    // a product solve between caller buffers (no conversion) runs plans bound
    // to them: the caller's y is only read, x only holds this solve's
    // recurrences and result.  A few pairs are kept (the solvers of this
    // library call with one fixed pair); others go through the copies.
    constexpr size_t MAX_BOUND = 4;
    const bool product = op == NSLX_OP_MMDAGGER || op == NSLX_OP_MDAGGERM;
    auto ours = [&](const void* p) {
        return p == B.in.p || p == B.mid.p || p == B.out.p;
    };
    if (product && !conv && !ours(y) && !ours(x)) {
        BoundPlans* bp = nullptr;
        for (auto& q : B.bound)
            if (q->op == op && q->y == y && q->x == x) bp = q.get();
        if (!bp && B.bound.size() < MAX_BOUND) {
            auto q = std::make_unique<BoundPlans>();
            q->op = op;
            q->y = y;
            q->x = x;
            if (op == NSLX_OP_MMDAGGER) {       // (M M^dag)^{-1} = M^{-dag} M^{-1}
                build_minv(q->first, y, B.mid.p, B, nb);
                build_mdinv(q->second, B.mid.p, x, B, nb);
            } else {                            // (M^dag M)^{-1} = M^{-1} M^{-dag}
                build_mdinv(q->first, y, B.mid.p, B, nb);
                build_minv(q->second, B.mid.p, x, B, nb);
            }
            q->first.upload(s);
            q->second.upload(s);
            bp = q.get();
            B.bound.push_back(std::move(q));
        }
        if (bp) {
            bp->first.run(s);
            bp->second.run(s);
            return;
        }
    }
    if (op == NSLX_OP_M) {
        in(B.in.p);
        B.Minv.run(s);
        out(B.mid.p);
    } else if (op == NSLX_OP_MDAGGER) {
        in(B.mid.p);
        B.Mdinv.run(s);
        out(B.out.p);
    } else if (op == NSLX_OP_MMDAGGER) {       // (M M^dag)^{-1} = M^{-dag} M^{-1}
        in(B.in.p);
        B.Minv.run(s);
        B.Mdinv.run(s);
        out(B.out.p);
    } else {                                   // (M^dag M)^{-1} = M^{-1} M^{-dag}
        in(B.mid.p);
        B.Mdinv.run(s);
        NSLX_HIP(hipMemcpyAsync(B.in.p, B.out.p, bytes, hipMemcpyDeviceToDevice, s));
        B.Minv.run(s);
        out(B.mid.p);
    }
}

}  // namespace nslx
