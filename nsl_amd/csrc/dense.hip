// Dense FP64 kernels (see dense.hpp).  These serve E = exp(delta K) (once per
// exponent), log det M and the force; they work on Nx x Nx matrices where
// the hop kernels work on the (nbatch*Nt) x Nx field block.
#include <algorithm>

#include "dense.hpp"

#include <cmath>

namespace nslx {
namespace {

constexpr int GT = 64;   // output tile
constexpr int GK = 16;   // k chunk

__device__ inline double2 cm(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

template <bool AC, bool BC>
__global__ void __launch_bounds__(256) k_gemm(int n, const void* A, const void* B, void* C, const double2* dl,
                                              const double2* dm, const double2* dr, double alpha, double gid) {
    __shared__ double sAr[GK][GT + 1], sAi[GK][GT + 1];
    __shared__ double sBr[GK][GT + 1], sBi[GK][GT + 1];
    const int tid = threadIdx.x;
    const int tx = tid % 16, ty = tid / 16;
    const int r0 = blockIdx.y * GT, c0 = blockIdx.x * GT;
    double cr[4][4], ci[4][4];
    #pragma unroll
    for (int i = 0; i < 4; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j) { cr[i][j] = 0.0; ci[i][j] = 0.0; }

    for (int k0 = 0; k0 < n; k0 += GK) {
        // A tile: rows r0..r0+63, cols k0..k0+15  (scaled by dm on k)
        #pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = e * 256 + tid;
            const int i = idx / GK, k = idx % GK;
            const int gi = r0 + i, gk = k0 + k;
            double2 v = make_double2(0.0, 0.0);
            if (gi < n && gk < n) {
                if (AC) v = reinterpret_cast<const double2*>(A)[(size_t)gi * n + gk];
                else v.x = reinterpret_cast<const double*>(A)[(size_t)gi * n + gk];
                if (dm) v = cm(v, dm[gk]);
            }
            sAr[k][i] = v.x;
            sAi[k][i] = v.y;
        }
        // B tile: rows k0..k0+15, cols c0..c0+63
        #pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = e * 256 + tid;
            const int k = idx / GT, j = idx % GT;
            const int gk = k0 + k, gj = c0 + j;
            double2 v = make_double2(0.0, 0.0);
            if (gk < n && gj < n) {
                if (BC) v = reinterpret_cast<const double2*>(B)[(size_t)gk * n + gj];
                else v.x = reinterpret_cast<const double*>(B)[(size_t)gk * n + gj];
            }
            sBr[k][j] = v.x;
            sBi[k][j] = v.y;
        }
        __syncthreads();
        #pragma unroll
        for (int k = 0; k < GK; ++k) {
            double ar[4], ai[4], br[4], bi[4];
            #pragma unroll
            for (int i = 0; i < 4; ++i) { ar[i] = sAr[k][ty * 4 + i]; ai[i] = sAi[k][ty * 4 + i]; }
            #pragma unroll
            for (int j = 0; j < 4; ++j) { br[j] = sBr[k][tx * 4 + j]; bi[j] = sBi[k][tx * 4 + j]; }
            #pragma unroll
            for (int i = 0; i < 4; ++i)
                #pragma unroll
                for (int j = 0; j < 4; ++j) {
                    cr[i][j] = fma(ar[i], br[j], cr[i][j]);
                    if (AC && BC) {
                        cr[i][j] = fma(-ai[i], bi[j], cr[i][j]);
                        ci[i][j] = fma(ar[i], bi[j], ci[i][j]);
                        ci[i][j] = fma(ai[i], br[j], ci[i][j]);
                    } else if (AC) {
                        ci[i][j] = fma(ai[i], br[j], ci[i][j]);
                    } else if (BC) {
                        ci[i][j] = fma(ar[i], bi[j], ci[i][j]);
                    }
                }
        }
        __syncthreads();
    }
    #pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = r0 + ty * 4 + i;
        if (gi >= n) continue;
        #pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = c0 + tx * 4 + j;
            if (gj >= n) continue;
            double2 v = make_double2(alpha * cr[i][j], alpha * ci[i][j]);
            if (dl) v = cm(dl[gi], v);
            if (dr) v = cm(v, dr[gj]);
            if (gi == gj) v.x += gid;
            if (AC || BC) reinterpret_cast<double2*>(C)[(size_t)gi * n + gj] = v;
            else reinterpret_cast<double*>(C)[(size_t)gi * n + gj] = v.x;
        }
    }
}

__global__ void k_set_identity_r(int n, double* A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (int64_t)n * n) A[i] = (i / n == i % n) ? 1.0 : 0.0;
}
__global__ void k_scale_r(int64_t m, const double* in, double a, double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = a * in[i];
}
__global__ void k_copy_r(int64_t m, const double* in, double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = in[i];
}
__global__ void k_scaled_plus_identity(int n, const double* K, double a, double id, double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (int64_t)n * n) out[i] = a * K[i] + ((i / n == i % n) ? id : 0.0);
}

// ---- LU with partial pivoting, one launch per column ----
// One LU step of the log determinant, from Ain into Aout (so that no
// workgroup reads what another writes): every workgroup finds the pivot row p
// of column k itself (largest modulus among rows >= k, ties to the lowest
// row, no swap when none is found); workgroup 0 adds log|pivot| and its phase
// (+ pi for a row swap) to acc; the workgroups write the trailing rows and
// columns > k: row i (row k where i = p) less l_i times row p, with l_i its
// column-k element over the pivot.  Rows and columns <= k are not written:
// no later step reads them, and the determinant needs only the pivots.
constexpr int LU_PER = 8;   // elements per thread
__global__ void __launch_bounds__(256) k_lu_step(int n, int k, const double2* __restrict__ Ain,
                                                 double2* __restrict__ Aout, double* acc) {
    __shared__ double bv[256];
    __shared__ int bi[256];
    const int tid = threadIdx.x;
    double best = -1.0;
    int besti = n;
    for (int i = k + tid; i < n; i += 256) {
        const double2 v = Ain[(size_t)i * n + k];
        const double mm = hypot(v.x, v.y);
        if (mm > best || (mm == best && i < besti)) { best = mm; besti = i; }
    }
    bv[tid] = best;
    bi[tid] = besti;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            const double o = bv[tid + s];
            const int oi = bi[tid + s];
            if (o > bv[tid] || (o == bv[tid] && oi < bi[tid])) { bv[tid] = o; bi[tid] = oi; }
        }
        __syncthreads();
    }
    const int p = bi[0] < n ? bi[0] : k;
    const double2 piv = Ain[(size_t)p * n + k];
    if (blockIdx.x == 0 && tid == 0) {
        acc[0] += log(hypot(piv.x, piv.y));
        acc[1] += atan2(piv.y, piv.x) + ((p != k) ? M_PI : 0.0);
    }
    const double den = piv.x * piv.x + piv.y * piv.y;
    const double2 inv = make_double2(piv.x / den, -piv.y / den);
    const int m = n - k - 1;
    const int64_t total = (int64_t)m * m;
    #pragma unroll
    for (int e = 0; e < LU_PER; ++e) {
        const int64_t idx = ((int64_t)blockIdx.x * LU_PER + e) * 256 + tid;
        if (idx >= total) break;
        const int i = k + 1 + (int)(idx / m), j = k + 1 + (int)(idx % m);
        const int src = (i == p) ? k : i;
        const double2 l = cm(Ain[(size_t)src * n + k], inv), u = Ain[(size_t)p * n + j];
        double2 a = Ain[(size_t)src * n + j];
        a.x -= l.x * u.x - l.y * u.y;
        a.y -= l.x * u.y + l.y * u.x;
        Aout[(size_t)i * n + j] = a;
    }
}

// ---- Gauss-Jordan on the augmented n x 2n matrix ----
// One Gauss-Jordan step on the augmented n x 2n matrix, from Min into Mout
// (so that no workgroup reads what another writes): every workgroup finds the
// pivot row p of column k itself (largest modulus among rows >= k, ties to
// the lowest row, no swap when none is found), then writes its share of
// columns j >= k (the columns left of k are unit columns that no later step
// reads): row k = row p / pivot, and every other row i, taken from row k
// where i = p, less its column-k element times the new row k.
constexpr int GJ_PER = 8;   // elements per thread
__global__ void __launch_bounds__(256) k_gj_step(int n, int k, const double2* __restrict__ Min,
                                                 double2* __restrict__ Mout) {
    __shared__ double bv[256];
    __shared__ int bi[256];
    const int tid = threadIdx.x;
    const int w = 2 * n;
    double best = -1.0;
    int besti = n;
    for (int i = k + tid; i < n; i += 256) {
        const double2 v = Min[(size_t)i * w + k];
        const double mm = hypot(v.x, v.y);
        if (mm > best || (mm == best && i < besti)) { best = mm; besti = i; }
    }
    bv[tid] = best;
    bi[tid] = besti;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            const double o = bv[tid + s];
            const int oi = bi[tid + s];
            if (o > bv[tid] || (o == bv[tid] && oi < bi[tid])) { bv[tid] = o; bi[tid] = oi; }
        }
        __syncthreads();
    }
    const int p = bi[0] < n ? bi[0] : k;
    const double2 piv = Min[(size_t)p * w + k];
    const double den = piv.x * piv.x + piv.y * piv.y;
    const double2 inv = make_double2(piv.x / den, -piv.y / den);
    const int wc = w - k;
    const int64_t total = (int64_t)n * wc;
    #pragma unroll
    for (int e = 0; e < GJ_PER; ++e) {
        const int64_t idx = ((int64_t)blockIdx.x * GJ_PER + e) * 256 + tid;
        if (idx >= total) break;
        const int i = (int)(idx / wc), j = k + (int)(idx % wc);
        const double2 u = cm(Min[(size_t)p * w + j], inv);
        double2 out = u;
        if (i != k) {
            const int src = (i == p) ? k : i;
            const double2 f = Min[(size_t)src * w + k];
            out = Min[(size_t)src * w + j];
            out.x -= f.x * u.x - f.y * u.y;
            out.y -= f.x * u.y + f.y * u.x;
        }
        Mout[(size_t)i * w + j] = out;
    }
}

__global__ void k_aug_init(int n, const double2* A, double2* M) {
    const int w = 2 * n;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * w) return;
    const int i = (int)(idx / w), j = (int)(idx % w);
    M[idx] = (j < n) ? A[(size_t)i * n + j] : make_double2((j - n == i) ? 1.0 : 0.0, 0.0);
}
__global__ void k_aug_extract(int n, const double2* M, double2* A) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * n) return;
    const int i = (int)(idx / n), j = (int)(idx % n);
    A[idx] = M[(size_t)i * 2 * n + n + j];
}

__global__ void k_diag_iprod(int n, const double2* Q, const double2* G, double2* d) {
    __shared__ double sr[256], si[256];
    const int x = blockIdx.x, tid = threadIdx.x;
    double r = 0.0, im = 0.0;
    for (int k = tid; k < n; k += 256) {
        const double2 a = Q[(size_t)x * n + k], b = G[(size_t)k * n + x];
        r += a.x * b.x - a.y * b.y;
        im += a.x * b.y + a.y * b.x;
    }
    sr[tid] = r;
    si[tid] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) { sr[tid] += sr[tid + s]; si[tid] += si[tid + s]; }
        __syncthreads();
    }
    if (tid == 0) d[x] = make_double2(-si[0], sr[0]);   // i * (re + i im)
}

__global__ void k_build_cyclic(int n, int nblk, const double2* C, double2* R) {
    const int N = n * nblk;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)N * N) return;
    const int I = (int)(idx / N), J = (int)(idx % N);
    const int bi = I / n, bj = J / n, i = I % n, j = J % n;
    double2 v = make_double2((I == J) ? 1.0 : 0.0, 0.0);
    if (nblk == 1) {
        const double2 c = C[(size_t)i * n + j];
        v.x += c.x; v.y += c.y;
    } else if (bi == 0 && bj == nblk - 1) {
        const double2 c = C[(size_t)i * n + j];
        v.x += c.x; v.y += c.y;
    } else if (bi >= 1 && bj == bi - 1) {
        const double2 c = C[(size_t)bi * n * n + (size_t)i * n + j];
        v.x -= c.x; v.y -= c.y;
    }
    R[idx] = v;
}

__global__ void k_set_identity_c(int n, double2* A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (int64_t)n * n) A[i] = make_double2((i / n == i % n) ? 1.0 : 0.0, 0.0);
}
__global__ void k_real_to_complex(int64_t m, const double* in, double2* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = make_double2(in[i], 0.0);
}
__global__ void k_add_identity_c(int n, double2* A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[(size_t)i * n + i].x += 1.0;
}

__global__ void k_dscale_rows(int n, const double2* dl, const double* E, double2* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * n) return;
    const double2 d = dl[i / n];
    const double e = E[i];
    out[i] = make_double2(d.x * e, d.y * e);
}

__global__ void k_conj(int64_t m, const double2* in, double2* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = make_double2(in[i].x, -in[i].y);
}

__global__ void k_dscale_cols(int n, const double* E, const double2* dr, double2* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * n) return;
    const double2 d = dr[i % n];
    const double e = E[i];
    out[i] = make_double2(d.x * e, d.y * e);
}

__global__ void k_transpose_c(int n, const double2* A, double2* AT) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * n) return;
    const int r = (int)(i / n), c = (int)(i % n);
    AT[(size_t)c * n + r] = A[i];
}

// d_b[x] = i * sum_k QT_b[k][x] G_b[k][x] for the nb matrix pairs b (QT_b =
// QT + b * stride, G_b likewise, d_b = d + b * dstride): a workgroup per 16
// columns of one pair, its 16 thread rows taking every 16th k (consecutive
// threads read consecutive elements of a row), the 16 partial sums of a
// column added in a fixed order
constexpr int DIP = 16;
__global__ void __launch_bounds__(256) k_diag_iprod_cols(int n, const double2* __restrict__ QT,
                                                         const double2* __restrict__ G, int64_t stride,
                                                         double2* __restrict__ d, int64_t dstride) {
    __shared__ double2 part[DIP][DIP + 1];
    const int c = threadIdx.x % DIP, ks = threadIdx.x / DIP;
    const int x = blockIdx.x * DIP + c;
    const int64_t off = (int64_t)blockIdx.y * stride;
    double r = 0.0, im = 0.0;
    if (x < n) {
        #pragma unroll 4
        for (int k = ks; k < n; k += DIP) {
            const double2 a = QT[off + (int64_t)k * n + x], b = G[off + (int64_t)k * n + x];
            r += a.x * b.x - a.y * b.y;
            im += a.x * b.y + a.y * b.x;
        }
    }
    part[ks][c] = make_double2(r, im);
    __syncthreads();
    if (ks == 0 && x < n) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < DIP; ++j) {
            sr += part[j][c].x;
            si += part[j][c].y;
        }
        d[(int64_t)blockIdx.y * dstride + x] = make_double2(-si, sr);
    }
}

inline unsigned g1(int64_t m) { return (unsigned)((m + 255) / 256); }

}  // namespace

void conj_c(int64_t m, const void* in, void* out, hipStream_t s) {
    hipLaunchKernelGGL(k_conj, dim3(g1(m)), dim3(256), 0, s, m, (const double2*)in, (double2*)out);
    NSLX_LAUNCH_CHECK();
}

void dscale_cols(int n, const double* E, const void* dr, void* out, hipStream_t s) {
    hipLaunchKernelGGL(k_dscale_cols, dim3(g1((int64_t)n * n)), dim3(256), 0, s, n, E, (const double2*)dr,
                       (double2*)out);
    NSLX_LAUNCH_CHECK();
}

void transpose_c(int n, const void* A, void* AT, hipStream_t s) {
    hipLaunchKernelGGL(k_transpose_c, dim3(g1((int64_t)n * n)), dim3(256), 0, s, n, (const double2*)A, (double2*)AT);
    NSLX_LAUNCH_CHECK();
}

void diag_iprod_cols(int n, const void* QT, const void* G, void* d, hipStream_t s) {
    diag_iprod_cols_batched(n, 1, QT, G, 0, d, 0, s);
}

void diag_iprod_cols_batched(int n, int nb, const void* QT, const void* G, int64_t stride, void* d, int64_t dstride,
                             hipStream_t s) {
    if (n <= 0 || nb <= 0) return;
    hipLaunchKernelGGL(k_diag_iprod_cols, dim3((n + DIP - 1) / DIP, nb), dim3(DIP * DIP), 0, s, n,
                       (const double2*)QT, (const double2*)G, stride, (double2*)d, dstride);
    NSLX_LAUNCH_CHECK();
}

void dscale_rows(int n, const void* dl, const double* E, void* out, hipStream_t s) {
    hipLaunchKernelGGL(k_dscale_rows, dim3(g1((int64_t)n * n)), dim3(256), 0, s, n, (const double2*)dl, E, (double2*)out);
    NSLX_LAUNCH_CHECK();
}

void gemm_sq(int n, const void* A, int a_cplx, const void* B, int b_cplx, void* C, const void* dl, const void* dm,
             const void* dr, double alpha, double gid, hipStream_t s) {
    dim3 grid((n + GT - 1) / GT, (n + GT - 1) / GT);
    const double2* l = (const double2*)dl;
    const double2* m = (const double2*)dm;
    const double2* r = (const double2*)dr;
    if (a_cplx && b_cplx) hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(256), 0, s, n, A, B, C, l, m, r, alpha, gid);
    else if (a_cplx) hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(256), 0, s, n, A, B, C, l, m, r, alpha, gid);
    else if (b_cplx) hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(256), 0, s, n, A, B, C, l, m, r, alpha, gid);
    else hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(256), 0, s, n, A, B, C, l, m, r, alpha, gid);
    NSLX_LAUNCH_CHECK();
}

// Taylor(18) of A' = (a / 2^sq) K by Horner, then sq squarings.
void expm_real_scaled(int n, const double* K, double a, int sq, double* E, double* w1, double* w2, hipStream_t s) {
    const int64_t m = (int64_t)n * n;
    const double ap = std::ldexp(a, -sq);
    hipLaunchKernelGGL(k_scale_r, dim3(g1(m)), dim3(256), 0, s, m, K, ap, w1);  // w1 = A'
    NSLX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_set_identity_r, dim3(g1(m)), dim3(256), 0, s, n, E);   // T = I
    NSLX_LAUNCH_CHECK();
    const int deg = 18;
    for (int k = deg; k >= 1; --k) {
        // T = I + (A'/k) T
        gemm_sq(n, w1, 0, E, 0, w2, nullptr, nullptr, nullptr, 1.0 / k, 1.0, s);
        hipLaunchKernelGGL(k_copy_r, dim3(g1(m)), dim3(256), 0, s, m, w2, E);
        NSLX_LAUNCH_CHECK();
    }
    for (int i = 0; i < sq; ++i) {
        gemm_sq(n, E, 0, E, 0, w2, nullptr, nullptr, nullptr, 1.0, 0.0, s);
        hipLaunchKernelGGL(k_copy_r, dim3(g1(m)), dim3(256), 0, s, m, w2, E);
        NSLX_LAUNCH_CHECK();
    }
}

void lu_logdet(int n, void* A, void* work, double* acc, hipStream_t s) {
    double2* buf[2] = {(double2*)A, (double2*)work};
    for (int k = 0; k < n; ++k) {
        const int64_t m = n - k - 1;
        const unsigned blocks = (unsigned)std::max<int64_t>(1, (m * m + 256 * LU_PER - 1) / (256 * LU_PER));
        hipLaunchKernelGGL(k_lu_step, dim3(blocks), dim3(256), 0, s, n, k, buf[k & 1], buf[(k + 1) & 1], acc);
    }
    NSLX_LAUNCH_CHECK();
}

void gj_inverse(int n, void* A, void* aug, void* work, hipStream_t s) {
    const int64_t m2 = (int64_t)n * 2 * n;
    hipLaunchKernelGGL(k_aug_init, dim3(g1(m2)), dim3(256), 0, s, n, (const double2*)A, (double2*)aug);
    double2* buf[2] = {(double2*)aug, (double2*)work};
    for (int k = 0; k < n; ++k) {
        const int64_t total = (int64_t)n * (2 * n - k);
        const unsigned blocks = (unsigned)((total + 256 * GJ_PER - 1) / (256 * GJ_PER));
        hipLaunchKernelGGL(k_gj_step, dim3(blocks), dim3(256), 0, s, n, k, buf[k & 1], buf[(k + 1) & 1]);
    }
    hipLaunchKernelGGL(k_aug_extract, dim3(g1((int64_t)n * n)), dim3(256), 0, s, n, (const double2*)buf[n & 1],
                       (double2*)A);
    NSLX_LAUNCH_CHECK();
}

void diag_iprod(int n, const void* Q, const void* G, void* d, hipStream_t s) {
    hipLaunchKernelGGL(k_diag_iprod, dim3(n), dim3(256), 0, s, n, (const double2*)Q, (const double2*)G, (double2*)d);
    NSLX_LAUNCH_CHECK();
}

void build_cyclic(int n, int nblk, const void* C, void* R, hipStream_t s) {
    const int64_t N = (int64_t)n * nblk;
    hipLaunchKernelGGL(k_build_cyclic, dim3(g1(N * N)), dim3(256), 0, s, n, nblk, (const double2*)C, (double2*)R);
    NSLX_LAUNCH_CHECK();
}

void set_identity_c(int n, void* A, hipStream_t s) {
    hipLaunchKernelGGL(k_set_identity_c, dim3(g1((int64_t)n * n)), dim3(256), 0, s, n, (double2*)A);
    NSLX_LAUNCH_CHECK();
}
void real_to_complex(int64_t m, const double* in, void* out, hipStream_t s) {
    hipLaunchKernelGGL(k_real_to_complex, dim3(g1(m)), dim3(256), 0, s, m, in, (double2*)out);
    NSLX_LAUNCH_CHECK();
}
void add_identity_c(int n, void* A, hipStream_t s) {
    hipLaunchKernelGGL(k_add_identity_c, dim3(g1(n)), dim3(256), 0, s, n, (double2*)A);
    NSLX_LAUNCH_CHECK();
}
void scaled_plus_identity(int n, const double* K, double a, double id, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_scaled_plus_identity, dim3(g1((int64_t)n * n)), dim3(256), 0, s, n, K, a, id, out);
    NSLX_LAUNCH_CHECK();
}

}  // namespace nslx
