// Dense FP64 matrix kernels for the setup / log det / force side of the
// path: square n x n row-major matrices, complex (c128) or real (double).
#pragma once

#include "common.hpp"

namespace nslx {

// C = alpha * diag(dl) A diag(dm) B diag(dr) + gid * I.
// a_cplx / b_cplx select complex (c128) or real (double) operands; C is
// complex unless both are real.  dl/dm/dr: complex vectors or nullptr.
// C must not alias A or B.
void gemm_sq(int n, const void* A, int a_cplx, const void* B, int b_cplx, void* C, const void* dl,
             const void* dm, const void* dr, double alpha, double gid, hipStream_t s);

// E = exp(a K) for real K (n x n, row-major, device): Taylor polynomial of
// degree 18 of A' = (a / 2^sq) K by Horner, then sq squarings.  The caller
// picks sq so that ||A'|| <= 1/2 (truncation < 1e-22 relative).  w1, w2:
// n*n scratch.
void expm_real_scaled(int n, const double* K, double a, int sq, double* E, double* w1, double* w2,
                      hipStream_t s);

// log det of a complex n x n matrix (destroyed) by LU with partial pivoting:
// acc[0] += sum log|u_kk|, acc[1] += sum arg(u_kk) + pi * (#swaps).  work: a
// second n x n c128 buffer (the steps alternate between A and work).
void lu_logdet(int n, void* A, void* work, double* acc, hipStream_t s);

// In-place inverse of a complex n x n matrix by Gauss-Jordan with partial
// pivoting, one launch per column.  aug and work: n x 2n c128 entries each
// (the steps alternate between them).
void gj_inverse(int n, void* A, void* aug, void* work, hipStream_t s);

// d[x] = i * sum_k Q[x][k] G[k][x]   (diagonal of i*Q*G), complex.
void diag_iprod(int n, const void* Q, const void* G, void* d, hipStream_t s);

// R (N x N, N = nblk*n, complex) = I with blocks: R[0][nblk-1] += C[0],
// R[j][j-1] -= C[j] (j >= 1); nblk == 1: R = I + C[0].  C: nblk matrices n x n.
void build_cyclic(int n, int nblk, const void* C, void* R, hipStream_t s);

// Fill a complex n x n matrix with the identity / copy a real matrix into a
// complex one / add the identity in place.
void set_identity_c(int n, void* A, hipStream_t s);
void real_to_complex(int64_t count, const double* in, void* out, hipStream_t s);
void add_identity_c(int n, void* A, hipStream_t s);
// out = a K + id I (real n x n).
void scaled_plus_identity(int n, const double* K, double a, double id, double* out, hipStream_t s);
// out[i][j] = dl[i] * E[i][j] (E real, out complex).
void dscale_rows(int n, const void* dl, const double* E, void* out, hipStream_t s);
// out[i][j] = E[i][j] * dr[j] (E real, out complex).
void dscale_cols(int n, const double* E, const void* dr, void* out, hipStream_t s);
// out = conj(in), m complex values.
void conj_c(int64_t m, const void* in, void* out, hipStream_t s);
// AT = A^T (complex n x n, no aliasing).
void transpose_c(int n, const void* A, void* AT, hipStream_t s);
// d[x] = i * sum_k QT[k][x] G[k][x]  (the diagonal of i * QT^T G), complex.
void diag_iprod_cols(int n, const void* QT, const void* G, void* d, hipStream_t s);
// the same for nb pairs: QT_b = QT + b * stride, G_b = G + b * stride, d_b = d + b * dstride
// (strides in complex elements)
void diag_iprod_cols_batched(int n, int nb, const void* QT, const void* G, int64_t stride, void* d, int64_t dstride,
                             hipStream_t s);

}  // namespace nslx
