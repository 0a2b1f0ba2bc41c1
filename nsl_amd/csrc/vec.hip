// Vector kernels with deterministic per-RHS reductions (see vec.hpp).
#include <algorithm>

#include "vec.hpp"

namespace nslx {
namespace {

constexpr int VNT = 256;
constexpr int VEPT = 4;
constexpr int VCH = VNT * VEPT;   // complex elements per block

template <class S>
struct CV { S re, im; };

template <class S>
__device__ inline CV<S> ld(const void* p, size_t i) {
    if constexpr (sizeof(S) == 8) { const double2 v = reinterpret_cast<const double2*>(p)[i]; return {v.x, v.y}; }
    else { const float2 v = reinterpret_cast<const float2*>(p)[i]; return {v.x, v.y}; }
}
template <class S>
__device__ inline void st(void* p, size_t i, CV<S> v) {
    if constexpr (sizeof(S) == 8) reinterpret_cast<double2*>(p)[i] = make_double2(v.re, v.im);
    else reinterpret_cast<float2*>(p)[i] = make_float2(v.re, v.im);
}

__device__ inline double wsum(double v) {
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block sum of (v0, v1); result valid in all threads.
__device__ inline void bsum2(double& v0, double& v1, double* red) {
    v0 = wsum(v0);
    v1 = wsum(v1);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[2 * w] = v0; red[2 * w + 1] = v1; }
    __syncthreads();
    double a = 0.0, b = 0.0;
    #pragma unroll
    for (int i = 0; i < VNT / 64; ++i) { a += red[2 * i]; b += red[2 * i + 1]; }
    __syncthreads();
    v0 = a; v1 = b;
}

// The per-RHS sums are two launches: the chunk kernel writes one partial
// (2 doubles) per block, k_finish sums them in index order.  (A last-block
// pass in the same launch needs a device-scope fence and an atomic on one
// counter in every block; at 128 x 64 x 288 that ran 6-12x slower than the
// kernels' memory traffic: k_resid64 292 us against 19.5 us now.)
// mode 0: out[b] = sum_k part[b][k].re; 1: out[2b], out[2b+1] = sum_k part[b][k];
// 2: out[0..1] = sum over b and k (thread-strided over b, then the block sum).
__global__ void k_finish(const double* __restrict__ part, int nchunk, int nb, double* out, int mode) {
    __shared__ double red[2 * VNT / 64];
    if (mode == 2) {
        double r = 0.0, im = 0.0;
        for (int bb = threadIdx.x; bb < nb; bb += VNT)
            for (int k = 0; k < nchunk; ++k) {
                r += part[2 * ((size_t)bb * nchunk + k)];
                im += part[2 * ((size_t)bb * nchunk + k) + 1];
            }
        bsum2(r, im, red);
        if (threadIdx.x == 0) { out[0] = r; out[1] = im; }
        return;
    }
    for (int bb = blockIdx.x * VNT + threadIdx.x; bb < nb; bb += gridDim.x * VNT) {
        double r = 0.0, im = 0.0;
        for (int k = 0; k < nchunk; ++k) {
            r += part[2 * ((size_t)bb * nchunk + k)];
            if (mode == 1) im += part[2 * ((size_t)bb * nchunk + k) + 1];
        }
        if (mode == 1) { out[2 * bb] = r; out[2 * bb + 1] = im; }
        else out[bb] = r;
    }
}

__global__ void k_phase(const double2* phi, int64_t n, int sgn, double mut, double mut_im, double2* P, double2* Pinv,
                        float2* P32) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 f = phi[i];
    // exp(i s phi + s mu~) with phi = fr + i fi, mu~ = mr + i mi:
    // exponent = (s mr - s fi) + i s (fr + mi)
    const double er = sgn * mut - sgn * f.y, ei = sgn * (f.x + mut_im);
    const double m = exp(er);
    double sn, cs;
    sincos(ei, &sn, &cs);
    const double2 p = make_double2(m * cs, m * sn);
    P[i] = p;
    const double mi = exp(-er);
    Pinv[i] = make_double2(mi * cs, -mi * sn);
    P32[i] = make_float2((float)p.x, (float)p.y);
}

template <class S>
__global__ void k_norm2(const void* v, int64_t len, int nchunk, double* part) {
    __shared__ double red[2 * VNT / 64];
    const int b = blockIdx.x / nchunk, c = blockIdx.x % nchunk;
    const size_t base = (size_t)b * len;
    double acc = 0.0;
    for (int e = 0; e < VEPT; ++e) {
        const int64_t i = (int64_t)c * VCH + e * VNT + threadIdx.x;
        if (i < len) { const CV<S> z = ld<S>(v, base + i); acc += double(z.re) * z.re + double(z.im) * z.im; }
    }
    double dummy = 0.0;
    bsum2(acc, dummy, red);
    if (threadIdx.x == 0) { part[2 * (size_t)blockIdx.x] = acc; part[2 * (size_t)blockIdx.x + 1] = 0.0; }
}

__global__ void k_set_tol2(const double* bn2, double* tol2, int64_t nb, double tol, int relative, int global) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (global) {
        if (b == 0) tol2[0] = tol * tol;
        return;
    }
    if (b < nb) tol2[b] = tol * tol * (relative ? bn2[b] : 1.0);
}

__global__ void k_resid64(const double2* bv, const double2* t, double2* r, int64_t len, int nchunk, double* part) {
    __shared__ double red[2 * VNT / 64];
    const int b = blockIdx.x / nchunk, c = blockIdx.x % nchunk;
    const size_t base = (size_t)b * len;
    double acc = 0.0;
    for (int e = 0; e < VEPT; ++e) {
        const int64_t i = (int64_t)c * VCH + e * VNT + threadIdx.x;
        if (i < len) {
            const double2 bb = bv[base + i], tt = t[base + i];
            const double2 z = make_double2(bb.x - tt.x, bb.y - tt.y);
            r[base + i] = z;
            acc += z.x * z.x + z.y * z.y;
        }
    }
    double dummy = 0.0;
    bsum2(acc, dummy, red);
    if (threadIdx.x == 0) { part[2 * (size_t)blockIdx.x] = acc; part[2 * (size_t)blockIdx.x + 1] = 0.0; }
}

__global__ void k_scale_to32(const double2* r64, float2* r32, const double* rn2, const double* otol2, int64_t len, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t b = i / len;
    const double nn = rn2[b];
    const double sc = (nn > otol2[b] && nn > 0.0) ? 1.0 / sqrt(nn) : 0.0;
    const double2 z = r64[i];
    r32[i] = make_float2((float)(z.x * sc), (float)(z.y * sc));
}

__global__ void k_update64(double2* x64, const float2* d32, const double* rn2, const double* otol2, int64_t len, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t b = i / len;
    const double nn = rn2[b];
    if (!(nn > otol2[b])) return;
    const double sc = sqrt(nn);
    const float2 d = d32[i];
    double2 x = x64[i];
    x.x += sc * (double)d.x;
    x.y += sc * (double)d.y;
    x64[i] = x;
}

__global__ void k_c128_to_c64(const double2* in, float2* out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { const double2 z = in[i]; out[i] = make_float2((float)z.x, (float)z.y); }
}
__global__ void k_c64_to_c128(const float2* in, double2* out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { const float2 z = in[i]; out[i] = make_double2(z.x, z.y); }
}

template <class S>
__global__ void k_dot(const void* a, const void* bv, int64_t len, int nchunk, double* part) {
    __shared__ double red[2 * VNT / 64];
    const int b = blockIdx.x / nchunk, c = blockIdx.x % nchunk;
    const size_t base = (size_t)b * len;
    double sr = 0.0, si = 0.0;
    for (int e = 0; e < VEPT; ++e) {
        const int64_t i = (int64_t)c * VCH + e * VNT + threadIdx.x;
        if (i < len) {
            const CV<S> x = ld<S>(a, base + i), y = ld<S>(bv, base + i);
            sr += double(x.re) * y.re + double(x.im) * y.im;
            si += double(x.re) * y.im - double(x.im) * y.re;
        }
    }
    bsum2(sr, si, red);
    if (threadIdx.x == 0) { part[2 * (size_t)blockIdx.x] = sr; part[2 * (size_t)blockIdx.x + 1] = si; }
}

template <class S>
__global__ void k_axpy(const void* x, const void* y, void* z, const double* cs, double sign, int64_t len, int64_t n,
                       int global, const int* done) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t b = i / len;
    const int64_t slot = global ? 0 : b;
    if (done && done[slot] >= 0) return;
    const double cr = sign * cs[2 * slot], ci = sign * cs[2 * slot + 1];
    const CV<S> xv = ld<S>(x, i), yv = ld<S>(y, i);
    CV<S> out;
    out.re = S(double(xv.re) + cr * yv.re - ci * yv.im);
    out.im = S(double(xv.im) + cr * yv.im + ci * yv.re);
    st<S>(z, i, out);
}

template <class S>
__global__ void k_ru_to_low(const double2* r64, void* rs, void* p, void* xs, int init, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 z = r64[i];
    const CV<S> v{S(z.x), S(z.y)};
    st<S>(rs, i, v);
    if (init) {
        st<S>(p, i, v);
        st<S>(xs, i, CV<S>{S(0), S(0)});
    }
}

template <class S>
__global__ void k_ru_accum(double2* x64, void* xs, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const CV<S> d = ld<S>(xs, i);
    double2 x = x64[i];
    x.x += double(d.re);
    x.y += double(d.im);
    x64[i] = x;
    st<S>(xs, i, CV<S>{S(0), S(0)});
}

__global__ void k_ru_scalars(const double* rn2, const double* xn2, double floor2, CgDev cg, int64_t nb, int nparts,
                             int cur, int iter, int init, double* best, int* stall, int stall_max, double* rstart,
                             int it0, int budget) {
    for (int64_t b = threadIdx.x; b < nb; b += VNT) {
        const int d = cg.done[b];
        if (init ? (d >= 0 && d != RU_ESCALATE) : d >= 0) continue;
        const double s = rn2[b];
        const int wanted = cg.want[b];
        cg.want[b] = 0;
        if (s <= cg.tol2[b] || (xn2 && s <= floor2 * xn2[b])) {
            cg.done[b] = iter;
            continue;
        }
        double* rp = cg.rpart[cur] + (size_t)b * nparts;
        if (init) {
            best[b] = s;
            stall[b] = 0;
            if (rstart) rstart[b] = s;
        } else if (stall_max > 0) {
            // the low-precision recurrence has lost track of the true residual
            // (true |r| 10x above the recursive one): its rounding errors
            // dominate, so hand the RHS to FP64 now
            double rec = 0.0;
            for (int k = 0; k < nparts; ++k) rec += rp[k];
            if (s > 100.0 * rec) { cg.done[b] = RU_ESCALATE; continue; }
            // stagnation: stall_max replacements this RHS asked for (its
            // recursive |r| fell by delta) without its true |r|^2 falling
            // below a quarter of its best value; replacements other RHS
            // asked for do not count
            if (s < 0.25 * best[b]) { best[b] = s; stall[b] = 0; }
            else if (wanted && ++stall[b] >= stall_max) { cg.done[b] = RU_ESCALATE; continue; }
            // projected budget: at the average rate so far the RHS would need
            // more low-precision iterations than the phase allows
            const int done_its = iter - it0;
            if (budget > 0 && rstart && done_its >= 100) {
                const double r0 = rstart[b];
                if (s >= r0) { cg.done[b] = RU_ESCALATE; continue; }
                const double rate = log(r0 / s) / done_its;
                const double need = log(s / cg.tol2[b]) / rate;
                if (done_its + need > budget) { cg.done[b] = RU_ESCALATE; continue; }
            }
        }
        cg.done[b] = -1;
        cg.rr_ref[b] = s;
        // the replaced residual's |r|^2 becomes the partial sum the next CGA reads
        rp[0] = s;
        for (int k = 1; k < nparts; ++k) rp[k] = 0.0;
    }
}

inline int nchunks(int64_t len) { return (int)((len + VCH - 1) / VCH); }
inline unsigned grid1(int64_t n, int nt = 256) { return (unsigned)((n + nt - 1) / nt); }

}  // namespace

void VecScratch::ensure(int64_t nb, int64_t len, hipStream_t s) {
    (void)s;
    const size_t need = (size_t)nb * nchunks(len) * 2 + 16;
    part.ensure(need);
}

static void finish(const double* part, int nchunk, int64_t nb, double* out, int mode, hipStream_t s) {
    const unsigned g = mode == 2 ? 1u : grid1(nb, VNT);
    hipLaunchKernelGGL(k_finish, dim3(g), dim3(VNT), 0, s, part, nchunk, (int)nb, out, mode);
    NSLX_LAUNCH_CHECK();
}

// max |Im phi| over n values: block maxima, then one block over those
// (no atomics, no memset: every slot is written)
constexpr int MIM_BLOCKS = 256;
__global__ void __launch_bounds__(256) k_max_abs_im(const double2* __restrict__ phi, int64_t n, double* part) {
    __shared__ double red[256];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        m = fmax(m, fabs(phi[i].y));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(256) k_max_part(const double* __restrict__ part, int np, double* out) {
    __shared__ double red[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) m = fmax(m, part[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

void max_abs_im(const void* phi, int64_t n, double* part, double* out, hipStream_t s) {
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(MIM_BLOCKS, (n + 255) / 256));
    hipLaunchKernelGGL(k_max_abs_im, dim3(g), dim3(256), 0, s, (const double2*)phi, n, part);
    NSLX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_max_part, dim3(1), dim3(256), 0, s, (const double*)part, (int)g, out);
    NSLX_LAUNCH_CHECK();
}

void phase_tables(const void* phi, int64_t n, int sgn, double mut, double mut_im, void* P64, void* Pinv64, void* P32,
                  hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_phase, dim3(grid1(n)), dim3(256), 0, s, (const double2*)phi, n, sgn, mut, mut_im, (double2*)P64,
                       (double2*)Pinv64, (float2*)P32);
    NSLX_LAUNCH_CHECK();
}

void norm2_perb(int f64, const void* v, int64_t nb, int64_t len, double* out, VecScratch& sc, hipStream_t s) {
    sc.ensure(nb, len, s);
    const int nc = nchunks(len);
    if (f64) hipLaunchKernelGGL(k_norm2<double>, dim3(nb * nc), dim3(VNT), 0, s, v, len, nc, sc.part.p);
    else hipLaunchKernelGGL(k_norm2<float>, dim3(nb * nc), dim3(VNT), 0, s, v, len, nc, sc.part.p);
    NSLX_LAUNCH_CHECK();
    finish(sc.part.p, nc, nb, out, 0, s);
}

void set_tol2(const double* bn2, double* tol2, int64_t nb, double tol, int relative, int global, hipStream_t s) {
    hipLaunchKernelGGL(k_set_tol2, dim3(grid1(nb)), dim3(256), 0, s, bn2, tol2, nb, tol, relative, global);
    NSLX_LAUNCH_CHECK();
}

void resid64(const void* b, const void* t, void* r, int64_t nb, int64_t len, double* rn2, VecScratch& sc, hipStream_t s) {
    sc.ensure(nb, len, s);
    const int nc = nchunks(len);
    hipLaunchKernelGGL(k_resid64, dim3(nb * nc), dim3(VNT), 0, s, (const double2*)b, (const double2*)t, (double2*)r, len, nc,
                       sc.part.p);
    NSLX_LAUNCH_CHECK();
    finish(sc.part.p, nc, nb, rn2, 0, s);
}

void scale_to32(const void* r64, void* r32, const double* rn2, const double* otol2, int64_t nb, int64_t len, hipStream_t s) {
    const int64_t n = nb * len;
    hipLaunchKernelGGL(k_scale_to32, dim3(grid1(n)), dim3(256), 0, s, (const double2*)r64, (float2*)r32, rn2, otol2, len, n);
    NSLX_LAUNCH_CHECK();
}

void update64(void* x64, const void* d32, const double* rn2, const double* otol2, int64_t nb, int64_t len, hipStream_t s) {
    const int64_t n = nb * len;
    hipLaunchKernelGGL(k_update64, dim3(grid1(n)), dim3(256), 0, s, (double2*)x64, (const float2*)d32, rn2, otol2, len, n);
    NSLX_LAUNCH_CHECK();
}

void ru_residual_to_low(int low_f64, const void* r64, void* rs, void* p, void* xs, int init, int64_t nb, int64_t len,
                        hipStream_t s) {
    const int64_t n = nb * len;
    if (low_f64) hipLaunchKernelGGL(k_ru_to_low<double>, dim3(grid1(n)), dim3(256), 0, s, (const double2*)r64, rs, p, xs, init, n);
    else hipLaunchKernelGGL(k_ru_to_low<float>, dim3(grid1(n)), dim3(256), 0, s, (const double2*)r64, rs, p, xs, init, n);
    NSLX_LAUNCH_CHECK();
}

void ru_accumulate(int low_f64, void* x64, void* xs, int64_t nb, int64_t len, hipStream_t s) {
    const int64_t n = nb * len;
    if (low_f64) hipLaunchKernelGGL(k_ru_accum<double>, dim3(grid1(n)), dim3(256), 0, s, (double2*)x64, xs, n);
    else hipLaunchKernelGGL(k_ru_accum<float>, dim3(grid1(n)), dim3(256), 0, s, (double2*)x64, xs, n);
    NSLX_LAUNCH_CHECK();
}

void ru_scalars(const double* rn2, const double* xn2, double floor2, const CgDev& cg, int64_t nb, int nparts, int cur,
                int iter, int init, double* best, int* stall, int stall_max, double* rstart, int it0, int budget,
                hipStream_t s) {
    hipLaunchKernelGGL(k_ru_scalars, dim3(1), dim3(VNT), 0, s, rn2, xn2, floor2, cg, nb, nparts, cur, iter, init,
                       best, stall, stall_max, rstart, it0, budget);
    NSLX_LAUNCH_CHECK();
}

void c128_to_c64(const void* in, void* out, int64_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_c128_to_c64, dim3(grid1(n)), dim3(256), 0, s, (const double2*)in, (float2*)out, n);
    NSLX_LAUNCH_CHECK();
}
void c64_to_c128(const void* in, void* out, int64_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_c64_to_c128, dim3(grid1(n)), dim3(256), 0, s, (const float2*)in, (double2*)out, n);
    NSLX_LAUNCH_CHECK();
}

void dot_perb(int f64, const void* a, const void* bv, int64_t nb, int64_t len, double* out, int global, VecScratch& sc,
              hipStream_t s) {
    sc.ensure(nb, len, s);
    const int nc = nchunks(len);
    if (f64) hipLaunchKernelGGL(k_dot<double>, dim3(nb * nc), dim3(VNT), 0, s, a, bv, len, nc, sc.part.p);
    else hipLaunchKernelGGL(k_dot<float>, dim3(nb * nc), dim3(VNT), 0, s, a, bv, len, nc, sc.part.p);
    NSLX_LAUNCH_CHECK();
    finish(sc.part.p, nc, nb, out, global ? 2 : 1, s);
}

void axpy_perb(int f64, const void* x, const void* y, void* z, const double* cs, double sign, int64_t nb, int64_t len,
               int global, const int* done, hipStream_t s) {
    const int64_t n = nb * len;
    if (!n) return;
    if (f64) hipLaunchKernelGGL(k_axpy<double>, dim3(grid1(n)), dim3(256), 0, s, x, y, z, cs, sign, len, n, global, done);
    else hipLaunchKernelGGL(k_axpy<float>, dim3(grid1(n)), dim3(256), 0, s, x, y, z, cs, sign, len, n, global, done);
    NSLX_LAUNCH_CHECK();
}

}  // namespace nslx
