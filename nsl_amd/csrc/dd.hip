// Double-double (about 106-bit) residuals of M M^dag and the double-double
// solution update, for solves whose target lies below the FP64 residual
// floor eps ||MM^dag|| |x| (random fields at large beta: |x| / |b| up to
// 1e8, so an FP64 residual cannot show 1e-12).
//
// A double-double value is a pair (hi, lo) with |lo| <= ulp(hi) / 2; the
// error-free transforms are Knuth's TwoSum and the fused-multiply-add
// TwoProduct (p = a b, e = fma(a, b, -p)).  The operator is the one the
// engine applies: the FP64 numbers of exp(s delta K) and of the phase table,
// taken as exact; every product and sum of M^dag and M is carried in
// double-double, and the residual b - M M^dag (x_hi + x_lo) is rounded to
// FP64 only at the end (its own relative precision is then 2^-53, which is
// all a residual needs).
#pragma clang fp contract(off)

#include <cfloat>

#include "dd.hpp"
#include "engine.hpp"
#include "prof.hpp"

namespace nslx {
namespace {

struct dd { double h, l; };

__device__ inline dd two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ inline dd quick(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
__device__ inline dd add(dd a, dd b) {
    dd s = two_sum(a.h, b.h);
    const dd t = two_sum(a.l, b.l);
    s.l = s.l + t.h;
    s = quick(s.h, s.l);
    s.l = s.l + t.l;
    return quick(s.h, s.l);
}
__device__ inline dd neg(dd a) { return {-a.h, -a.l}; }
__device__ inline dd mul_d(dd a, double b) {
    const double p = a.h * b;
    double e = __builtin_fma(a.h, b, -p);
    e = __builtin_fma(a.l, b, e);
    return quick(p, e);
}

struct cdd { dd re, im; };

__device__ inline cdd cadd(cdd a, cdd b) { return {add(a.re, b.re), add(a.im, b.im)}; }
__device__ inline cdd neg_c(cdd a) { return {neg(a.re), neg(a.im)}; }
// a * p for a complex double p; conj_p uses conj(p)
__device__ inline cdd cmul_d(cdd a, double2 p, bool conj_p) {
    const double pr = p.x, pi = conj_p ? -p.y : p.y;
    const cdd out{add(mul_d(a.re, pr), neg(mul_d(a.im, pi))), add(mul_d(a.re, pi), mul_d(a.im, pr))};
    return out;
}
__device__ inline cdd cscale(cdd a, double s) { return {mul_d(a.re, s), mul_d(a.im, s)}; }

__device__ inline cdd load(const double2* h, const double2* l, long long i) {
    const double2 a = h[i];
    const double2 b = l ? l[i] : make_double2(0.0, 0.0);
    return {quick(a.x, b.x), quick(a.y, b.y)};
}

// Accumulator of one output element on a fixed grid (an error-free
// extraction, as in Rump, Ogita and Oishi's ExtractScalar): g = 1.5 * 2^m with
// 2^(m-1) above the row's bound B on sum |e| |v.h| (k_dd_rowsig, B >= 2^(m-2)).
// fma(e, v.h, g) lies in [2^m, 2^(m+1)], so ph = fma(e, v.h, g) - g is the
// exact product rounded to a multiple of 2^(m-52), and the ph sum exactly in s
// (every partial sum below 2^m).  The rest, e v.h - ph (below 2^(m-53)), is
// added to c rounded once (error 2^(m-106) per product, a factor n below the
// rounding of c itself), and e * v.l after it.  Six FP64 operations per
// product against eleven with a TwoProduct and a TwoSum; the error bound is of
// the same order, n^2 2^-104 B, from the rounding in c.
struct Acc {
    double s = 0.0, c = 0.0;
    __device__ inline void fma(double e, dd v, double g) {
        const double ph = __builtin_fma(e, v.h, g) - g;
        s = s + ph;
        c = c + __builtin_fma(e, v.h, -ph);
        c = __builtin_fma(e, v.l, c);
    }
    __device__ inline dd value() const { return quick(s, c); }
};

// threads per block, one output column each: one wave, so that Nx = 288
// fills 4.5 blocks of columns (90 % of the threads busy; 75 % with 128)
constexpr int DCOLS = 64;
constexpr int DROWS = 8;     // output rows (rhs, t) per block
constexpr int DKB = 32;      // K chunk staged in LDS

// One hop of M^dag (OP 1) or M (OP 0) on a double-double field, rows
// g = b*nt + t of a batch:
//   M^dag: out_t = in_t - b'_t conj(P_t) (E in_{t+1})      b'_{nt-1} = -1
//   M:     out_t = in_t - b_t E (P_{t-1} in_{t-1})          b_0 = -1
// FINAL (M only): r = bvec - out, rounded; part[g * ncol + blockIdx.x] = |r_g|^2
// over this block's columns.  ET: the transpose of the hop's E (ET[x][y] =
// E[y][x], so that a block's threads read consecutive columns).  E is not
// used as its own transpose: the FP64 exp(s delta K) is symmetric only to an
// ulp, and at |x| / |b| ~ 1e8 that ulp is a 1e-9 change of the residual.
template <int OP, bool FINAL>
__global__ void __launch_bounds__(DCOLS) k_dd_hop(int n, int nt, long long nrows, int rpf, const double* __restrict__ ET,
                                                  const double2* __restrict__ P, const double2* inh, const double2* inl,
                                                  const double* __restrict__ sig, double2* outh, double2* outl,
                                                  const double2* bvec, double2* r, double* part) {
    __shared__ double sh[4][DROWS][DKB + 1];   // re.h, re.l, im.h, im.l of the staged input rows
    const int tid = threadIdx.x;
    const int y = blockIdx.x * DCOLS + tid;
    const long long g0 = (long long)blockIdx.y * DROWS;
    Acc ar[DROWS], ai[DROWS];
    // the accumulation grid of each output row (uniform over the block)
    double gr[DROWS];
    #pragma unroll
    for (int rr = 0; rr < DROWS; ++rr) gr[rr] = g0 + rr < nrows ? sig[g0 + rr] : 1.5;
    // this thread's staged elements: rows rr = (tid + i DCOLS) / DKB at K
    // offset kk = tid % DKB of every chunk (DCOLS a multiple of DKB), so the
    // input and phase row offsets are computed once, not per chunk
    static_assert(DCOLS % DKB == 0, "staging rows fixed per thread");
    constexpr int NSTG = DROWS * DKB / DCOLS;
    const int kk = tid % DKB;
    long long in_off[NSTG], ph_off[NSTG];
    bool g_ok[NSTG];
    #pragma unroll
    for (int i = 0; i < NSTG; ++i) {
        const long long g = g0 + (tid + i * DCOLS) / DKB;
        g_ok[i] = g < nrows;
        const long long gg = g_ok[i] ? g : 0;
        const long long b = gg / nt;
        const int t = (int)(gg % nt);
        const int tin = OP == 0 ? (t == 0 ? nt - 1 : t - 1) : (t + 1 == nt ? 0 : t + 1);
        in_off[i] = (b * nt + tin) * n;
        ph_off[i] = OP == 0 ? ((b / rpf) * nt + tin) * n : 0;
    }
    for (int x0 = 0; x0 < n; x0 += DKB) {
        const int x = x0 + kk;
        #pragma unroll
        for (int i = 0; i < NSTG; ++i) {
            const int rr = (tid + i * DCOLS) / DKB;
            cdd v{{0.0, 0.0}, {0.0, 0.0}};
            if (g_ok[i] && x < n) {
                v = load(inh, inl, in_off[i] + x);
                if (OP == 0) v = cmul_d(v, P[ph_off[i] + x], false);
            }
            sh[0][rr][kk] = v.re.h;
            sh[1][rr][kk] = v.re.l;
            sh[2][rr][kk] = v.im.h;
            sh[3][rr][kk] = v.im.l;
        }
        __syncthreads();
        if (y < n) {
            const int kmax = min(DKB, n - x0);
            for (int kk = 0; kk < kmax; ++kk) {
                const double e = ET[(long long)(x0 + kk) * n + y];
                #pragma unroll
                for (int rr = 0; rr < DROWS; ++rr) {
                    ar[rr].fma(e, dd{sh[0][rr][kk], sh[1][rr][kk]}, gr[rr]);
                    ai[rr].fma(e, dd{sh[2][rr][kk], sh[3][rr][kk]}, gr[rr]);
                }
            }
        }
        __syncthreads();
    }
    double rsq[DROWS];
    #pragma unroll
    for (int rr = 0; rr < DROWS; ++rr) {
        rsq[rr] = 0.0;
        const long long g = g0 + rr;
        if (g >= nrows || y >= n) continue;
        const long long b = g / nt;
        const int t = (int)(g % nt);
        const long long idx = g * n + y;
        cdd res{ar[rr].value(), ai[rr].value()};
        double sg;
        if (OP == 1) {
            const long long f = b / rpf;
            res = cmul_d(res, P[(f * nt + t) * n + y], true);
            sg = t == nt - 1 ? -1.0 : 1.0;
        } else {
            sg = t == 0 ? -1.0 : 1.0;
        }
        const cdd in = load(inh, inl, idx);
        const cdd out = cadd(in, neg_c(cscale(res, sg)));
        if (FINAL) {
            const double2 bv = bvec[idx];
            const cdd bb{{bv.x, 0.0}, {bv.y, 0.0}};
            const cdd rd = cadd(bb, neg_c(out));
            const double rr_ = rd.re.h + rd.re.l, ri = rd.im.h + rd.im.l;
            r[idx] = make_double2(rr_, ri);
            rsq[rr] = rr_ * rr_ + ri * ri;
        } else {
            outh[idx] = make_double2(out.re.h, out.im.h);
            outl[idx] = make_double2(out.re.l, out.im.l);
        }
    }
    if (FINAL) {
        // the row sums reuse the staging buffer (free since the K loop's
        // last barrier): no extra LDS, so as many blocks fit a CU as without
        static_assert(DROWS * DCOLS <= 4 * DROWS * (DKB + 1), "row sums fit the staging buffer");
        double* red = &sh[0][0][0];
        #pragma unroll
        for (int rr = 0; rr < DROWS; ++rr) red[rr * DCOLS + tid] = rsq[rr];
        __syncthreads();
        if (tid < DROWS) {
            const long long g = g0 + tid;
            double sum = 0.0;
            for (int k = 0; k < DCOLS; ++k) sum += red[tid * DCOLS + k];
            if (g < nrows) part[g * gridDim.x + blockIdx.x] = sum;
        }
    }
}

// rn2[b] = sum over its nt rows and ncol column blocks, in index order
__global__ void k_dd_rsum(const double* part, int nt, int ncol, long long nb, double* rn2) {
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double s = 0.0;
    const long long m = (long long)nt * ncol;
    for (long long k = 0; k < m; ++k) s += part[b * m + k];
    rn2[b] = s;
}

// (xh, xl) += d for RHS b with done[b] < 0 (done may be null)
__global__ void k_dd_accumulate(double2* xh, double2* xl, const double2* d, long long len, long long n,
                                const int* done) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (done && done[i / len] >= 0) return;
    const double2 h = xh[i], l = xl[i], v = d[i];
    dd re = two_sum(h.x, v.x), im = two_sum(h.y, v.y);
    re.l = re.l + l.x;
    im.l = im.l + l.y;
    re = quick(re.h, re.l);
    im = quick(im.h, im.l);
    xh[i] = make_double2(re.h, im.h);
    xl[i] = make_double2(re.l, im.l);
}

// *out = max over rows y of sum_x |E[y][x]| (one block)
__global__ void __launch_bounds__(256) k_dd_enorm(int n, const double* E, double* out) {
    __shared__ double red[256];
    double m = 0.0;
    for (int y = threadIdx.x; y < n; y += 256) {
        double sum = 0.0;
        for (int x = 0; x < n; ++x) sum += fabs(E[(long long)y * n + x]);
        m = fmax(m, sum);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// The accumulation grid of each row g of a hop's input (one wave per row):
// sig[g] = 1.5 * 2^m with 2^(m-2) <= B < 2^(m-1), B = ||E||_inf * max_x
// (|v_x.re| + |v_x.im|) * (1 + 2^-40), v the row the hop stages
// (P_{t-1} . in_{t-1} for M, in_{t+1} for M^dag; |v| bounded through the
// double-double parts' magnitudes), so that B bounds sum_x |E[y][x]| |v_x|
// for every output column y, rounding of the bound included.
template <int OP>
__global__ void __launch_bounds__(256) k_dd_rowsig(int n, int nt, long long nrows, int rpf,
                                                   const double2* __restrict__ P, const double2* inh,
                                                   const double2* inl, const double* enorm, double* sig) {
    const long long g = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (g >= nrows) return;
    const long long b = g / nt;
    const int t = (int)(g % nt);
    const int tin = OP == 0 ? (t == 0 ? nt - 1 : t - 1) : (t + 1 == nt ? 0 : t + 1);
    const long long base = (b * nt + tin) * n;
    const long long pbase = OP == 0 ? ((b / rpf) * nt + tin) * n : 0;
    double m = 0.0;
    for (int x = lane; x < n; x += 64) {
        const double2 h = inh[base + x];
        const double2 l = inl ? inl[base + x] : make_double2(0.0, 0.0);
        double a = fabs(h.x) + fabs(l.x) + fabs(h.y) + fabs(l.y);
        if (OP == 0) {
            const double2 ph = P[pbase + x];
            a *= fabs(ph.x) + fabs(ph.y);
        }
        m = fmax(m, a);
    }
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (lane == 0) {
        const double bound = (*enorm) * m * (1.0 + 0x1p-40);
        // an all-zero row takes any grid; a non-finite bound makes the row
        // non-finite, as its products would
        sig[g] = !(bound > 0.0) ? (bound == 0.0 ? 1.5 : bound)
                 : bound <= DBL_MAX ? 1.5 * ldexp(1.0, ilogb(bound) + 2) : bound;
    }
}

__global__ void k_transpose(int n, const double* E, double* ET) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * n) return;
    const int y = (int)(e / n), x = (int)(e % n);
    ET[(long long)x * n + y] = E[e];
}

inline unsigned blocks1(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void DdWork::ensure(int64_t nb, int64_t len, int n, int nt) {
    const size_t v = (size_t)nb * len;
    uh.ensure(v);
    ul.ensure(v);
    const int ncol = (n + DCOLS - 1) / DCOLS;
    part.ensure((size_t)nb * nt * ncol);
    sig.ensure((size_t)nb * nt);
    enorm.ensure(1);
}

void dd_residual(FermionMatrix& fm, int op, const void* b, const void* xh, const void* xl, void* r, double* rn2,
                 int64_t nb, DdWork& w, hipStream_t s) {
    NSLX_REQUIRE(op == NSLX_OP_MMDAGGER || op == NSLX_OP_MDAGGERM, "dd residual: MMdagger or MdaggerM");
    NSLX_REQUIRE(fm.disc == DISC_EXP, "dd residual: HubbardExp only");
    fm.require_populated();
    NSLX_REQUIRE(nb >= 1 && nb % fm.nfield == 0, "dd residual: nbatch must be a positive multiple of the fields");
    const int n = fm.nx, nt = fm.nt;
    const int rpf = (int)(nb / fm.nfield);
    const int64_t len = fm.len();
    w.ensure(nb, len, n, nt);
    const double* E = fm.E(fm.sgn() * fm.delta(), s).dense.p;
    if (w.et_src != E) {
        w.et.ensure((size_t)n * n);
        hipLaunchKernelGGL(k_transpose, dim3(blocks1((long long)n * n)), dim3(256), 0, s, n, E, w.et.p);
        NSLX_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_dd_enorm, dim3(1), dim3(256), 0, s, n, E, w.enorm.p);
        NSLX_LAUNCH_CHECK();
        w.et_src = E;
    }
    const double* ET = w.et.p;
    const double2* P = fm.P64.p;
    const long long nrows = (long long)nb * nt;
    const int ncol = (n + DCOLS - 1) / DCOLS;
    dim3 g(ncol, (unsigned)((nrows + DROWS - 1) / DROWS));
    const double2* ih = static_cast<const double2*>(xh);
    const double2* il = static_cast<const double2*>(xl);
    // the accumulation grids of a hop's rows, from its input
    const dim3 gsig((unsigned)((nrows + 3) / 4));
    auto rowsig = [&](int hop_op, const double2* inh, const double2* inl) {
        if (hop_op == 0)
            hipLaunchKernelGGL(k_dd_rowsig<0>, gsig, dim3(256), 0, s, n, nt, nrows, rpf, P, inh, inl, w.enorm.p,
                               w.sig.p);
        else
            hipLaunchKernelGGL(k_dd_rowsig<1>, gsig, dim3(256), 0, s, n, nt, nrows, rpf, P, inh, inl, w.enorm.p,
                               w.sig.p);
        NSLX_LAUNCH_CHECK();
    };
    // one hop's algorithmic work: per output element and site two real
    // double-double multiply-adds (TwoProduct with FMA + TwoSum + the low
    // parts, 10 flops each); bytes: the input and output fields in
    // double-double (32 B per complex element), E once
    const double flops = 20.0 * (double)nrows * n * n;
    const double bytes = 64.0 * (double)nrows * n + 8.0 * (double)n * n;
    // first hop into (uh, ul), second into the residual
    {
        ProfScope prof(prof_on() ? std::string(op == NSLX_OP_MMDAGGER ? "dd_hop_mdag" : "dd_hop_m") : std::string(), s);
        prof.work(flops, bytes);
        rowsig(op == NSLX_OP_MMDAGGER ? 1 : 0, ih, il);
        if (op == NSLX_OP_MMDAGGER)
            launch_timed(k_dd_hop<1, false>, g, dim3(DCOLS), 0, s, n, nt, nrows, rpf, ET, P, ih, il, w.sig.p, w.uh.p,
                         w.ul.p, (const double2*)nullptr, (double2*)nullptr, (double*)nullptr);
        else
            launch_timed(k_dd_hop<0, false>, g, dim3(DCOLS), 0, s, n, nt, nrows, rpf, ET, P, ih, il, w.sig.p, w.uh.p,
                         w.ul.p, (const double2*)nullptr, (double2*)nullptr, (double*)nullptr);
        NSLX_LAUNCH_CHECK();
    }
    {
        ProfScope prof(prof_on() ? std::string(op == NSLX_OP_MMDAGGER ? "dd_hop_m_resid" : "dd_hop_mdag_resid")
                                 : std::string(), s);
        prof.work(flops, bytes);
        const double2* uh = w.uh.p;
        const double2* ul = w.ul.p;
        rowsig(op == NSLX_OP_MMDAGGER ? 0 : 1, uh, ul);
        if (op == NSLX_OP_MMDAGGER)
            launch_timed(k_dd_hop<0, true>, g, dim3(DCOLS), 0, s, n, nt, nrows, rpf, ET, P, uh, ul, w.sig.p,
                         (double2*)nullptr, (double2*)nullptr, static_cast<const double2*>(b),
                         static_cast<double2*>(r), w.part.p);
        else
            launch_timed(k_dd_hop<1, true>, g, dim3(DCOLS), 0, s, n, nt, nrows, rpf, ET, P, uh, ul, w.sig.p,
                         (double2*)nullptr, (double2*)nullptr, static_cast<const double2*>(b),
                         static_cast<double2*>(r), w.part.p);
        NSLX_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_dd_rsum, dim3(blocks1(nb)), dim3(256), 0, s, w.part.p, nt, ncol, (long long)nb, rn2);
    NSLX_LAUNCH_CHECK();
}

void dd_accumulate(void* xh, void* xl, const void* d, int64_t nb, int64_t len, const int* done, hipStream_t s) {
    const long long n = (long long)nb * len;
    if (!n) return;
    hipLaunchKernelGGL(k_dd_accumulate, dim3(blocks1(n)), dim3(256), 0, s, static_cast<double2*>(xh),
                       static_cast<double2*>(xl), static_cast<const double2*>(d), (long long)len, n, done);
    NSLX_LAUNCH_CHECK();
}

}  // namespace nslx
