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

#include <algorithm>
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

// A block is one wave computing a tile of DTR output rows (rhs, t) by
// DTY = 8 DCL output columns y; lane (rg, cl) = (tid / 8, tid % 8) owns rows
// rg + 8 i (i < DRL) and columns DCL cl + j (j < DCL).  Every input value
// read from LDS serves DCL products and every E value DRL.  DCL = 4 (32
// columns, up to 168 registers: 3 waves per SIMD) or 6 (48 columns, 2 waves
// per SIMD); dd_residual picks by how the tiles fill the device's wave
// slots (hop_columns).
constexpr int DTR = 16;             // output rows per block
constexpr int DRL = 2;              // rows per lane
constexpr int DKB = 8;              // K chunk staged in LDS
constexpr int DRS = 4 * DKB + 2;    // staged row stride in doubles (rows 16 B apart modulo 128 B)
constexpr int DTY_MIN = 32;         // the narrowest tile (partial sums per row)
static_assert(8 * (DTR / DRL) == 64, "one wave per tile");
static_assert(DTR * DKB % 64 == 0, "staging evenly split");

// One hop of M^dag (OP 1) or M (OP 0) on a double-double field, rows
// g = b*nt + t of a batch:
//   M^dag: out_t = in_t - b'_t conj(P_t) (E in_{t+1})      b'_{nt-1} = -1
//   M:     out_t = in_t - b_t E (P_{t-1} in_{t-1})          b_0 = -1
// FINAL (M only): r = bvec - out, rounded; part[g * ncol + blockIdx.x] = |r_g|^2
// over this block's columns.  ET: the transpose of the hop's E (ET[x][y] =
// E[y][x], so that a block stages rows of consecutive columns).  E is not
// used as its own transpose: the FP64 exp(s delta K) is symmetric only to an
// ulp, and at |x| / |b| ~ 1e8 that ulp is a 1e-9 change of the residual.
// sig: the accumulation grid of each row (k_dd_rowsig).
template <int OP, bool FINAL, int DCL>
__global__ void __launch_bounds__(64, DCL == 4 ? 3 : 2) k_dd_hop(int n, int nt, long long nrows, int rpf, const double* __restrict__ ET,
                                               const double2* __restrict__ P, const double2* inh, const double2* inl,
                                               const double* __restrict__ sig, double2* outh, double2* outl,
                                               const double2* bvec, double2* r, double* part) {
    constexpr int DTY = 8 * DCL;
    static_assert(DCL % 2 == 0 && DKB * DTY % 64 == 0, "E tile read in pairs, staged evenly");
    // staged input rows (re.h, re.l, im.h, im.l per x) and the ET tile
    __shared__ __align__(16) double sv[DTR * DRS];
    __shared__ __align__(16) double se[DKB * DTY];
    const int tid = threadIdx.x;
    const int cl = tid % (DTY / DCL), rg = tid / (DTY / DCL);
    const int y0 = blockIdx.x * DTY;
    const long long g0 = (long long)blockIdx.y * DTR;
    Acc ar[DRL][DCL], ai[DRL][DCL];
    double gr[DRL];
    #pragma unroll
    for (int i = 0; i < DRL; ++i) {
        const long long g = g0 + rg + 8 * i;
        gr[i] = g < nrows ? sig[g] : 1.5;
    }
    // this thread's staged input elements: rows (tid + 64 s) / DKB at K
    // offset kk = tid % DKB of every chunk; its ET elements: (k, c) =
    // ((tid + 64 s) / DTY, (tid + 64 s) % DTY)
    constexpr int NSTG = DTR * DKB / 64;
    constexpr int NSE = DKB * DTY / 64;
    const int kk = tid % DKB;
    long long in_off[NSTG], ph_off[NSTG];
    bool g_ok[NSTG];
    #pragma unroll
    for (int i = 0; i < NSTG; ++i) {
        const long long g = g0 + (tid + i * 64) / DKB;
        g_ok[i] = g < nrows;
        const long long gg = g_ok[i] ? g : 0;
        const long long b = gg / nt;
        const int t = (int)(gg % nt);
        const int tin = OP == 0 ? (t == 0 ? nt - 1 : t - 1) : (t + 1 == nt ? 0 : t + 1);
        in_off[i] = (b * nt + tin) * n;
        ph_off[i] = OP == 0 ? ((b / rpf) * nt + tin) * n : 0;
    }
    // One chunk's global loads go to registers a chunk ahead of their use,
    // so their latency overlaps the previous chunk's products.  The loads
    // are unconditional (addresses clamped into the arrays, out-of-range
    // values replaced when staged): the compiler then waits for each load
    // only where its value is used.
    const bool has_l = inl != nullptr;
    const double2* inl_ = has_l ? inl : inh;
    double2 fh[NSTG], fl[NSTG], fp[NSTG];
    double fe[NSE];
    auto fetch = [&](int x0) {
        const int x = min(x0 + kk, n - 1);
        #pragma unroll
        for (int i = 0; i < NSTG; ++i) {
            fh[i] = inh[in_off[i] + x];
            fl[i] = inl_[in_off[i] + x];
            if (OP == 0) fp[i] = P[ph_off[i] + x];
        }
        #pragma unroll
        for (int i = 0; i < NSE; ++i) {
            const int q = tid + i * 64;
            const int xe = min(x0 + q / DTY, n - 1), ye = min(y0 + q % DTY, n - 1);
            fe[i] = ET[(long long)xe * n + ye];
        }
    };
    auto stage = [&](int x0) {
        const int x = x0 + kk;
        #pragma unroll
        for (int i = 0; i < NSTG; ++i) {
            const int row = (tid + i * 64) / DKB;
            cdd v{{0.0, 0.0}, {0.0, 0.0}};
            if (g_ok[i] && x < n) {
                const double2 l = has_l ? fl[i] : make_double2(0.0, 0.0);
                v = cdd{quick(fh[i].x, l.x), quick(fh[i].y, l.y)};
                if (OP == 0) v = cmul_d(v, fp[i], false);
            }
            double2* d = reinterpret_cast<double2*>(&sv[row * DRS + kk * 4]);
            d[0] = make_double2(v.re.h, v.re.l);
            d[1] = make_double2(v.im.h, v.im.l);
        }
        #pragma unroll
        for (int i = 0; i < NSE; ++i) {
            const int q = tid + i * 64;
            const int xe = x0 + q / DTY, ye = y0 + q % DTY;
            se[q] = xe < n && ye < n ? fe[i] : 0.0;
        }
    };
    fetch(0);
    for (int x0 = 0; x0 < n; x0 += DKB) {
        stage(x0);
        __syncthreads();
        fetch(x0 + DKB);    // past the end: clamped re-reads, never staged
        // padded x (beyond n) hold zeros in both operands: exact zero
        // products.  Unrolled by 2: fully unrolled, the hoisted LDS reads
        // take the registers of a fourth wave per SIMD
        #pragma unroll 2
        for (int k = 0; k < DKB; ++k) {
            double e[DCL];
            #pragma unroll
            for (int j = 0; j < DCL; j += 2) {
                const double2 ep = *reinterpret_cast<const double2*>(&se[k * DTY + DCL * cl + j]);
                e[j] = ep.x;
                e[j + 1] = ep.y;
            }
            #pragma unroll
            for (int i = 0; i < DRL; ++i) {
                const double* vr = &sv[(rg + 8 * i) * DRS + 4 * k];
                const double2 vre = *reinterpret_cast<const double2*>(vr);
                const double2 vim = *reinterpret_cast<const double2*>(vr + 2);
                #pragma unroll
                for (int j = 0; j < DCL; ++j) {
                    ar[i][j].fma(e[j], dd{vre.x, vre.y}, gr[i]);
                    ai[i][j].fma(e[j], dd{vim.x, vim.y}, gr[i]);
                }
            }
        }
        __syncthreads();
    }
    double rsq[DRL];
    #pragma unroll
    for (int i = 0; i < DRL; ++i) {
        rsq[i] = 0.0;
        const long long g = g0 + rg + 8 * i;
        if (g >= nrows) continue;
        const long long b = g / nt;
        const int t = (int)(g % nt);
        #pragma unroll
        for (int j = 0; j < DCL; ++j) {
            const int y = y0 + DCL * cl + j;
            if (y >= n) continue;
            const long long idx = g * n + y;
            cdd res{ar[i][j].value(), ai[i][j].value()};
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
                rsq[i] += rr_ * rr_ + ri * ri;
            } else {
                outh[idx] = make_double2(out.re.h, out.im.h);
                outl[idx] = make_double2(out.re.l, out.im.l);
            }
        }
    }
    if (FINAL) {
        // a row's partial over the block's columns: the 8 lanes of its row
        // group (consecutive lanes), in a fixed tree order
        #pragma unroll
        for (int i = 0; i < DRL; ++i) {
            double v = rsq[i];
            #pragma unroll
            for (int o = 1; o < DTY / DCL; o <<= 1) v += __shfl_xor(v, o, 64);
            const long long g = g0 + rg + 8 * i;
            if (cl == 0 && g < nrows) part[g * gridDim.x + blockIdx.x] = v;
        }
    }
}

// rn2[b] = sum over its nt rows and ncol column blocks (one wave per b: lane
// l sums partials l, l + 64, ... in order, then a fixed xor tree)
__global__ void __launch_bounds__(256) k_dd_rsum(const double* part, int nt, int ncol, long long nb, double* rn2) {
    const long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b >= nb) return;
    double s = 0.0;
    const long long m = (long long)nt * ncol;
    for (long long k = lane; k < m; k += 64) s += part[b * m + k];
    #pragma unroll
    for (int o = 1; o < 64; o <<= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) rn2[b] = s;
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

template <int OP, bool FINAL>
void launch_hop(int dcl, int n, int nt, long long nrows, int rpf, const double* ET, const double2* P,
                const double2* inh, const double2* inl, const double* sig, double2* outh, double2* outl,
                const double2* bvec, double2* r, double* part, hipStream_t s) {
    const dim3 g((unsigned)((n + 8 * dcl - 1) / (8 * dcl)), (unsigned)((nrows + DTR - 1) / DTR));
    if (dcl == 6)
        launch_timed(k_dd_hop<OP, FINAL, 6>, g, dim3(64), 0, s, n, nt, nrows, rpf, ET, P, inh, inl, sig, outh, outl,
                     bvec, r, part);
    else
        launch_timed(k_dd_hop<OP, FINAL, 4>, g, dim3(64), 0, s, n, nt, nrows, rpf, ET, P, inh, inl, sig, outh, outl,
                     bvec, r, part);
    NSLX_LAUNCH_CHECK();
}

// Columns per lane of the hop tiles: each variant's rounds of waves (tiles
// over the device's resident waves) weighed by the round's length (resident
// waves per SIMD x work per wave); the resident counts come from the
// runtime's occupancy query for the kernels as compiled.  Ties go to the
// wider tile (fewer LDS bytes per product).
int hop_columns(int n, long long nrows) {
    static int occ4 = 0, occ6 = 0, ncu = 0;
    if (!ncu) {
        int dev = 0;
        NSLX_HIP(hipGetDevice(&dev));
        NSLX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        NSLX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ4, k_dd_hop<0, true, 4>, 64, 0));
        NSLX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ6, k_dd_hop<0, true, 6>, 64, 0));
    }
    const long long rt = (nrows + DTR - 1) / DTR;
    auto cost = [&](int dcl, int occ) {
        const long long tiles = rt * ((n + 8 * dcl - 1) / (8 * dcl));
        const long long slots = (long long)std::max(ncu * occ, 1);
        return (double)((tiles + slots - 1) / slots) * occ * dcl;
    };
    return cost(6, occ6) <= cost(4, occ4) ? 6 : 4;
}

}  // namespace

void DdWork::ensure(int64_t nb, int64_t len, int n, int nt) {
    const size_t v = (size_t)nb * len;
    uh.ensure(v);
    ul.ensure(v);
    const int ncol = (n + DTY_MIN - 1) / DTY_MIN;
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
    NSLX_REQUIRE((nrows + DTR - 1) / DTR <= 65535, "dd residual: at most 65535 * 16 rows (nbatch * nt) per call");
    const int dcl = hop_columns(n, nrows);
    const int ncol = (n + 8 * dcl - 1) / (8 * dcl);
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
            launch_hop<1, false>(dcl, n, nt, nrows, rpf, ET, P, ih, il, w.sig.p, w.uh.p, w.ul.p, nullptr, nullptr,
                                 nullptr, s);
        else
            launch_hop<0, false>(dcl, n, nt, nrows, rpf, ET, P, ih, il, w.sig.p, w.uh.p, w.ul.p, nullptr, nullptr,
                                 nullptr, s);
    }
    {
        ProfScope prof(prof_on() ? std::string(op == NSLX_OP_MMDAGGER ? "dd_hop_m_resid" : "dd_hop_mdag_resid")
                                 : std::string(), s);
        prof.work(flops, bytes);
        const double2* uh = w.uh.p;
        const double2* ul = w.ul.p;
        rowsig(op == NSLX_OP_MMDAGGER ? 0 : 1, uh, ul);
        if (op == NSLX_OP_MMDAGGER)
            launch_hop<0, true>(dcl, n, nt, nrows, rpf, ET, P, uh, ul, w.sig.p, nullptr, nullptr,
                                static_cast<const double2*>(b), static_cast<double2*>(r), w.part.p, s);
        else
            launch_hop<1, true>(dcl, n, nt, nrows, rpf, ET, P, uh, ul, w.sig.p, nullptr, nullptr,
                                static_cast<const double2*>(b), static_cast<double2*>(r), w.part.p, s);
    }
    hipLaunchKernelGGL(k_dd_rsum, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, s, w.part.p, nt, ncol, (long long)nb,
                       rn2);
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
