// Persistent recurrence launches ("chains") of the time-slice solver.  The
// in-chunk passes x_t = y_t + b_t E diag(P_{t-1}) x_{t-1} (and the M^dag
// recurrence) are sequences of batched GEMM steps in which step s + 1 reads
// as A exactly the rows step s wrote as Y.  One launch runs every step:
//
//   * a workgroup owns the 16-row block rt and the 32-column tile m of one
//     chain (chunk); its slice of B (= exp(delta K), the same in every step)
//     stays in LDS for the whole launch;
//   * per step each of its 8 waves loads its share of the 16 A rows (K steps
//     of 4 columns, every 8th one) straight into registers, applies the phase
//     there and runs v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32 against
//     both 16-column blocks of the tile; the 8 partial tiles are summed
//     through LDS and the 16 x 32 output tile is written;
//   * the column tiles of one (chain, row block) -- a group -- meet at a
//     counter in device memory between steps.  Rows of different groups
//     never meet, so there is no grid-wide barrier.
//
// Visibility of one step's tiles to the group's next step.  A group's
// workgroups are numbered so that the round-robin dispatch of workgroups over
// the 8 XCDs (blockIdx mod 8) puts all of them on one XCD, i.e. behind one L2,
// and each workgroup checks where it actually runs (its XCC id register).
// The first exchange (step 0 -> 1) always goes through memory: every
// workgroup writes its L2 back (release, agent scope) before signalling, and
// the signal carries a mark when the workgroup is not on the XCD its number
// implies.  Once the first wait is over every member knows whether the whole
// group shares one L2.  If it does, the writer's side of later exchanges
// needs only each wave's stores to be complete (s_waitcnt vmcnt(0)) before
// the signal: the L2 is the coherence point of its XCD and the L1 is
// write-through.  Otherwise every signal keeps the L2 write-back.  The
// reader's side always takes one acquire (agent scope: L1 and stale L2 lines
// invalidated) by one wave before the workgroup barrier: a 128-byte line can
// hold the end of one slice row and the start of the next (FP32 rows of 72
// sites are 576 bytes), so reading step s's input can leave an old copy of
// part of step s + 1's input in the L1.
//
#include <algorithm>
#include <cstdlib>

#include "prof.hpp"
#include "tss.hpp"

namespace nslx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int CBM = 16, CBN = 32, CT = 512;
constexpr int CHAIN_SPIN = 1 << 20;   // polls before a wait gives up (~1 s)
// group counter: arrivals in the low 16 bits, misplaced members above
constexpr unsigned ARRIVALS = 0xFFFFu, MISPLACED = 1u << 16;
// s_getreg_b32 operand of the XCC id (hardware register 20, bits [3:0])
constexpr int HWREG_XCC_ID = 20 | (0 << 6) | ((4 - 1) << 11);
// s_waitcnt vmcnt(0) with lgkmcnt / expcnt left free (simm16 encoding)
constexpr int WAIT_VMCNT0 = 0x0F70;

template <class S> struct ChainT;
template <> struct ChainT<double> { using Acc = d4; static constexpr int KMAX = 304; };
template <> struct ChainT<float> { using Acc = f4; static constexpr int KMAX = 608; };

template <class S> struct Z2 { S re, im; };

template <class S>
__device__ inline Z2<S> ldz2(const void* base, long long idx) {
    if constexpr (sizeof(S) == 8) {
        const double2 v = reinterpret_cast<const double2*>(base)[idx];
        return {v.x, v.y};
    } else {
        const float2 v = reinterpret_cast<const float2*>(base)[idx];
        return {v.x, v.y};
    }
}
template <class S>
__device__ inline void stz2(void* base, long long idx, Z2<S> v) {
    if constexpr (sizeof(S) == 8) reinterpret_cast<double2*>(base)[idx] = make_double2(v.re, v.im);
    else reinterpret_cast<float2*>(base)[idx] = make_float2(v.re, v.im);
}

template <class S>
__device__ inline typename ChainT<S>::Acc cmfma(S a, S b, typename ChainT<S>::Acc c) {
    if constexpr (sizeof(S) == 8) return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// jobs[s * nchain + i]: step s of chain i (real B, the same B, rows, cols and
// K in every job).  sync: [ngroups] step counters, then [ngroups] finish
// counters, zero on entry and on exit.
template <class S, int PH>
__global__ void __launch_bounds__(CT) k_chain(const BJob* __restrict__ jobs, int steps, int nchain, int nrt, int nct,
                                              unsigned* __restrict__ sync, int* __restrict__ err, int spin) {
    using Acc = typename ChainT<S>::Acc;
    constexpr int KP = ChainT<S>::KMAX + 1;
    __shared__ S sE[CBN][KP];
    __shared__ S sRed[8 * 2 * CBM * CBM * 2];   // per-wave partial products
    __shared__ int s_abort;
    __shared__ unsigned s_cnt;
    const int ngroups = nchain * nrt;
    const int b = blockIdx.x;
    const int g = (b & 7) + 8 * ((b >> 3) / nct);     // group (chain, row block)
    const int m = (b >> 3) % nct;                     // column tile
    if (g >= ngroups) return;
    const int chain = g / nrt, rt = g % nrt;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int r16 = lane & 15, q = lane >> 4;
    const BJob& J0 = jobs[chain];
    const int rows = J0.rows, cols = J0.cols, K = J0.K;
    const int r0 = rt * CBM, c0 = m * CBN;
    const int nks = (K + 3) >> 2, Kp = nks * 4;
    unsigned* cnt = sync + g;
    unsigned* fin = sync + ngroups + g;

    // this tile's columns of B, zero beyond cols and K (one column per 16 lanes)
    {
        const int c = tid >> 4;
        const bool live = c0 + c < cols;
        const S* Bp = reinterpret_cast<const S*>(J0.B) + (long long)(c0 + c) * J0.ldb;
        for (int k = tid & 15; k < Kp; k += 16) sE[c][k] = (live && k < K) ? Bp[k] : S(0);
    }
    if (tid == 0) s_abort = 0;
    const bool placed = (unsigned)__builtin_amdgcn_s_getreg(HWREG_XCC_ID) == (unsigned)(b & 7);
    bool fenced = true;     // L2 write-back before each signal, until the group is known to share one L2

    __syncthreads();   // sE complete before the first step's products

    constexpr int JN = (ChainT<S>::KMAX / 4 + 7) / 8;  // K steps of 4 per wave per step
    const int r = r0 + r16;                              // this lane's A row
    const bool rlive = r < rows;
    const int orow = tid >> 5, ocol = tid & 31;          // epilogue: one output per thread
    for (int s = 0; s < steps; ++s) {
        const BJob& J = jobs[(long long)s * nchain + chain];
        // everything the step reads that no step writes, loaded before the
        // wait: the A-row phases (PH 1), the Add element and its phase (PH 2)
        Z2<S> ph[JN];
        if constexpr (PH == 1) {
            const long long prow = rlive ? (long long)(r / J.phdiv) * J.ldph : 0;
            #pragma unroll
            for (int j = 0; j < JN; ++j) {
                const int k = 4 * (w + 8 * j) + q;
                ph[j] = (rlive && k < K) ? ldz2<S>(J.ph, prow + k) : Z2<S>{S(0), S(0)};
            }
        }
        const int orr = r0 + orow, occ = c0 + ocol;
        const bool olive = orr < rows && occ < cols;
        Z2<S> add{S(0), S(0)}, oph{S(1), S(0)};
        if (olive) {
            if (J.Add) add = ldz2<S>(J.Add, (long long)orr * J.ldadd + occ);
            if constexpr (PH == 2) oph = ldz2<S>(J.ph, (long long)(orr / J.phdiv) * J.ldph + occ);
        }
        if (s > 0) {
            // wait until the group's nct tiles of step s - 1 are written
            if (tid == 0) {
                const unsigned target = (unsigned)s * (unsigned)nct;
                int polls = 0;
                unsigned v;
                // relaxed polls: an acquire per poll would invalidate the L2
                // on every iteration
                // (a bound of 0 gives up at once: the host's test of this path)
                while (((v = __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) & ARRIVALS) < target ||
                       spin == 0) {
                    if (++polls > spin || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        s_abort = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
                s_cnt = v;
            }
            __syncthreads();
            if (s_abort) break;
            // every member's step-0 signal (with its mark) is in the count
            if (s == 1) fenced = !placed || (s_cnt & ~ARRIVALS) != 0;
            if (w == 0) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
            }
            __syncthreads();
        }
        // this wave's K steps ks = w + 8 j of the 16 A rows, straight into
        // registers (all loads in flight at once; each product waits only for
        // its own operand), phase applied (PH 1), against both column blocks
        Z2<S> z[JN];
        #pragma unroll
        for (int j = 0; j < JN; ++j) {
            const int k = 4 * (w + 8 * j) + q;
            z[j] = Z2<S>{S(0), S(0)};
            if (rlive && k < K)
                z[j] = k < J.kseg ? ldz2<S>(J.A0, (long long)r * J.lda0 + k)
                                  : ldz2<S>(J.A1, (long long)r * J.lda1 + (k - J.kseg));
        }
        Acc are0 = {0, 0, 0, 0}, aim0 = {0, 0, 0, 0}, are1 = {0, 0, 0, 0}, aim1 = {0, 0, 0, 0};
        #pragma unroll
        for (int j = 0; j < JN; ++j) {
            const int ks = w + 8 * j;
            if (ks < nks) {
                const int k = 4 * ks + q;
                Z2<S> v = z[j];
                if constexpr (PH == 1) {
                    v.re = ph[j].re * z[j].re - ph[j].im * z[j].im;
                    v.im = ph[j].re * z[j].im + ph[j].im * z[j].re;
                }
                const S b0 = sE[r16][k], b1 = sE[16 + r16][k];
                are0 = cmfma<S>(v.re, b0, are0);
                aim0 = cmfma<S>(v.im, b0, aim0);
                are1 = cmfma<S>(v.re, b1, are1);
                aim1 = cmfma<S>(v.im, b1, aim1);
            }
        }
        // partials [w][cb][row][col][re, im]; the previous step's reads of
        // sRed ended at its closing barrier
        #pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = (sizeof(S) == 8) ? 4 * v + q : 4 * q + v;
            const int o0 = ((((w * 2 + 0) * 16 + row) * 16) + r16) * 2;
            const int o1 = ((((w * 2 + 1) * 16 + row) * 16) + r16) * 2;
            sRed[o0] = are0[v];
            sRed[o0 + 1] = aim0[v];
            sRed[o1] = are1[v];
            sRed[o1 + 1] = aim1[v];
        }
        __syncthreads();
        {
            const int cbx = ocol >> 4, cc = ocol & 15;
            S re = S(0), im = S(0);
            #pragma unroll
            for (int k8 = 0; k8 < 8; ++k8) {
                const int o = ((((k8 * 2 + cbx) * 16 + orow) * 16) + cc) * 2;
                re += sRed[o];
                im += sRed[o + 1];
            }
            if (olive) {
                if constexpr (PH == 2) {
                    const S tr = oph.re * re + oph.im * im;   // conj(p) * (re + i im)
                    const S ti = oph.re * im - oph.im * re;
                    re = tr;
                    im = ti;
                }
                const S sg = S(J.sgn);
                const Z2<S> out{sg * re + add.re, sg * im + add.im};
                stz2<S>(J.Y, (long long)orr * J.ldy + occ, out);
            }
        }
        // publish the tile: every wave's stores complete, then one signal
        // (the barrier also ends this step's reads of sRed)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
        __syncthreads();
        if (s + 1 < steps && tid == 0) {
            if (fenced) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            const unsigned add = 1u + ((s == 0 && !placed) ? MISPLACED : 0u);
            __hip_atomic_fetch_add(cnt, add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // the group's last workgroup to finish clears its counters
    __syncthreads();
    if (tid == 0) {
        const unsigned d = __hip_atomic_fetch_add(fin, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (d + 1 == (unsigned)nct) {
            __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(fin, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <class S>
void chain_launch_s(int ph, dim3 grid, hipStream_t s, const BJob* j, int steps, int nchain, int nrt, int nct,
                    unsigned* sync, int* err, int spin) {
    if (ph == 0) launch_timed(k_chain<S, 0>, grid, dim3(CT), 0, s, j, steps, nchain, nrt, nct, sync, err, spin);
    else if (ph == 1) launch_timed(k_chain<S, 1>, grid, dim3(CT), 0, s, j, steps, nchain, nrt, nct, sync, err, spin);
    else launch_timed(k_chain<S, 2>, grid, dim3(CT), 0, s, j, steps, nchain, nrt, nct, sync, err, spin);
}

// NSLX_CHAIN_SPIN overrides the poll bound; 0 makes every wait give up at
// once (how the tests force the error path: every workgroup still leaves
// normally)
int chain_spin() {
    const char* e = std::getenv("NSLX_CHAIN_SPIN");
    if (!e || !e[0]) return CHAIN_SPIN;
    const long v = std::strtol(e, nullptr, 10);
    return v < 0 ? 0 : (v > CHAIN_SPIN ? CHAIN_SPIN : (int)v);
}

int chain_grid(int nchain, int rows, int cols) {
    const int nrt = (rows + CBM - 1) / CBM, nct = (cols + CBN - 1) / CBN;
    const int ng = nchain * nrt;
    return 8 * ((ng + 7) / 8) * nct;
}

}  // namespace

std::string chain_class(const BVariant& v) {
    return std::string("chain_") + (v.prec == NSLX_F64 ? "f64" : "f32") + "_ph" + std::to_string(v.ph);
}

bool chain_supported(const BVariant& v, int steps, int nchain, int rows, int cols, int K) {
    if (v.cb || (v.prec != NSLX_F64 && v.prec != NSLX_F32) || nchain < 1 || rows < 1 || cols < 1 || K < 1) return false;
    if ((long long)steps * ((cols + CBN - 1) / CBN) > (long long)ARRIVALS) return false;
    if (K > (v.prec == NSLX_F64 ? ChainT<double>::KMAX : ChainT<float>::KMAX)) return false;
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return false;
    const hipError_t e = v.prec == NSLX_F64
                             ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_chain<double, 1>, CT, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_chain<float, 1>, CT, 0);
    if (e != hipSuccess || per < 1) return false;
    // one workgroup per CU, the grid at most one XCD's share of CUs per XCD
    const int grid = chain_grid(nchain, rows, cols);
    return grid <= (ncu / 8) * 8;
}

size_t chain_sync_words(int nchain, int rows) { return 2 * (size_t)nchain * ((rows + CBM - 1) / CBM); }

void launch_chain(const BVariant& v, const BJob* d_jobs, int steps, int nchain, int rows, int cols, unsigned* sync,
                  int* err, hipStream_t s, double flops, double bytes) {
    if (steps <= 0) return;
    ProfScope prof(prof_on() ? chain_class(v) : std::string(), s);
    prof.work(flops, bytes);
    const int nrt = (rows + CBM - 1) / CBM, nct = (cols + CBN - 1) / CBN;
    const dim3 grid(chain_grid(nchain, rows, cols));
    const int spin = chain_spin();
    if (v.prec == NSLX_F64) chain_launch_s<double>(v.ph, grid, s, d_jobs, steps, nchain, nrt, nct, sync, err, spin);
    else chain_launch_s<float>(v.ph, grid, s, d_jobs, steps, nchain, nrt, nct, sync, err, spin);
    NSLX_LAUNCH_CHECK();
}

}  // namespace nslx
