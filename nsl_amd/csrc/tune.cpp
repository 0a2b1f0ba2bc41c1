// Hop tile autotuner: times every tile variant that fits the lattice on the
// fermion matrix's populated fields and records the fastest per
// (precision, nx, nt, nbatch class) in the context's tile table.
#include <limits>

#include "engine.hpp"

namespace nslx {

int fm_tune(FermionMatrix& fm, int prec, int64_t nb, int reps, int* best_tile, double* best_us, hipStream_t s) {
    NSLX_REQUIRE(prec == NSLX_F64 || prec == NSLX_F32 || prec == NSLX_BF16X3, "tune: bad precision");
    NSLX_REQUIRE(reps >= 1, "tune: reps must be >= 1");
    HopArgs a = fm.hop_args(prec, nb, s);
    const size_t esz = prec == NSLX_F64 ? 16 : 8;
    const size_t bytes = (size_t)nb * fm.len() * esz;
    DevBuf<unsigned char> in(bytes), out(bytes);
    NSLX_HIP(hipMemsetAsync(in.p, 0, bytes, s));
    a.X = in.p;
    a.Y = out.p;
    hipEvent_t e0, e1;
    NSLX_HIP(hipEventCreate(&e0));
    NSLX_HIP(hipEventCreate(&e1));
    int best = -1;
    double best_t = std::numeric_limits<double>::infinity();
    try {
        const int n = (int)tile_menu(prec).size();
        for (int t = 0; t < n; ++t) {
            if (!tile_fits(prec, t, fm.nx)) continue;
            if (fm.disc == DISC_DIAG && t >= 4) continue;   // HubbardDiag: classic tiles only
            launch_hop(prec, HOP_M, MODE_PLAIN, a, t, s);   // warm-up
            launch_hop(prec, HOP_MD, MODE_PLAIN, a, t, s);
            NSLX_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) {
                launch_hop(prec, HOP_M, MODE_PLAIN, a, t, s);
                launch_hop(prec, HOP_MD, MODE_PLAIN, a, t, s);
            }
            NSLX_HIP(hipEventRecord(e1, s));
            NSLX_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            NSLX_HIP(hipEventElapsedTime(&ms, e0, e1));
            const double us = 1e3 * ms / reps;
            if (us < best_t) {
                best_t = us;
                best = t;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    NSLX_REQUIRE(best >= 0, "tune: no tile fits this lattice");
    fm.ctx->tiles.set(prec, fm.nx, fm.nt, nb, best);
    if (best_tile) *best_tile = best;
    if (best_us) *best_us = best_t;
    return NSLX_OK;
}

}  // namespace nslx
