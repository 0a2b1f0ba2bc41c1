// Double-double residuals and solution updates (dd.hip).
#pragma once

#include "common.hpp"

namespace nslx {

struct FermionMatrix;

struct DdWork {
    DevBuf<double2> uh, ul;    // the first hop's double-double output
    DevBuf<double> part;       // per (row, column block) |r|^2 partials
    DevBuf<double> et;         // transpose of the E it was built from
    DevBuf<double> enorm;      // max row sum of |E| (one value)
    DevBuf<double> sig;        // per input row: the accumulation grid of its hop
    const double* et_src = nullptr;
    void ensure(int64_t nb, int64_t len, int n, int nt);
};

// r = b - A (xh + xl) with A = M M^dag (op NSLX_OP_MMDAGGER) or M^dag M
// (NSLX_OP_MDAGGERM) of a HubbardExp fm, every product and sum in
// double-double; r rounded to c128, rn2[k] = |r_k|^2 (device).  xl may be
// null (x = xh exactly).  All device buffers, nb x nt x nx c128.
void dd_residual(FermionMatrix& fm, int op, const void* b, const void* xh, const void* xl, void* r, double* rn2,
                 int64_t nb, DdWork& w, hipStream_t s);
// (xh, xl) += d in double-double for the RHS with done[k] < 0 (done may be null).
void dd_accumulate(void* xh, void* xl, const void* d, int64_t nb, int64_t len, const int* done, hipStream_t s);

}  // namespace nslx
