// log det M and its gradient (the HMC force kernel) for HubbardExp, FP64.
//
// logDetM (reference src/NSL/FermionMatrix/Impl/hubbardExp.tpp:128-150)
// forms det(1 + B_{Nt-1}...B_0), B_t = E(s delta) diag(P_{t+1}), which
// overflows FP64 once log det > 709 (the 12x12 honeycomb at beta = 8 has
// log det ~ 1815).  Here the Nt slices are grouped into chunks whose products
// C_j stay well conditioned, and
//      det M = det(1 + C_{n-1} ... C_0) = det R,
// R = block-cyclic matrix (I on the diagonal, -C_j below it, +C_0 in the
// corner), factorised by LU with partial pivoting; log det = sum log u_kk.
// With one chunk this is the reference's formula.
//
// gradLogDetM follows hubbardExp.tpp:152-211 literally: Fk_t =
// diag(Pinv_{t-1}) E(-s delta), suffix products Q_t = Fk_t Q_{t+1},
// G = (1 + Q_0)^{-1}, pi_{Nt-1} = i diag(Q_0 G), then G <- G Fk_t and
// pi_t = i diag(Q_{t+1} G).
//
// HubbardDiag (hubbardDiag.tpp:166-190, 193-252) uses the same two
// algorithms with D = 1 - s delta K in place of E: logDetM = i s sum(phi) +
// log det(1 + fk_0 fk_1 ... fk_{Nt-1}), fk_t = D diag(Pinv_{t+1}) (the
// product built left to right from t = 0, no mu term, as the reference);
// gradLogDetM with Fk_t = diag(Pinv_{t-1}) D.
#include <cmath>

#include "engine.hpp"

namespace nslx {

namespace {
int auto_chunks(FermionMatrix& fm, int64_t field, hipStream_t s) {
    // growth of the slice product ~ exp(beta ||K|| + max_x sum_t |Im phi|)
    std::vector<double2> h((size_t)fm.len());
    NSLX_HIP(hipMemcpyAsync(h.data(), fm.phi.p + field * fm.len(), h.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    double im = 0.0;
    for (int x = 0; x < fm.nx; ++x) {
        double a = 0.0;
        for (int t = 0; t < fm.nt; ++t) a += std::fabs(h[(size_t)t * fm.nx + x].y);
        im = std::max(im, a);
    }
    const double growth = std::fabs(fm.beta) * fm.normK + im + fm.mut_abs() * fm.nt;
    int n = (int)std::ceil(growth / 8.0);
    return std::max(1, std::min(n, fm.nt));
}
}  // namespace

void fm_logdet(FermionMatrix& fm, int64_t field, int nchunks, double* out2, hipStream_t s) {
    fm.require_populated();
    NSLX_REQUIRE(field >= 0 && field < fm.nfield, "logdet: field index out of range");
    NSLX_REQUIRE(out2, "logdet: null output");
    const int nx = fm.nx, nt = fm.nt;
    if (nchunks <= 0) nchunks = auto_chunks(fm, field, s);
    nchunks = std::max(1, std::min(nchunks, nt));
    const bool diag = fm.disc == DISC_DIAG;
    EEntry& e = diag ? fm.Elin(-fm.sgn() * fm.delta(), 1.0, s) : fm.E(fm.sgn() * fm.delta(), s);
    const double2* P = (diag ? fm.Pinv64.p : fm.P64.p) + field * fm.len();
    const size_t m = (size_t)nx * nx;
    DevBuf<double2> C(m * nchunks), tmp(m), cur(m);
    for (int j = 0; j < nchunks; ++j) {
        const int t_lo = (int)std::llround((double)j * nt / nchunks);
        const int t_hi = (int)std::llround((double)(j + 1) * nt / nchunks);
        set_identity_c(nx, cur.p, s);
        if (diag) {
            // cur = cur * D * diag(Pinv_{t+1}) for t ascending (hubbardDiag.tpp:176-182)
            for (int t = t_lo; t < t_hi; ++t) {
                const double2* Pn = P + (size_t)((t + 1) % nt) * nx;
                gemm_sq(nx, cur.p, 1, e.dense.p, 0, tmp.p, nullptr, nullptr, Pn, 1.0, 0.0, s);
                std::swap(cur.p, tmp.p);
            }
        } else {
            for (int t = t_hi - 1; t >= t_lo; --t) {
                // cur = cur * E * diag(P_{t+1})   (sausage.mat_mul(prod[t]))
                const double2* Pn = P + (size_t)((t + 1) % nt) * nx;
                gemm_sq(nx, cur.p, 1, e.dense.p, 0, tmp.p, nullptr, nullptr, Pn, 1.0, 0.0, s);
                std::swap(cur.p, tmp.p);
            }
        }
        // build_cyclic forms det(1 + C_{n-1} ... C_0): the ascending Diag
        // product C_0 ... C_{n-1} is stored in reverse chunk order
        const int slot = diag ? nchunks - 1 - j : j;
        NSLX_HIP(hipMemcpyAsync(C.p + (size_t)slot * m, cur.p, m * sizeof(double2), hipMemcpyDeviceToDevice, s));
    }
    const int N = nx * nchunks;
    DevBuf<double2> R((size_t)N * N);
    build_cyclic(nx, nchunks, C.p, R.p, s);
    DevBuf<double> acc(2);
    NSLX_HIP(hipMemsetAsync(acc.p, 0, 2 * sizeof(double), s));
    lu_logdet(N, R.p, acc.p, s);
    double h[2];
    NSLX_HIP(hipMemcpyAsync(h, acc.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    out2[0] = h[0];
    double a = std::remainder(h[1], 2.0 * M_PI);   // (-pi, pi]
    if (a <= -M_PI) a += 2.0 * M_PI;
    out2[1] = a;
    if (diag) {
        // + i s sum(phi) (hubbardDiag.tpp:186), outside the principal branch
        std::vector<double2> ph((size_t)fm.len());
        NSLX_HIP(hipMemcpyAsync(ph.data(), fm.phi.p + field * fm.len(), ph.size() * sizeof(double2),
                                hipMemcpyDeviceToHost, s));
        NSLX_HIP(hipStreamSynchronize(s));
        double sr = 0.0, si = 0.0;
        for (const double2& z : ph) { sr += z.x; si += z.y; }
        out2[0] -= fm.sgn() * si;
        out2[1] += fm.sgn() * sr;
    }
}

void fm_grad_logdet(FermionMatrix& fm, int64_t field, void* out, hipStream_t s) {
    fm.require_populated();
    NSLX_REQUIRE(field >= 0 && field < fm.nfield, "grad_logdet: field index out of range");
    NSLX_REQUIRE(out, "grad_logdet: null output");
    const int nx = fm.nx, nt = fm.nt;
    EEntry& en = fm.disc == DISC_DIAG ? fm.Elin(-fm.sgn() * fm.delta(), 1.0, s) : fm.E(-fm.sgn() * fm.delta(), s);
    const double2* Pi = fm.Pinv64.p + field * fm.len();
    auto Pprev = [&](int t) { return Pi + (size_t)((t - 1 + nt) % nt) * nx; };   // Pinv_{t-1}
    const size_t m = (size_t)nx * nx;
    DevBuf<double2> Q(m * nt), G(m), tmp(m), aug(2 * m), work((size_t)nx);
    // Q_{Nt-1} = Fk_{Nt-1}
    dscale_rows(nx, Pprev(nt - 1), en.dense.p, Q.p + (size_t)(nt - 1) * m, s);
    for (int t = nt - 2; t >= 0; --t)
        gemm_sq(nx, en.dense.p, 0, Q.p + (size_t)(t + 1) * m, 1, Q.p + (size_t)t * m, Pprev(t), nullptr, nullptr, 1.0,
                0.0, s);
    NSLX_HIP(hipMemcpyAsync(G.p, Q.p, m * sizeof(double2), hipMemcpyDeviceToDevice, s));
    add_identity_c(nx, G.p, s);
    gj_inverse(nx, G.p, aug.p, work.p, s);
    double2* pi = reinterpret_cast<double2*>(out);
    diag_iprod(nx, Q.p, G.p, pi + (size_t)(nt - 1) * nx, s);
    for (int t = 0; t < nt - 1; ++t) {
        // G = G diag(Pinv_{t-1}) E(-s delta)
        gemm_sq(nx, G.p, 1, en.dense.p, 0, tmp.p, nullptr, Pprev(t), nullptr, 1.0, 0.0, s);
        std::swap(G.p, tmp.p);
        diag_iprod(nx, Q.p + (size_t)(t + 1) * m, G.p, pi + (size_t)t * nx, s);
    }
    NSLX_HIP(hipStreamSynchronize(s));   // temporaries are freed on return
}

}  // namespace nslx
