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
//
// The Nx x Nx products run on the FP64 matrix cores as launches of the
// batched GEMM (bgemm.hip): every chunk's product advances one slice per
// launch (all chunks in one launch), and the force's two sequential chains
// are one persistent recurrence launch each (chain.hip; one launch per slice
// where that launch does not fit the device or is not allowed), their
// diagonals one batched launch.  Both use that E (and D) are symmetric
// (K is), so E[k][c] = E[c][k]: a product X E diag(p) is the GEMM's
// Y = X B^T with B = E and the phase on the output columns.
#include <algorithm>
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
    // chunk j covers slices [t_lo_j, t_hi_j); its product starts at I and
    // takes one slice per launch: HubbardExp cur = cur E diag(P_{t+1}) for t
    // descending (sausage.mat_mul(prod[t])), HubbardDiag cur = cur D
    // diag(Pinv_{t+1}) for t ascending (hubbardDiag.tpp:176-182).  The GEMM's
    // output-column phase is conjugated, so it reads conj(P).
    std::vector<int> tlo(nchunks), len(nchunks);
    int maxlen = 0;
    for (int j = 0; j < nchunks; ++j) {
        tlo[j] = (int)std::llround((double)j * nt / nchunks);
        len[j] = (int)std::llround((double)(j + 1) * nt / nchunks) - tlo[j];
        maxlen = std::max(maxlen, len[j]);
    }
    DevBuf<double2> buf[2] = {DevBuf<double2>(m * nchunks), DevBuf<double2>(m * nchunks)};
    DevBuf<double2> Pc((size_t)fm.len());
    conj_c(fm.len(), P, Pc.p, s);
    for (int j = 0; j < nchunks; ++j) set_identity_c(nx, buf[0].p + (size_t)j * m, s);
    BPlan plan;
    for (int q = 0; q < maxlen; ++q) {
        std::vector<BJob> jobs;
        for (int j = 0; j < nchunks; ++j) {
            if (q >= len[j]) continue;
            const int t = diag ? tlo[j] + q : tlo[j] + len[j] - 1 - q;
            BJob b;
            b.A0 = buf[q & 1].p + (size_t)j * m;
            b.lda0 = nx;
            b.rows = b.cols = b.K = nx;
            b.B = e.dense.p;
            b.ldb = nx;
            b.Y = buf[(q + 1) & 1].p + (size_t)j * m;
            b.ldy = nx;
            b.ph = Pc.p + (size_t)((t + 1) % nt) * nx;
            b.ldph = 0;
            jobs.push_back(b);
        }
        plan.gemm(jobs, BVariant{NSLX_F64, false, 2, 0});
    }
    plan.upload(s);
    plan.run(s);
    // build_cyclic forms det(1 + C_{n-1} ... C_0): the ascending Diag
    // product C_0 ... C_{n-1} is stored in reverse chunk order
    DevBuf<double2> C(m * nchunks);
    for (int j = 0; j < nchunks; ++j) {
        const int slot = diag ? nchunks - 1 - j : j;
        NSLX_HIP(hipMemcpyAsync(C.p + (size_t)slot * m, buf[len[j] & 1].p + (size_t)j * m, m * sizeof(double2),
                                hipMemcpyDeviceToDevice, s));
    }
    const int N = nx * nchunks;
    DevBuf<double2> R((size_t)N * N), Rw((size_t)N * N);
    build_cyclic(nx, nchunks, C.p, R.p, s);
    DevBuf<double> acc(2);
    NSLX_HIP(hipMemsetAsync(acc.p, 0, 2 * sizeof(double), s));
    lu_logdet(N, R.p, Rw.p, acc.p, s);
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
    // The chain is kept transposed, QT_t = Q_t^T = QT_{t+1} E diag(Pinv_{t-1})
    // (E symmetric): a GEMM with the phase on its output columns (given
    // conjugated).  diag(Q_{t+1} G)[x] = sum_k QT_{t+1}[k][x] G[k][x].
    // QT holds every QT_t; Gs every G_t (slot t: G after t steps), so that
    // both recurrences are one persistent launch each and the diagonals one
    // batched launch after them
    DevBuf<double2> QT(m * nt), Gs(m * nt), H(m), aug(2 * m), work(2 * m), Pic((size_t)fm.len());
    conj_c(fm.len(), Pi, Pic.p, s);
    auto Pcprev = [&](int t) { return Pic.p + (size_t)((t - 1 + nt) % nt) * nx; };
    const double* E = en.dense.p;
    // QT_{Nt-1} = Fk_{Nt-1}^T = E diag(Pinv_{Nt-2})
    dscale_cols(nx, E, Pprev(nt - 1), QT.p + (size_t)(nt - 1) * m, s);
    std::vector<std::vector<BJob>> steps_q, steps_g;
    for (int t = nt - 2; t >= 0; --t) {
        BJob b;
        b.A0 = QT.p + (size_t)(t + 1) * m;
        b.lda0 = nx;
        b.rows = b.cols = b.K = nx;
        b.B = E;
        b.ldb = nx;
        b.Y = QT.p + (size_t)t * m;
        b.ldy = nx;
        b.ph = Pcprev(t);
        b.ldph = 0;
        steps_q.push_back({b});
    }
    // G <- G diag(Pinv_{t-1}) E(-s delta) (the phase on the GEMM's A columns),
    // pi_t = i diag(Q_{t+1} G_{t+1})
    for (int t = 0; t < nt - 1; ++t) {
        BJob b;
        b.A0 = Gs.p + (size_t)t * m;
        b.lda0 = nx;
        b.rows = b.cols = b.K = nx;
        b.B = E;
        b.ldb = nx;
        b.Y = Gs.p + (size_t)(t + 1) * m;
        b.ldy = nx;
        b.ph = Pprev(t);
        b.ldph = 0;
        steps_g.push_back({b});
    }
    double2* pi = reinterpret_cast<double2*>(out);
    // a recurrence launch that gave up waiting for its workgroups leaves its
    // input (QT_{Nt-1}, G_0) intact: the recurrence is run again one launch
    // per slice
    auto run_chain = [&](const std::vector<std::vector<BJob>>& steps, int ph) {
        for (int attempt = 0; attempt < 2; ++attempt) {
            BPlan plan;
            plan.chain(steps, BVariant{NSLX_F64, false, ph, 0}, attempt == 0 && fm.chains_ok);
            plan.upload(s);
            plan.run(s);
            if (!plan.chain_failed(s)) {
                NSLX_HIP(hipStreamSynchronize(s));   // the plan's device arrays are freed on return
                return;
            }
        }
    };
    if (!steps_q.empty()) run_chain(steps_q, 2);
    // G_0 = (1 + Q_0)^{-1} = ((1 + QT_0)^{-1})^T
    NSLX_HIP(hipMemcpyAsync(H.p, QT.p, m * sizeof(double2), hipMemcpyDeviceToDevice, s));
    add_identity_c(nx, H.p, s);
    gj_inverse(nx, H.p, aug.p, work.p, s);
    transpose_c(nx, H.p, Gs.p, s);
    diag_iprod_cols(nx, QT.p, Gs.p, pi + (size_t)(nt - 1) * nx, s);
    if (!steps_g.empty()) {
        run_chain(steps_g, 1);
        diag_iprod_cols_batched(nx, nt - 1, QT.p + m, Gs.p + m, (int64_t)m, pi, nx, s);
    }
    NSLX_HIP(hipStreamSynchronize(s));   // temporaries and the plans' device arrays are freed on return
}

}  // namespace nslx
