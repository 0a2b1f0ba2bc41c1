// Batched time-slice-preconditioned solves (nslx_batch_solve_tss): many
// HubbardExp systems MM^dag x = b (or M^dag M x = b) of different Nx, Nt, beta and fields, each
// solved by solve_tss (CG preconditioned by its own time-slice solve, with
// double-double correction), spread over several host threads that each
// drive one HIP stream.  With 2-3 preconditioned iterations per system the
// cost is the factorisation and the preconditioner applies, which for small
// lattices are far from filling the GPU; running the systems' launch
// sequences side by side on separate streams is what lets them share it.
//
// Items are taken largest first (longest-processing-time order, by the
// factorisation plus apply work) from a shared counter, so the long items
// start early and the short ones fill in.  Each worker owns its stream; an
// item's computation does not depend on which worker runs it or what runs
// beside it.  The recurrences use one launch per slice (chains_ok off): the
// persistent recurrence launch needs the whole device.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <numeric>
#include <set>
#include <thread>

#include "engine.hpp"

namespace nslx {

namespace {
// relative work of one item: factorisation (Nt n^3 per field) + applies
double item_work(const FermionMatrix& fm, int64_t nb) {
    const double n = fm.nx, nt = fm.nt;
    return (double)fm.nfield * nt * n * n * n + 8.0 * (double)nb * nt * n * n;
}
}  // namespace

int batch_solve_tss(Ctx& ctx, int64_t nitems, const nslx_batch_item* items, void* const* x_lo, double tol,
                    int pprec, int64_t maxit, int nstreams, nslx_batch_report* rep) {
    const auto t0 = std::chrono::steady_clock::now();
    NSLX_REQUIRE(nitems >= 1 && items, "batch_solve_tss: no items");
    NSLX_REQUIRE(pprec == NSLX_F64 || pprec == NSLX_F32, "batch_solve_tss: precision must be F64 or F32");
    NSLX_REQUIRE(tol > 0.0 && maxit >= 1, "batch_solve_tss: bad tol/maxit");
    std::set<const void*> seen;
    for (int64_t i = 0; i < nitems; ++i) {
        const nslx_batch_item& it = items[i];
        NSLX_REQUIRE(it.fm && it.b && it.x && it.nbatch >= 1, "batch_solve_tss: incomplete item");
        NSLX_REQUIRE(it.op == NSLX_OP_MMDAGGER || it.op == NSLX_OP_MDAGGERM, "batch_solve_tss: op must be MMdagger or MdaggerM");
        FermionMatrix& fm = it.fm->m;
        NSLX_REQUIRE(fm.ctx == &ctx, "batch_solve_tss: an item's fermion matrix belongs to another context");
        NSLX_REQUIRE(fm.disc == DISC_EXP, "batch_solve_tss: HubbardExp items only");
        fm.require_populated();
        NSLX_REQUIRE(it.nbatch % fm.nfield == 0, "batch_solve_tss: nbatch must be a multiple of nfield");
        NSLX_REQUIRE(seen.insert(it.fm).second, "batch_solve_tss: a fermion matrix appears in two items");
    }
    if (nstreams <= 0) nstreams = 8;
    nstreams = (int)std::min<int64_t>(nstreams, nitems);
    std::vector<int64_t> order(nitems);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return item_work(items[a].fm->m, items[a].nbatch) > item_work(items[b].fm->m, items[b].nbatch);
    });
    // inputs written on any stream are complete before the workers start
    NSLX_HIP(hipDeviceSynchronize());
    std::vector<bool> chains(nitems);
    for (int64_t i = 0; i < nitems; ++i) {
        chains[i] = items[i].fm->m.chains_ok;
        items[i].fm->m.chains_ok = nstreams == 1 && chains[i];
    }

    std::vector<nslx_solve_report> reps(nitems);
    std::vector<int> status(nitems, NSLX_OK);
    std::vector<std::string> errs(nitems);
    std::atomic<int64_t> next{0};
    auto worker = [&]() {
        hipStream_t s = nullptr;
        if (hipSetDevice(ctx.device) != hipSuccess ||
            hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            for (int64_t k; (k = next++) < nitems;) {
                status[order[k]] = NSLX_E_DEVICE;
                errs[order[k]] = "batch_solve_tss: could not create a worker stream";
            }
            return;
        }
        for (int64_t k; (k = next++) < nitems;) {
            const int64_t i = order[k];
            const nslx_batch_item& it = items[i];
            try {
                status[i] = solve_tss(it.fm->m, it.op, it.b, it.x, x_lo ? x_lo[i] : nullptr, nullptr,
                                      it.nbatch, tol, maxit, pprec, 0, &reps[i], s);
            } catch (const Error& e) {
                status[i] = e.status;
                errs[i] = e.what();
            } catch (const std::exception& e) {
                status[i] = NSLX_E_INTERNAL;
                errs[i] = e.what();
            }
        }
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    };
    std::vector<std::thread> th;
    for (int w = 0; w < nstreams; ++w) th.emplace_back(worker);
    for (auto& t : th) t.join();
    for (int64_t i = 0; i < nitems; ++i) items[i].fm->m.chains_ok = chains[i];

    nslx_batch_report out{};
    bool all_ok = true;
    for (int64_t i = 0; i < nitems; ++i) {
        if (status[i] != NSLX_OK && status[i] != NSLX_NOT_CONVERGED)
            throw Error(status[i], "batch_solve_tss: item " + std::to_string(i) + ": " + errs[i]);
        const nslx_solve_report& r = reps[i];
        out.iterations = std::max(out.iterations, r.iterations);
        out.replays += r.iterations_total;      // preconditioner applications
        out.restarts += r.outer_steps;          // double-double residual evaluations
        out.max_rel_residual = std::max(out.max_rel_residual, r.max_rel_residual);
        if (status[i] == NSLX_OK) ++out.converged_items;
        else all_ok = false;
    }
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rep) *rep = out;
    return all_ok ? NSLX_OK : NSLX_NOT_CONVERGED;
}

}  // namespace nslx
