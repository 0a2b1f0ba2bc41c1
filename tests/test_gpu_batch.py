"""Ragged batched CG (nslx_batch_solve): many fermion matrices of different
Nx, Nt and fields in one grouped launch per Krylov iteration, with and
without HIP-graph replay.  Tolerances: every RHS to its FP64 true residual
target max(tol, attainable accuracy), checked with the oracle's own
MMdagger; an item's solution is bit-identical alone and in a mix, and with
and without the graph."""

import math

import pytest
import torch

import nsl_amd as nsl
from oracle import fermion as F, lattice as olat, rng

pytestmark = pytest.mark.gpu

DEV = "cuda"


def make(kind, size, Nt, beta, U, nfield, nrhs, seed, ctx):
    if kind == "honeycomb":
        lat_d, lat_o = nsl.Lattice.honeycomb(*size), olat.honeycomb(*size)
    elif kind == "ring":
        lat_d, lat_o = nsl.Lattice.ring(size), olat.ring(size)
    else:
        lat_d, lat_o = nsl.Lattice.armchair_nanotube(*size), olat.armchair_nanotube(*size)
    Nx = lat_d.sites()
    phis = [torch.as_tensor(rng.real_field(seed + f, 0, (Nt, Nx), math.sqrt(beta / Nt * U))) for f in range(nfield)]
    fm = nsl.HubbardExp(lat_d, Nt, beta, 0.0, ctx=ctx)
    fm.populate(torch.stack(phis).to(DEV))
    ofs = []
    for p in phis:
        o = F.HubbardExp(lat_o, Nt, beta)
        o.populate(p)
        ofs.append(o)
    b = torch.as_tensor(rng.complex_normal(seed + 500, 0, (nfield * nrhs, Nt, Nx)))
    return fm, ofs, b


SPECS = [
    ("honeycomb", (3, 3), 16, 4.0, 2.0, 2, 2, 100),   # Nx 18
    ("honeycomb", (6, 6), 32, 6.0, 2.0, 1, 3, 200),   # Nx 72
    ("nanotube", (3, 4), 20, 3.0, 1.5, 1, 2, 300),    # Nx 48
    ("ring", 10, 12, 2.0, 3.0, 2, 1, 400),            # Nx 10
]


def check_residuals(ofs, b, x, target, name="MMdagger"):
    rpf = b.shape[0] // len(ofs)
    for k in range(b.shape[0]):
        r = b[k] - getattr(ofs[k // rpf], name)(x[k].cpu())
        assert float(r.norm() / b[k].norm()) < 2 * target, k


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32, nsl.BF16X3])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_true_residual(prec, use_graph):
    ctx = nsl.Context(0)
    probs = [make(*s, ctx=ctx) for s in SPECS]
    xs, rep = nsl.batch_solve([(fm, b.to(DEV)) for fm, _, b in probs], tol=1e-10, precision=prec,
                              use_graph=use_graph, ctx=ctx)
    assert rep["converged"], rep
    assert rep["converged_items"] == len(SPECS)
    assert rep["iterations"] > 0 and rep["replays"] > 0
    target = max(1e-10, rep["max_floor_rel"])
    assert rep["max_rel_residual"] <= target
    for (fm, ofs, b), x in zip(probs, xs):
        check_residuals(ofs, b, x, target)


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32])
def test_batch_mdagm_items_in_the_mix(prec):
    """MdaggerM systems (the two-body correlator's second CG,
    twoBodyCorrelationFunction.tpp:256-283) batched beside MMdagger ones:
    each item checked with the oracle's own operator, and an MdaggerM item
    bit-identical alone and in the mix."""
    ctx = nsl.Context(0)
    probs = [make(*s, ctx=ctx) for s in SPECS]
    ops = [nsl.MdaggerM, nsl.MMdagger, nsl.MdaggerM, nsl.MMdagger]
    xs, rep = nsl.batch_solve([(fm, b.to(DEV), op) for (fm, _, b), op in zip(probs, ops)], tol=1e-10,
                              precision=prec, ctx=ctx)
    assert rep["converged"] and rep["converged_items"] == len(SPECS), rep
    target = max(1e-10, rep["max_floor_rel"])
    for (fm, ofs, b), x, op in zip(probs, xs, ops):
        check_residuals(ofs, b, x, target, "MdaggerM" if op == nsl.MdaggerM else "MMdagger")
    alone, _ = nsl.batch_solve([(probs[0][0], probs[0][2].to(DEV), nsl.MdaggerM)], tol=1e-10, precision=prec,
                               ctx=ctx)
    assert torch.equal(alone[0], xs[0])


def test_batch_tss_mdagm_items():
    """The time-slice-preconditioned batch with MdaggerM and MMdagger items:
    every RHS to 1e-12 in double-double, re-checked on the CPU."""
    from tests import ddarith as dd
    ctx = nsl.Context.default()
    probs = [make(*s, ctx=ctx) for s in SPECS[:3]]
    ops = [nsl.MdaggerM, nsl.MMdagger, nsl.MdaggerM]
    xs, los, rep = nsl.batch_solve_tss([(fm, b.to(DEV), op) for (fm, _, b), op in zip(probs, ops)], tol=1e-12,
                                       nstreams=3, return_lo=True)
    assert rep["converged"], rep
    for (fm, ofs, b), x, lo, op in zip(probs, xs, los, ops):
        E, P = fm.export()
        _, relv = dd.mmdag_residual(E, P, b.numpy(), x.cpu().numpy(), lo.cpu().numpy(), rpf=b.shape[0] // len(ofs),
                                    op="MdaggerM" if op == nsl.MdaggerM else "MMdagger")
        assert relv.max() <= 1.05e-12, relv


def test_batch_alone_equals_in_mix():
    ctx = nsl.Context(0)
    probs = [make(*s, ctx=ctx) for s in SPECS]
    mine = probs[1]
    alone, _ = nsl.batch_solve([(mine[0], mine[2].to(DEV))], tol=1e-10, ctx=ctx)
    mix, _ = nsl.batch_solve([(fm, b.to(DEV)) for fm, _, b in probs], tol=1e-10, ctx=ctx)
    nograph, _ = nsl.batch_solve([(fm, b.to(DEV)) for fm, _, b in probs[::-1]], tol=1e-10, use_graph=False, ctx=ctx)
    assert torch.equal(alone[0], mix[1])
    assert torch.equal(alone[0], nograph[2])


def test_batch_matches_single_solver():
    """Against the single-problem per-RHS solver (reliable-update FP64):
    both reach the FP64 true residual target, so the solutions agree to
    target x cond; here 1e-6 relative."""
    ctx = nsl.Context(0)
    fm, ofs, b = make(*SPECS[1], ctx=ctx)
    xs, rep = nsl.batch_solve([(fm, b.to(DEV))], tol=1e-11, ctx=ctx)
    cg = nsl.CG(fm, nsl.MMdagger, 1e-11, 100000, tol_kind=nsl.TOL_REL_PER_RHS)
    x1 = cg(b.to(DEV))
    err = float((xs[0] - x1).abs().max() / x1.abs().max())
    assert err < 1e-6, err


def test_batch_errors():
    ctx = nsl.Context(0)
    fm, ofs, b = make(*SPECS[0], ctx=ctx)
    with pytest.raises(nsl.NslxError):                       # nbatch not a multiple of the fields
        nsl.batch_solve([(fm, b[:3].to(DEV))], ctx=ctx)
    other = nsl.Context(0)
    with pytest.raises(nsl.NslxError):                       # fm of another context
        nsl.batch_solve([(fm, b.to(DEV))], ctx=other)
    with pytest.raises(nsl.NslxError):                       # M is not a batched system
        nsl.batch_solve([(fm, b.to(DEV), nsl.M)], ctx=ctx)
    fd = nsl.HubbardDiag(nsl.Lattice.honeycomb(3, 3), 16, 4.0, ctx=ctx)
    fd.populate(torch.zeros((16, 18), dtype=torch.complex128, device=DEV))
    with pytest.raises(nsl.NslxError):                       # HubbardDiag items are MMdagger systems
        nsl.batch_solve([(fd, b[:2].to(DEV), nsl.MdaggerM)], ctx=ctx)


def test_batch_large_lattice_grows_partials():
    """One Nx = 1152 (24 x 24 honeycomb), Nt = 64 item with 4 RHS: the 16 x 32
    tile it gets needs 4 * 36 = 144 partial-sum slots per RHS, more than the
    pool's initial 128 per RHS slot, so the partial arrays must grow.  Every
    RHS to 1e-8 (F32 recurrence, FP64 reliable updates); one RHS checked
    against the oracle's MMdagger."""
    ctx = nsl.Context(0)
    fm, ofs, b = make("honeycomb", (24, 24), 64, 2.0, 1.0, 1, 4, 500, ctx=ctx)
    xs, rep = nsl.batch_solve([(fm, b.to(DEV))], tol=1e-8, precision=nsl.F32, ctx=ctx)
    assert rep["converged"], rep
    assert rep["max_rel_residual"] <= 1e-8
    r = b[0] - ofs[0].MMdagger(xs[0][0].cpu())
    assert float(r.norm() / b[0].norm()) < 2e-8


# ---- time-slice-preconditioned batch (nslx_batch_solve_tss) -----------------

def test_batch_tss_full_mix():
    """The batching form's full mix (bench.batch_mix_specs: 64 fields,
    honeycomb Nx 18-1152, (6,6) nanotubes of 4-32 cells, Nt 16-128,
    beta 2-12, U 1-4, 1-16 RHS) on 8 worker streams: every item converges,
    every RHS to 1e-10 in double-double; the six smallest items' (x, x_lo)
    are re-checked on the CPU in double-double from the exported operator."""
    import bench
    from tests import ddarith as dd
    probs = bench.build_mix_problems(bench.batch_mix_specs(64, 7000))
    xs, los, rep = nsl.batch_solve_tss([(fm, b) for fm, b, _ in probs], tol=1e-10, precision=nsl.F64, nstreams=8,
                                       return_lo=True)
    assert rep["converged"] and rep["converged_items"] == 64, rep
    assert rep["max_rel_residual"] <= 1e-10, rep
    assert rep["iterations"] <= 8, rep
    small = sorted(range(64), key=lambda k: probs[k][0].Nx * probs[k][0].Nt)[:6]
    for k in small:
        fm, b, _ = probs[k]
        E, P = fm.export()
        _, relv = dd.mmdag_residual(E, P, b.cpu().numpy(), xs[k].cpu().numpy(), los[k].cpu().numpy(), rpf=b.shape[0])
        assert relv.max() <= 1.05e-10, (k, relv.max())


def test_batch_tss_item_independent_of_mix():
    """An item's solution does not depend on the other items or on which
    worker runs it: four items solved among the whole 16-item mix and alone
    in a batch of four (both on several streams) are identical bit for bit."""
    import bench
    specs = bench.batch_mix_specs(16, 7100)
    keep = [k for k in range(16) if specs[k]["lattice"][1] <= 12][:4]
    full = bench.build_mix_problems(specs)
    xs_full, rep = nsl.batch_solve_tss([(fm, b) for fm, b, _ in full], tol=1e-10, nstreams=6)
    assert rep["converged"], rep
    part = bench.build_mix_problems([specs[k] for k in keep])
    xs_part, rep2 = nsl.batch_solve_tss([(fm, b) for fm, b, _ in part], tol=1e-10, nstreams=3)
    assert rep2["converged"], rep2
    for j, k in enumerate(keep):
        assert torch.equal(xs_full[k], xs_part[j]), k


def test_batch_tss_errors():
    ctx = nsl.Context.default()
    fm, _, b = make("honeycomb", (3, 3), 16, 4.0, 2.0, 1, 2, 100, ctx)
    with pytest.raises(nsl.NslxError):      # the same fermion matrix twice
        nsl.batch_solve_tss([(fm, b.to(DEV)), (fm, b.to(DEV))])
    fd = nsl.HubbardDiag(nsl.Lattice.honeycomb(3, 3), 16, 4.0)
    fd.populate(torch.zeros((16, 18), dtype=torch.complex128, device=DEV))
    with pytest.raises(nsl.NslxError):      # HubbardExp only
        nsl.batch_solve_tss([(fd, b.to(DEV))])
