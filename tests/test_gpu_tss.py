"""The time-slice direct solve (nslx_fm_solve_direct) and CG on MM^dag
preconditioned by it inside a double-double defect correction
(nslx_solve_tss), through the C ABI.

Parity: the direct solves are checked by their residual against the
oracle's own M / M^dag (hubbardExp.tpp:12-75 restated): |M x - y| / |y| <=
1e-10 in FP64 (the lattices here have cond(M) < 1e4) and <= 1e-3 in FP32.
The preconditioned solver must bring every right-hand side to
|b - MM^dag (x + x_lo)| <= 1e-12 |b|; that residual is recomputed on the CPU
in double-double (tests/ddarith.py, pinned against exact rational
arithmetic) from the FP64 E and phase table the engine exported, so the
check does not trust the device's own residual kernel."""

import ctypes
import math

import numpy as np
import pytest
import torch

import nsl_amd as nsl
from oracle import fermion as F, lattice as olat, rng
from tests import ddarith as dd

pytestmark = pytest.mark.gpu

DEV = "cuda"


def make(kind, size, Nt, beta, nfield=1, seed=1000, std=0.7, mu=0.0, species=nsl.Particle, im_std=0.0):
    if kind == "honeycomb":
        lat_d, lat_o = nsl.Lattice.honeycomb(*size), olat.honeycomb(*size)
    elif kind == "ring":
        lat_d, lat_o = nsl.Lattice.ring(size), olat.ring(size)
    else:
        lat_d, lat_o = nsl.Lattice.armchair_nanotube(*size), olat.armchair_nanotube(*size)
    Nx = lat_d.sites()
    phis = []
    for f in range(nfield):
        p = torch.as_tensor(rng.real_field(seed + f, 0, (Nt, Nx), std))
        if im_std:
            p = p + 1j * torch.as_tensor(rng.real_field(seed + 500 + f, 0, (Nt, Nx), im_std))
        phis.append(p.to(torch.complex128))
    fm = nsl.HubbardExp(lat_d, Nt, beta, mu, species)
    fm.populate(torch.stack(phis).to(DEV))
    ofs = []
    for p in phis:
        o = F.HubbardExp(lat_o, Nt, beta, mu, species)
        o.populate(p)
        ofs.append(o)
    return fm, ofs


def oracle_apply(ofs, name, psi):
    rpf = psi.shape[0] // len(ofs)
    return torch.stack([getattr(ofs[k // rpf], name)(psi[k]) for k in range(psi.shape[0])])


def rel(a, b):
    return float(((a - b).abs() ** 2).sum().sqrt() / (b.abs() ** 2).sum().sqrt())


# (kind, size, Nt, beta, extra): the chunk count comes out as 4, 8, 2, 4, 11
# (Nt prime: one slice per chunk) and, forced, 1 (the whole product)
CASES = [
    ("honeycomb", (3, 3), 16, 4.0, {}),
    ("honeycomb", (6, 6), 32, 6.0, {}),
    ("ring", 10, 12, 2.0, {}),
    ("nanotube", (3, 4), 20, 3.0, {}),
    ("honeycomb", (3, 3), 11, 2.0, {}),
    ("ring", 8, 8, 1.0, {"nchunks": 1}),
    # complex phi, hole species, nonzero mu (the reference's complex-phi tests,
    # test_fermionMatrixHubbardExp.cpp:150-151)
    ("honeycomb", (3, 3), 16, 4.0, {"im_std": 0.3, "species": nsl.Hole, "mu": 0.3}),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32])
def test_direct_solve_residual(case, prec):
    kind, size, Nt, beta, extra = CASES[case]
    nch = extra.get("nchunks", 0)
    kw = {k: v for k, v in extra.items() if k != "nchunks"}
    fm, ofs = make(kind, size, Nt, beta, nfield=2, seed=1100 + case, **kw)
    Nx = fm.Nx
    y = torch.as_tensor(rng.complex_normal(1200 + case, 0, (4, Nt, Nx)))
    tol = 1e-10 if prec == nsl.F64 else 1e-3
    for op, name in ((nsl.MatrixCombination.M, "M"), (nsl.MatrixCombination.Mdagger, "Mdagger")):
        x = fm.solve_direct(op, y.to(DEV), precision=prec, nchunks=nch).to("cpu", torch.complex128)
        r = rel(oracle_apply(ofs, name, x), y)
        assert r < tol, (name, r)
    # (M M^dag)^{-1} = M^{-dag} M^{-1}: residual of MM^dag, FP64 only (FP32's
    # error is cond(M) * eps32 amplified once more)
    if prec == nsl.F64:
        for op, name in ((nsl.MatrixCombination.MMdagger, "MMdagger"), (nsl.MatrixCombination.MdaggerM, "MdaggerM")):
            x = fm.solve_direct(op, y.to(DEV), nchunks=nch).cpu()
            r = rel(oracle_apply(ofs, name, x), y)
            assert r < 1e-8, (name, r)


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32])
def test_fused_gauss_jordan_matches_separate_launches(monkeypatch, prec):
    """The factorisation with each block's trailing update and the next
    block's pivot panel in one launch (the default) against the form with
    every panel and update in launches of their own (NSLX_GJ_FUSE=0, read
    when the plan is built): the same direct solves to rounding (the chunk
    products' condition, up to ~1e8 for FP64's 16-slice chunks, sets the
    tolerance).  Honeycomb 9x9 (Nx = 162), Nt = 32, 2 fields: 324-row block
    systems (panels of 257-324 and <= 256 live rows) and a 162-row last one."""
    Nt, beta, U = 32, 6.0, 3.0
    y = torch.as_tensor(rng.complex_normal(1401, 0, (4, Nt, 162)))
    nch = 2 if prec == nsl.F64 else 0
    xs = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("NSLX_GJ_FUSE", fuse)
        fm, _ = make("honeycomb", (9, 9), Nt, beta, nfield=2, seed=1400, std=math.sqrt(beta / Nt * U))
        xs[fuse] = fm.solve_direct(nsl.MatrixCombination.MMdagger, y.to(DEV), precision=prec, nchunks=nch).cpu()
    tol = 1e-7 if prec == nsl.F64 else 1e-3
    assert rel(xs["1"], xs["0"]) < tol, rel(xs["1"], xs["0"])


def cpu_dd_check(fm, b, x, x_lo, rpf):
    E, P = fm.export()
    _, relv = dd.mmdag_residual(E, P, b.cpu().numpy(), x.cpu().numpy(), x_lo.cpu().numpy(), rpf=rpf)
    return relv


@pytest.mark.parametrize("species", [nsl.Particle, nsl.Hole])
def test_dd_residual_matches_cpu(species):
    """nslx_dd_residual against the CPU double-double residual on the same
    exported operator, at a near-solution (the FP64 direct solve) whose
    residual is dominated by cancellation, with a nonzero low part."""
    Nt, beta, U = 32, 6.0, 3.0
    fm, _ = make("honeycomb", (6, 6), Nt, beta, nfield=2, seed=1300, std=math.sqrt(beta / Nt * U),
                 species=species, im_std=0.1)
    b = torch.as_tensor(rng.complex_normal(1301, 0, (4, Nt, 72)))
    x = fm.solve_direct(nsl.MatrixCombination.MMdagger, b.to(DEV))
    lo = (x * 1e-17).contiguous()
    dev = nsl.dd_residual(fm, b.to(DEV), x, lo)
    relv = cpu_dd_check(fm, b, x, lo, rpf=2)
    assert np.allclose(dev, relv, rtol=1e-9, atol=0.0), (dev, relv)


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32])
def test_tss_pcg_reaches_1e12(prec):
    """Honeycomb 6x6 (Nx = 72), Nt = 64, beta = 8, U = 3: random fields whose
    MM^dag has condition numbers up to ~1e10, where an FP64 residual cannot
    go below ~1e-7; the double-double solution must reach 1e-12."""
    Nt, beta, U = 64, 8.0, 3.0
    fm, ofs = make("honeycomb", (6, 6), Nt, beta, nfield=2, seed=1000, std=math.sqrt(beta / Nt * U))
    b = torch.as_tensor(rng.complex_normal(2000, 0, (6, Nt, 72)))
    cg = nsl.CGTimeSlice(fm, 1e-12, 2000, precision=prec)
    x = cg(b.to(DEV))
    rep = cg.report
    assert cg.converged, rep
    assert rep["max_rel_residual"] <= 1e-12, rep
    relv = cpu_dd_check(fm, b, x, cg.x_lo, rpf=3)
    assert relv.max() <= 1.05e-12, relv
    # the device's double-double residual agrees with the CPU's
    dev = nsl.dd_residual(fm, b.to(DEV), x, cg.x_lo)
    assert np.allclose(dev, relv, rtol=1e-6, atol=1e-20), (dev, relv)
    # x alone (rounded to FP64) is at the FP64 floor, far above 1e-12 for the
    # ill-conditioned field: x_lo carries the digits
    xonly = cpu_dd_check(fm, b, x, torch.zeros_like(x), rpf=3)
    assert xonly.max() > 1e-11


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32])
def test_tss_pcg_mdagm_reaches_1e12(prec):
    """The same fields with M^dag M (the two-body correlator's second
    system, twoBodyCorrelationFunction.tpp:256-283): preconditioned by
    (M^dag M)^{-1} = M^{-1} M^{-dag}, every RHS to 1e-12 in double-double,
    re-checked on the CPU with the M^dag M double-double residual."""
    Nt, beta, U = 64, 8.0, 3.0
    fm, ofs = make("honeycomb", (6, 6), Nt, beta, nfield=2, seed=1000, std=math.sqrt(beta / Nt * U))
    b = torch.as_tensor(rng.complex_normal(2100, 0, (4, Nt, 72)))
    cg = nsl.CGTimeSlice(fm, 1e-12, 2000, precision=prec, matrixCombination=nsl.MatrixCombination.MdaggerM)
    x = cg(b.to(DEV))
    assert cg.converged, cg.report
    assert cg.report["max_rel_residual"] <= 1e-12, cg.report
    E, P = fm.export()
    _, relv = dd.mmdag_residual(E, P, b.numpy(), x.cpu().numpy(), cg.x_lo.cpu().numpy(), rpf=2, op="MdaggerM")
    assert relv.max() <= 1.05e-12, relv
    dev = nsl.dd_residual(fm, b.to(DEV), x, cg.x_lo, op=nsl.MatrixCombination.MdaggerM)
    assert np.allclose(dev, relv, rtol=1e-6, atol=1e-20), (dev, relv)


@pytest.mark.parametrize("prec", [nsl.F32, nsl.F64])
def test_tss_full_size_bench_workload(prec):
    """The bench workload (12x12 honeycomb, Nx = 288, Nt = 64, beta = 8, U = 3,
    16 fields x 8 RHS) with either preconditioner (FP64: 2 chunks, so the
    first block system has 576 rows and its early pivot panels more than 512
    live rows): every RHS to 1e-12, and one RHS of the first field checked on
    the CPU in double-double."""
    L, Nt, beta, U = 12, 64, 8.0, 3.0
    Nx = 2 * L * L
    std = math.sqrt(beta / Nt * U)
    phi = torch.stack([torch.as_tensor(rng.real_field(1000 + f, 0, (Nt, Nx), std)) for f in range(16)])
    fm = nsl.HubbardExp(nsl.Lattice.honeycomb(L, L), Nt, beta)
    fm.populate(phi.to(torch.complex128).to(DEV))
    b = torch.stack([torch.as_tensor(rng.complex_normal(2000 + 16 * f + k, 0, (Nt, Nx)))
                     for f in range(16) for k in range(8)])
    cg = nsl.CGTimeSlice(fm, 1e-12, 2000, precision=prec)
    x = cg(b.to(DEV))
    assert cg.converged, cg.report
    assert cg.report["max_rel_residual"] <= 1e-12
    E, P = fm.export()
    _, relv = dd.mmdag_residual(E, P[:1], b[:1].numpy(), x[:1].cpu().numpy(), cg.x_lo[:1].cpu().numpy(), rpf=1)
    assert relv.max() <= 1.05e-12, relv


def test_tss_errors():
    lat = nsl.Lattice.honeycomb(3, 3)
    fd = nsl.HubbardDiag(lat, 16, 4.0)
    fd.populate(torch.zeros((16, 18), dtype=torch.complex128, device=DEV))
    y = torch.zeros((1, 16, 18), dtype=torch.complex128, device=DEV)
    with pytest.raises(nsl.NslxError):
        fd.solve_direct(nsl.MatrixCombination.M, y)
    fm = nsl.HubbardExp(lat, 16, 4.0)
    fm.populate(torch.zeros((16, 18), dtype=torch.complex128, device=DEV))
    with pytest.raises(nsl.NslxError):                  # CG on M is not a product system
        nsl.CGTimeSlice(fm, 1e-12, 100, matrixCombination=nsl.MatrixCombination.M)(y)
    with pytest.raises(nsl.NslxError):                  # nchunks must divide Nt
        fm.solve_direct(nsl.MatrixCombination.M, y, nchunks=3)


def _prof_classes(fn):
    """Runs fn with every launch timed; returns the kernel classes launched."""
    L_ = nsl._lib.lib()
    torch.cuda.synchronize()
    L_.nslx_prof_reset()
    L_.nslx_prof_enable(1)
    try:
        out = fn()
        torch.cuda.synchronize()
    finally:
        L_.nslx_prof_enable(0)
    buf = ctypes.create_string_buffer(1 << 16)
    L_.nslx_prof_read(buf, len(buf))
    L_.nslx_prof_reset()
    return out, {line.split()[0] for line in buf.value.decode().splitlines() if line.strip()}


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32])
@pytest.mark.parametrize("case", [1, 3, 6])
def test_chain_launch_matches_per_slice_launches(case, prec, monkeypatch):
    """The persistent recurrence launch (one launch per pass) against one
    batched GEMM launch per slice (NSLX_CHAINS=0) on the same fields: the same
    M^{-1} y and M^{-dag} y up to the summation order (cond(M) < 1e4 here:
    1e-10 in FP64, 1e-3 in FP32), and the kernel classes show which path ran."""
    kind, size, Nt, beta, extra = CASES[case]
    kw = {k: v for k, v in extra.items() if k != "nchunks"}
    res = {}
    y = None
    ops = (nsl.MatrixCombination.M, nsl.MatrixCombination.Mdagger)
    for chains in ("1", "0"):
        monkeypatch.setenv("NSLX_CHAINS", chains)
        fm, _ = make(kind, size, Nt, beta, nfield=2, seed=1100 + case, **kw)
        if y is None:
            y = torch.as_tensor(rng.complex_normal(1400 + case, 0, (4, Nt, fm.Nx))).to(DEV)
        xs, classes = _prof_classes(lambda: [fm.solve_direct(op, y, precision=prec) for op in ops])
        chain_classes = {c for c in classes if c.startswith("chain_")}
        if chains == "1":
            assert len(chain_classes) == 2, classes     # ph1 (M^-1) and ph2 (M^-dag) passes
        else:
            assert not chain_classes, classes
        res[chains] = [x.to("cpu", torch.complex128) for x in xs]
    tol = 1e-10 if prec == nsl.F64 else 1e-3
    for a, b in zip(res["1"], res["0"]):
        assert rel(a, b) < tol

def test_chain_give_up_is_reported_and_recovers(monkeypatch):
    """A wait bound of 0 makes every wait of the persistent recurrence launch
    give up: the launch still ends normally, the solve reports
    NSLX_E_INTERNAL, and the next solve (normal bound) is correct again."""
    Nt, beta = 32, 6.0
    fm, ofs = make("honeycomb", (6, 6), Nt, beta, nfield=2, seed=1500)
    y = torch.as_tensor(rng.complex_normal(1501, 0, (4, Nt, 72)))
    monkeypatch.setenv("NSLX_CHAIN_SPIN", "0")
    with pytest.raises(nsl.NslxError):
        fm.solve_direct(nsl.MatrixCombination.M, y.to(DEV))
    monkeypatch.delenv("NSLX_CHAIN_SPIN")
    x = fm.solve_direct(nsl.MatrixCombination.M, y.to(DEV)).cpu()
    assert rel(oracle_apply(ofs, "M", x), y) < 1e-10


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32])
@pytest.mark.parametrize("kind,size,Nt,beta,nf", [("honeycomb", (3, 3), 16, 4.0, 2),      # 1 column block per wave
                                                  ("honeycomb", (6, 6), 32, 6.0, 2),      # 2
                                                  ("nanotube", (3, 4), 20, 3.0, 1),       # Nx = 48, odd block count
                                                  ("honeycomb", (12, 12), 64, 8.0, 1)])   # Nx = 288, 5
def test_chunk_products_row_kernel_matches_batched_gemm(kind, size, Nt, beta, nf, prec, monkeypatch):
    """The chunk products built by row blocks that keep their rows for the
    whole chunk (one launch) against one batched GEMM launch per slice
    (NSLX_CHUNKPROD=0): the same M^{-1} y and M^{-dag} y up to summation order
    (1e-10 FP64, 1e-3 FP32 relative), and the kernel classes show which path
    built them.  The Nx = 288 case uses the bench's beta, Nt and field
    width (FP64 only: cond(M) ~ 1e5 there, beyond what an FP32 comparison
    at 1e-3 resolves)."""
    if prec == nsl.F32 and size == (12, 12):
        pytest.skip("FP32 solves of the Nx = 288, beta = 8 fields differ by eps32 * cond(M) ~ 1e-2")
    res = {}
    y = None
    ops = (nsl.MatrixCombination.M, nsl.MatrixCombination.Mdagger)
    std = math.sqrt(beta / Nt * 3.0)
    for mode in ("1", "0"):
        monkeypatch.setenv("NSLX_CHUNKPROD", mode)
        fm, ofs = make(kind, size, Nt, beta, nfield=nf, seed=1600, std=std)
        if y is None:
            y = torch.as_tensor(rng.complex_normal(1601, 0, (2 * nf, Nt, fm.Nx)))
        yd = y.to(DEV) if prec == nsl.F64 else y.to(torch.complex64).to(DEV)
        xs, classes = _prof_classes(lambda: [fm.solve_direct(op, yd, precision=prec) for op in ops])
        name = "chunkprod_f64" if prec == nsl.F64 else "chunkprod_f32"
        assert (name in classes) == (mode == "1"), classes
        res[mode] = [x.to("cpu", torch.complex128) for x in xs]
        if mode == "1" and prec == nsl.F64 and fm.Nx <= 72:
            for x, nm in zip(res[mode], ("M", "Mdagger")):
                assert rel(oracle_apply(ofs, nm, x), y) < 1e-9, nm
    tol = 1e-10 if prec == nsl.F64 else 1e-3
    for a, b in zip(res["1"], res["0"]):
        assert rel(a, b) < tol, rel(a, b)


def test_force_chains_match_per_slice_launches(monkeypatch):
    """gradLogDetM with its two recurrences as persistent launches against one
    launch per slice (NSLX_CHAINS=0), and with every wait of the persistent
    launch giving up (NSLX_CHAIN_SPIN=0: the recurrence is run again one
    launch per slice): the same force up to the summation order, and the
    kernel classes show which path ran."""
    fm, ofs = make("honeycomb", (6, 6), 16, 4.0, nfield=1, seed=1600)
    res = {}
    for mode in ("chains", "per_slice", "give_up"):
        monkeypatch.setenv("NSLX_CHAINS", "0" if mode == "per_slice" else "1")
        if mode == "give_up":
            monkeypatch.setenv("NSLX_CHAIN_SPIN", "0")
        g, classes = _prof_classes(lambda: fm.gradLogDetM())
        chain_classes = {c for c in classes if c.startswith("chain_")}
        assert bool(chain_classes) == (mode != "per_slice"), (mode, classes)
        res[mode] = g.to("cpu")
        monkeypatch.delenv("NSLX_CHAIN_SPIN", raising=False)
    ref = ofs[0].gradLogDetM()
    # the recurrences grow by up to exp(delta ||K|| Nt) = exp(12): rounding of
    # a different summation order reaches ~1e-11 of the force
    assert rel(res["chains"], ref) < 1e-10
    assert rel(res["chains"], res["per_slice"]) < 1e-10
    assert rel(res["give_up"], res["per_slice"]) < 1e-10
