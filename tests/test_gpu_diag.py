"""HubbardDiag on the device (nslx_fm_create_diag) against the oracle's
restatement of hubbardDiag.tpp: M, M^dagger, MM^dagger, M^dagger M in every
precision and every classic tile, the fused CG (reference semantics and
reliable updates), preconditioned CG, log det and force (both species), the
two-point correlator, HMC Delta H, and the batching engine.
Tolerances as for HubbardExp: 1e-13 relative (FP64), 2e-6 (FP32), 4e-5
(BF16x3) for applications."""

import math

import pytest
import torch

import nsl_amd as nsl
from oracle import correlator as ocorr, fermion as F, hmc as ohmc, lattice as olat, rng, solvers as osol

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = {nsl.F64: 1e-13, nsl.F32: 2e-6, nsl.BF16X3: 4e-5}
NAMES = {nsl.M: "M", nsl.Mdagger: "Mdagger", nsl.MMdagger: "MMdagger", nsl.MdaggerM: "MdaggerM"}


def relerr(a, b):
    a = a.detach().to("cpu", torch.complex128)
    b = torch.as_tensor(b).detach().to("cpu", torch.complex128)
    return float((a - b).abs().max() / b.abs().max())


def make(kind, size, Nt, beta, nfield=1, mu=0.0, species=nsl.Particle, std=0.7, seed=5000, ctx=None):
    if kind == "honeycomb":
        lat_d, lat_o = nsl.Lattice.honeycomb(*size), olat.honeycomb(*size)
    elif kind == "ring":
        lat_d, lat_o = nsl.Lattice.ring(size), olat.ring(size)
    else:
        lat_d, lat_o = nsl.Lattice.armchair_nanotube(*size), olat.armchair_nanotube(*size)
    Nx = lat_d.sites()
    phis = [torch.as_tensor(rng.real_field(seed + f, 0, (Nt, Nx), std)) for f in range(nfield)]
    fm = nsl.HubbardDiag(lat_d, Nt, beta, mu, species, ctx=ctx)
    fm.populate(torch.stack(phis).to(DEV) if nfield > 1 else phis[0].to(DEV))
    ofs = []
    for p in phis:
        o = F.HubbardDiag(lat_o, Nt, beta, mu, species)
        o.populate(p)
        ofs.append(o)
    return fm, ofs, lat_d, lat_o, phis


def oracle_apply(ofs, name, psi):
    rpf = psi.shape[0] // len(ofs)
    return torch.stack([getattr(ofs[k // rpf], name)(psi[k]) for k in range(psi.shape[0])])


@pytest.mark.parametrize("kind,size,Nt,nfield,nrhs,species", [
    ("honeycomb", (3, 3), 16, 1, 4, nsl.Particle),
    ("honeycomb", (3, 3), 16, 2, 4, nsl.Hole),
    ("ring", 10, 12, 1, 3, nsl.Particle),
    ("nanotube", (6, 2), 20, 1, 2, nsl.Particle),
    ("honeycomb", (6, 6), 32, 1, 8, nsl.Particle),
])
def test_diag_apply_f64(kind, size, Nt, nfield, nrhs, species):
    fm, ofs, _, _, _ = make(kind, size, Nt, 3.0, nfield, mu=0.05, species=species)
    Nx = fm.Nx
    psi = torch.as_tensor(rng.complex_normal(77, 0, (nrhs, Nt, Nx)))
    for op, name in NAMES.items():
        y = fm.apply(op, psi.to(DEV))
        assert relerr(y, oracle_apply(ofs, name, psi)) < 1e-13, name


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32, nsl.BF16X3])
def test_diag_every_classic_tile(prec):
    ctx = nsl.Context()
    Nt, nb = 32, 6
    fm, ofs, _, _, _ = make("honeycomb", (6, 6), Nt, 4.0, 2, mu=0.03, ctx=ctx)
    Nx = fm.Nx
    psi = torch.as_tensor(rng.complex_normal(78, 0, (nb, Nt, Nx)))
    dt = torch.complex128 if prec == nsl.F64 else torch.complex64
    refs = {name: oracle_apply(ofs, name, psi) for name in ("M", "Mdagger")}
    for tile in range(4):
        ctx.set_tile(prec, Nx, Nt, nb, tile)
        for op, name in ((nsl.M, "M"), (nsl.Mdagger, "Mdagger")):
            y = fm.apply(op, psi.to(DEV, dt), precision=prec)
            assert relerr(y, refs[name]) < TOL[prec], (tile, name)
    # a slice tile from the table is replaced by a classic one for HubbardDiag
    n = nsl.Context.tile_count(prec)
    for t in range(4, n):
        try:
            ctx.set_tile(prec, Nx, Nt, nb, t)
        except nsl.NslxError:
            continue
        y = fm.apply(nsl.M, psi.to(DEV, dt), precision=prec)
        assert relerr(y, refs["M"]) < TOL[prec], t
    ctx.set_tile(prec, Nx, Nt, nb, None)


def test_diag_cg_reference_semantics_and_per_rhs():
    fm, ofs, _, _, _ = make("honeycomb", (3, 3), 16, 4.0, std=math.sqrt(0.5))
    b = torch.as_tensor(rng.complex_normal(2000, 0, (6, 16, 18)))
    cg = nsl.CG(fm, nsl.MMdagger, 1e-8, 10000)
    x = cg(b.to(DEV))
    xo, info = osol.cg(ofs[0].MMdagger, b, 1e-8)
    assert cg.converged and info["converged"]
    assert abs(cg.report["iterations"] - info["iterations"]) <= 2
    assert relerr(x, xo) < 1e-7
    for prec in (nsl.F64, nsl.F32, nsl.BF16X3):
        cg = nsl.CG(fm, nsl.MdaggerM, 1e-11, 20000, precision=prec, tol_kind=nsl.TOL_REL_PER_RHS_FLOOR,
                    check_residual=True)
        x = cg(b.to(DEV))
        assert cg.converged, prec
        res = (ofs[0].MdaggerM(x.cpu()) - b).flatten(1).norm(dim=1) / b.flatten(1).norm(dim=1)
        assert float(res.max()) < 1e-10, prec


def test_diag_bicgstab_and_precond():
    fm, ofs, lat_d, lat_o, _ = make("ring", 8, 12, 2.0, std=0.6, seed=5100)
    # BiCGStab on M itself: M = 1 - C with C block-cyclic has its eigenvalues
    # on circles through the origin, so even at cond(M) ~ 18 the reference's
    # recurrence needs a few hundred steps, and with several RHS in its
    # batch-global products it breaks down (the same for HubbardExp; the
    # oracle does too).  One RHS here.
    b1 = torch.as_tensor(rng.complex_normal(2001, 0, (1, 12, 8)))
    bi = nsl.BiCGStab(fm, nsl.M, 1e-9, 5000)
    x = bi(b1.to(DEV))
    xo, info = osol.bicgstab(ofs[0].M, b1, eps=1e-9, maxIter=5000)
    assert bi.converged and info["converged"]
    res = (ofs[0].M(x.cpu()) - b1).norm() / b1.norm()
    assert float(res) < 1e-9
    assert relerr(x, xo) < 1e-6
    b = torch.as_tensor(rng.complex_normal(2001, 0, (3, 12, 8)))
    f0 = nsl.HubbardDiag(lat_d, 12, 2.0)
    f0.populate(torch.zeros(12, 8, dtype=torch.float64, device=DEV))
    o0 = F.HubbardDiag(lat_o, 12, 2.0)
    o0.populate(torch.zeros(12, 8))
    pc = nsl.CGpreconditioned(fm, f0, nsl.MMdagger, eps=1e-9)
    x = pc(b.to(DEV))
    xo, info = osol.cg_preconditioned(ofs[0].MMdagger, o0.MMdagger, b, eps=1e-9)
    assert pc.converged and info["converged"]
    # each inner solve stops at a rounding-dependent step, so the two
    # preconditioners differ slightly and the outer counts by a few
    assert abs(pc.report["iterations"] - info["iterations"]) <= 3
    assert relerr(x, xo) < 1e-6


@pytest.mark.parametrize("species", [nsl.Particle, nsl.Hole])
def test_diag_logdet_and_grad(species):
    fm, ofs, _, _, _ = make("honeycomb", (3, 3), 16, 4.0, species=species, std=math.sqrt(0.5))
    ld = fm.logDetM()
    ref = ofs[0].logDetM()
    assert abs(ld.real - ref.real) < 1e-10 * max(1.0, abs(ref.real))
    assert abs(math.remainder(ld.imag - ref.imag, 2 * math.pi)) < 1e-9
    g = fm.gradLogDetM()
    assert relerr(g, ofs[0].gradLogDetM()) < 1e-10


def test_diag_logdet_large_beta_stable():
    """Honeycomb 6x6 at beta = 12, Nt = 96: the reference's single product
    overflows; the chunked form against the oracle's logdet_stable."""
    fm, ofs, _, _, _ = make("honeycomb", (6, 6), 96, 12.0, std=0.3, seed=5200)
    ld = fm.logDetM()
    ref = ofs[0].logdet_stable(12)
    assert abs(ld.real - ref.real) < 1e-9 * abs(ref.real)
    assert abs(math.remainder(ld.imag - ref.imag, 2 * math.pi)) < 1e-7


def test_diag_correlator_vs_oracle():
    fm, ofs, _, _, _ = make("honeycomb", (3, 3), 16, 4.0, std=math.sqrt(0.5), seed=5300)
    tpc = nsl.TwoPointCorrelator(fm, eps=1e-12)
    corr = tpc.measure(2)
    ref, _ = ocorr.two_point(ofs[0], 16, 18, 2)
    assert relerr(corr, ref) < 1e-8


def test_diag_hmc_delta_h():
    """Honeycomb 3x3, Nt = 16, beta = 4, U = 2 with the diagonal
    discretisation: 20 trajectories (length 1, 10 steps), Philox momenta."""
    lat_d, lat_o = nsl.Lattice.honeycomb(3, 3), olat.honeycomb(3, 3)
    U, beta, Nt = 2.0, 4.0, 16
    phi0 = torch.as_tensor(rng.real_field(1000, 0, (Nt, 18), math.sqrt(beta / Nt * U)))
    A = nsl.HubbardAction(lat_d, Nt, beta, U, fermion_matrix=nsl.HubbardDiag)
    phi = phi0.clone().to(DEV)
    reps = A.chain(phi, 20, 1.0, 10, 3000)
    Ao = ohmc.HubbardAction(lat_o, Nt, beta, U, fermion_matrix=F.HubbardDiag)
    outs = ohmc.hmc_chain(Ao, phi0, 20, 1.0, 10, 3000)
    for k, (r, o) in enumerate(zip(reps, outs)):
        ref = o["dH"].real
        tol = 1e-8 if abs(ref) < 20.0 else 1e-7 * abs(ref)
        assert abs(r["dH_re"] - ref) < tol, (k, r["dH_re"], ref)
        assert bool(r["accepted"]) == bool(o["accepted"]), k
    assert relerr(phi, outs[-1]["phi"]) < 1e-9


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32, nsl.BF16X3])
def test_diag_batching_engine(prec):
    """HubbardDiag matrices in the batching engine, mixed with a HubbardExp
    one (separate grouped classes): every solution's true residual with the
    oracle's own MMdagger, and a HubbardDiag item solved alone gives the same
    bits as in the mix."""
    ctx = nsl.Context(0)
    d1 = make("honeycomb", (3, 3), 16, 4.0, nfield=2, ctx=ctx, seed=5100)
    d2 = make("nanotube", (6, 2), 20, 3.0, ctx=ctx, seed=5200)
    lat_e = nsl.Lattice.honeycomb(3, 3)
    phi_e = torch.as_tensor(rng.real_field(5300, 0, (12, 18), 0.5))
    fe = nsl.HubbardExp(lat_e, 12, 2.0, ctx=ctx)
    fe.populate(phi_e.to(DEV))
    oe = F.HubbardExp(olat.honeycomb(3, 3), 12, 2.0)
    oe.populate(phi_e)
    b1 = torch.as_tensor(rng.complex_normal(5101, 0, (4, 16, 18)))
    b2 = torch.as_tensor(rng.complex_normal(5201, 0, (3, 20, d2[2].sites())))
    be = torch.as_tensor(rng.complex_normal(5301, 0, (2, 12, 18)))
    items = [(d1[0], b1.to(DEV)), (fe, be.to(DEV)), (d2[0], b2.to(DEV))]
    xs, rep = nsl.batch_solve(items, tol=1e-10, precision=prec, ctx=ctx)
    assert rep["converged"], rep
    target = max(1e-10, rep["max_floor_rel"])
    for (ofs, b), x in zip([(d1[1], b1), ([oe], be), (d2[1], b2)], xs):
        r = b - oracle_apply(ofs, "MMdagger", x.cpu())
        rel = r.flatten(1).norm(dim=1) / b.flatten(1).norm(dim=1)
        assert float(rel.max()) < 2 * target
    alone, _ = nsl.batch_solve([items[2]], tol=1e-10, precision=prec, ctx=ctx)
    assert torch.equal(alone[0], xs[2])
