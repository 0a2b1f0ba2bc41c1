"""GPU parity: the HIP path (libnslx.so via the C ABI) against the oracle on
the same seeded inputs.  Tolerances are written in each test: 1e-13 relative
for FP64 applications of M (the target for M.psi), solutions to the level the
condition number allows, 1e-12 FP64 true residual for the solvers, 1e-8 on
Delta H per HMC trajectory.
"""

import math

import numpy as np
import pytest
import torch

import nsl_amd as nsl
from oracle import fermion as F, hmc as ohmc, lattice as olat, rng, solvers as osol

pytestmark = pytest.mark.gpu

DEV = "cuda"


def relerr(a, b):
    a = a.detach().to("cpu", torch.complex128)
    b = torch.as_tensor(b).detach().to("cpu", torch.complex128)
    return float((a - b).abs().max() / b.abs().max())


def make(kind, size, Nt, beta, nfield=1, seed=1000, std=None, mu=0.0, species=nsl.Particle):
    if kind == "honeycomb":
        lat_d, lat_o = nsl.Lattice.honeycomb(*size), olat.honeycomb(*size)
    elif kind == "ring":
        lat_d, lat_o = nsl.Lattice.ring(size), olat.ring(size)
    else:
        lat_d, lat_o = nsl.Lattice.armchair_nanotube(*size), olat.armchair_nanotube(*size)
    Nx = lat_d.sites()
    std = std if std is not None else 0.7
    phis = [torch.as_tensor(rng.real_field(seed + f, 0, (Nt, Nx), std)) for f in range(nfield)]
    phi = torch.stack(phis)
    fm = nsl.HubbardExp(lat_d, Nt, beta, mu, species)
    fm.populate(phi.to(DEV))
    ofs = []
    for f in range(nfield):
        o = F.HubbardExp(lat_o, Nt, beta, mu, species)
        o.populate(phis[f])
        ofs.append(o)
    return fm, ofs, phi, lat_d, lat_o


def oracle_apply(ofs, name, psi):
    nf = len(ofs)
    nb = psi.shape[0]
    rpf = nb // nf
    out = []
    for k in range(nb):
        out.append(getattr(ofs[k // rpf], name)(psi[k]))
    return torch.stack(out)


OPS = ["M", "Mdagger", "MMdagger", "MdaggerM"]


@pytest.mark.parametrize("kind,size,Nt,nfield,nrhs", [
    ("honeycomb", (3, 3), 16, 1, 3),       # the CPU-runnable configuration (Nx = 18)
    ("ring", 2, 2, 1, 1),
    ("ring", 10, 14, 2, 4),
    ("ring", 16, 12, 1, 2),
    ("honeycomb", (12, 12), 64, 2, 4),     # Nx = 288, Nt = 64 shape, 2 fields x 2 RHS
    ("nanotube", (6, 4), 32, 1, 2),        # Nx = 96
])
@pytest.mark.parametrize("species", [nsl.Particle, nsl.Hole])
def test_apply_f64_parity(kind, size, Nt, nfield, nrhs, species):
    fm, ofs, phi, lat, _ = make(kind, size, Nt, 4.0, nfield, mu=0.05, species=species)
    psi = torch.as_tensor(rng.complex_normal(77, 0, (nrhs, Nt, lat.sites())))
    for name in OPS:
        got = getattr(fm, name)(psi.to(DEV))
        ref = oracle_apply(ofs, name, psi)
        assert relerr(got, ref) < 1e-13, name


def test_apply_edge_shapes():
    # single RHS as a 2-D tensor, Nx not a multiple of any tile, Nt odd
    fm, ofs, phi, lat, _ = make("ring", 7, 5, 2.0)
    psi = torch.as_tensor(rng.complex_normal(5, 0, (5, 7)))
    for name in OPS:
        got = getattr(fm, name)(psi.to(DEV))
        assert got.shape == (5, 7)
        assert relerr(got, getattr(ofs[0], name)(psi)) < 1e-13


def test_apply_errors():
    fm, ofs, phi, lat, _ = make("ring", 6, 4, 2.0, nfield=2)
    psi = torch.zeros((3, 4, 6), dtype=torch.complex128, device=DEV)   # 3 not a multiple of 2 fields
    with pytest.raises(nsl.NslxError):
        fm.M(psi)
    fm2 = nsl.HubbardExp(lat, 4, 2.0)
    with pytest.raises(nsl.NslxError):                                 # apply before populate
        fm2.M(torch.zeros((4, 6), dtype=torch.complex128, device=DEV))


@pytest.mark.parametrize("prec,tol", [(nsl.F32, 2e-6), (nsl.BF16X3, 4e-5)])
@pytest.mark.parametrize("kind,size,Nt", [("honeycomb", (3, 3), 16), ("honeycomb", (12, 12), 64)])
def test_apply_low_precision(prec, tol, kind, size, Nt):
    """FP32 and BF16x3 (hi*hi+hi*lo+lo*hi, FP32 accumulate) applications:
    relative error against the FP64 oracle stated per precision."""
    fm, ofs, phi, lat, _ = make(kind, size, Nt, 4.0)
    psi = torch.as_tensor(rng.complex_normal(78, 0, (2, Nt, lat.sites())))
    for name in OPS:
        got = getattr(fm, name)(psi.to(DEV), precision=prec)
        assert got.dtype == torch.complex64
        e = relerr(got, oracle_apply(ofs, name, psi))
        assert e < tol, (name, e)


def test_cg_reference_semantics_cfg0():
    """CG(fm, MMdagger, eps): x0 = b, batch-global alpha/beta, absolute
    |r|^2 <= eps^2 stop on the recursive residual (CG.tpp), against the
    oracle's CG.  At eps = 1e-8 the two recurrences take the same number of
    steps; at eps = 1e-12 (|r|/|b| ~ 1e-14, below the FP64 floor of the true
    residual) the count depends on rounding, so only the solution is compared."""
    fm, ofs, phi, lat, _ = make("honeycomb", (3, 3), 16, 4.0, std=math.sqrt(0.5))
    b = torch.as_tensor(rng.complex_normal(2000, 0, (8, 16, 18)))
    cg = nsl.CG(fm, nsl.MMdagger, 1e-8, 10000)
    x = cg(b.to(DEV))
    xo, info = osol.cg(ofs[0].MMdagger, b, 1e-8)
    assert cg.converged and info["converged"]
    assert abs(cg.report["iterations"] - info["iterations"]) <= 2
    assert relerr(x, xo) < 1e-8
    cg = nsl.CG(fm, nsl.MMdagger, 1e-12, 10000, check_residual=True)
    x = cg(b.to(DEV))
    xo, info = osol.cg(ofs[0].MMdagger, b, 1e-12)
    assert cg.converged and info["converged"]
    assert relerr(x, xo) < 1e-9
    assert cg.report["rsqr"] <= 1e-24
    assert cg.report["max_rel_residual"] < 1e-10


@pytest.mark.parametrize("op", [nsl.MMdagger, nsl.MdaggerM])
def test_cg_per_rhs_relative(op):
    fm, ofs, phi, lat, _ = make("honeycomb", (3, 3), 16, 4.0, nfield=2)
    b = torch.as_tensor(rng.complex_normal(2001, 0, (4, 16, 18)))
    b[3] *= 1e-6                               # per-RHS scaling must not matter
    cg = nsl.CG(fm, op, 1e-12, 20000, tol_kind=nsl.TOL_REL_PER_RHS_FLOOR, check_residual=True)
    x = cg(b.to(DEV))
    assert cg.converged
    name = "MMdagger" if op == nsl.MMdagger else "MdaggerM"
    r = b - oracle_apply(ofs, name, x.cpu())
    target = max(1e-12, cg.report["max_floor_rel"])     # attainable FP64 accuracy
    for k in range(4):
        assert float(r[k].norm() / b[k].norm()) < 2 * target
    assert cg.report["max_rel_residual"] <= target


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32, nsl.BF16X3])
@pytest.mark.parametrize("kind,size,Nt,beta", [("honeycomb", (3, 3), 16, 4.0), ("honeycomb", (12, 12), 64, 8.0)])
def test_cg_reliable_update_true_residual(prec, kind, size, Nt, beta):
    """CG with FP64 reliable updates: the recurrence in FP64 / FP32 / BF16x3,
    stop on the FP64 TRUE relative residual 1e-12 per RHS (checked here again
    with the oracle's own MMdagger to 2e-12)."""
    fm, ofs, phi, lat, _ = make(kind, size, Nt, beta, nfield=2, std=math.sqrt(beta / Nt * 3.0))
    b = torch.as_tensor(rng.complex_normal(2002, 0, (4, Nt, lat.sites())))
    cg = nsl.CG(fm, nsl.MMdagger, 1e-12, 200000, precision=prec, tol_kind=nsl.TOL_REL_PER_RHS_FLOOR)
    x = cg(b.to(DEV))
    assert cg.converged, cg.report
    target = max(1e-12, cg.report["max_floor_rel"])     # attainable FP64 accuracy
    assert cg.report["max_rel_residual"] <= target
    assert cg.report["outer_steps"] >= 1
    r = b - oracle_apply(ofs, "MMdagger", x.cpu())
    for k in range(4):
        assert float(r[k].norm() / b[k].norm()) < 2 * target


def test_mixed_solution_matches_fp64_cfg0():
    fm, ofs, phi, lat, _ = make("honeycomb", (3, 3), 16, 4.0, nfield=2)
    b = torch.as_tensor(rng.complex_normal(2003, 0, (4, 16, 18)))
    x64 = nsl.CG(fm, nsl.MMdagger, 1e-12, 20000, tol_kind=nsl.TOL_REL_PER_RHS)(b.to(DEV))
    for prec in (nsl.F32, nsl.BF16X3):
        x = nsl.CG(fm, nsl.MMdagger, 1e-12, 20000, precision=prec)(b.to(DEV))
        assert relerr(x, x64) < 1e-9


def test_cg_floor_stop_is_not_converged():
    """TOL_REL_PER_RHS on a system whose 1e-12 lies below the FP64 floor
    (6x6 honeycomb, Nt = 64, beta = 8, U = 3, seed 1000: cond(MM^dag) ~ 6e9):
    the RHS stop at the attainable accuracy, the report counts them, and the
    call returns NOT_CONVERGED; TOL_REL_PER_RHS_FLOOR accepts the same stop,
    and the double-double solver reaches 1e-12."""
    fm, ofs, phi, lat, _ = make("honeycomb", (6, 6), 64, 8.0, nfield=1, std=math.sqrt(8.0 / 64 * 3.0))
    b = torch.as_tensor(rng.complex_normal(2003, 0, (2, 64, 72))).to(DEV)
    cg = nsl.CG(fm, nsl.MMdagger, 1e-12, 100000, tol_kind=nsl.TOL_REL_PER_RHS)
    cg(b)
    assert not cg.converged and cg.report["floor_stops"] >= 1, cg.report
    cg = nsl.CG(fm, nsl.MMdagger, 1e-12, 100000, tol_kind=nsl.TOL_REL_PER_RHS_FLOOR)
    cg(b)
    assert cg.converged and cg.report["floor_stops"] >= 1
    cg = nsl.CG(fm, nsl.MMdagger, 1e-12, 100000, double_double=True)
    x = cg(b)
    assert cg.converged and cg.report["floor_stops"] == 0, cg.report
    assert cg.report["max_rel_residual"] <= 1e-12
    rel = nsl.dd_residual(fm, b, x, cg.x_lo)
    assert rel.max() <= 1e-12


def test_cg_not_converged_returns_iterate():
    fm, ofs, phi, lat, _ = make("honeycomb", (3, 3), 16, 4.0)
    b = torch.as_tensor(rng.complex_normal(3, 0, (2, 16, 18)))
    cg = nsl.CG(fm, nsl.MMdagger, 1e-12, 5)
    x = cg(b.to(DEV))
    assert not cg.converged
    xo, info = osol.cg(ofs[0].MMdagger, b, 1e-12, maxIter=5)
    assert relerr(x, xo) < 1e-12


def test_bicgstab_reference_semantics():
    fm, ofs, phi, lat, _ = make("honeycomb", (3, 3), 16, 4.0)
    b = torch.as_tensor(rng.complex_normal(4, 0, (2, 16, 18)))
    bi = nsl.BiCGStab(fm, nsl.MMdagger, 1e-10, 2000)
    x = bi(b.to(DEV))
    xo, info = osol.bicgstab(ofs[0].MMdagger, b, 1e-10, 2000)
    assert bi.converged and info["converged"]
    assert relerr(x, xo) < 1e-8
    r = b - ofs[0].MMdagger(x.cpu())
    assert float(r.norm() / b.norm()) < 1e-10


def test_cg_on_M_generic_path():
    # CG on the non-Hermitian M is allowed by the reference's Solver; same
    # iterates as the oracle for a few steps
    fm, ofs, phi, lat, _ = make("ring", 6, 4, 1.0)
    b = torch.as_tensor(rng.complex_normal(6, 0, (1, 4, 6)))
    cg = nsl.CG(fm, nsl.M, 1e-12, 4)
    x = cg(b.to(DEV))
    xo, _ = osol.cg(ofs[0].M, b, 1e-12, maxIter=4)
    assert relerr(x, xo) < 1e-10


@pytest.mark.parametrize("species", [nsl.Particle, nsl.Hole])
def test_logdet_and_grad_cfg0(species):
    fm, ofs, phi, lat, lo = make("honeycomb", (3, 3), 16, 4.0, std=math.sqrt(0.5), species=species)
    ld = fm.logDetM()
    ref = ofs[0].logDetM()
    assert abs(ld.real - ref.real) < 1e-10 * abs(ref.real)
    assert abs(math.remainder(ld.imag - ref.imag, 2 * math.pi)) < 1e-9
    for nc in (1, 2, 4):
        assert abs(fm.logDetM(nchunks=nc) - ld) < 1e-10
    g = fm.gradLogDetM()
    assert relerr(g, ofs[0].gradLogDetM()) < 1e-11


def test_logdet_large_beta_closed_form():
    """12x12 honeycomb, beta = 8: log det ~ 1815 (the reference's det
    overflows); phi = 0 closed form sum log(1 + e^{beta eps})."""
    lat = nsl.Lattice.honeycomb(12, 12)
    fm = nsl.HubbardExp(lat, 64, 8.0)
    fm.populate(torch.zeros((64, 288), dtype=torch.complex128, device=DEV))
    eps = np.linalg.eigvalsh(lat.K)
    exact = float(np.sum(np.logaddexp(0.0, 8.0 * eps)))
    ld = fm.logDetM()
    assert abs(ld.real - exact) < 1e-10 * exact
    # uniform-slice closed form with a random phase per slice
    th = rng.uniform(9, 0, 64)
    phi = torch.as_tensor(np.repeat(th[:, None], 288, 1).astype(np.complex128))
    fm.populate(phi.to(DEV))
    ex = complex(np.sum(np.log(1 + np.exp(1j * th.sum()) * np.exp(8.0 * eps))))
    got = fm.logDetM()
    assert abs(got.real - ex.real) < 1e-10 * abs(ex.real)
    assert abs(math.remainder(got.imag - ex.imag, 2 * math.pi)) < 1e-7


def test_action_force_cfg0():
    lat_d, lat_o = nsl.Lattice.honeycomb(3, 3), olat.honeycomb(3, 3)
    A = nsl.HubbardAction(lat_d, 16, 4.0, 2.0)
    Ao = ohmc.HubbardAction(lat_o, 16, 4.0, 2.0)
    phi = torch.as_tensor(rng.real_field(1000, 0, (16, 18), math.sqrt(0.5)))
    S = A.eval(phi.to(DEV))
    So = Ao.eval(phi)
    assert abs(S - So) < 1e-9 * abs(So)
    f = A.force(phi.to(DEV))
    assert relerr(f, Ao.force(phi)) < 1e-11


def test_hmc_delta_h_cfg0_100_trajectories():
    """The CPU-runnable configuration: honeycomb 3x3, Nt=16, beta=4, U=2,
    100 trajectories of length 1 with 10 leapfrog steps; momenta and
    Metropolis uniforms from Philox seed 3000.  Delta H per trajectory must
    agree with the oracle to 1e-8."""
    lat_d, lat_o = nsl.Lattice.honeycomb(3, 3), olat.honeycomb(3, 3)
    U, beta, Nt = 2.0, 4.0, 16
    phi0 = torch.as_tensor(rng.real_field(1000, 0, (Nt, 18), math.sqrt(beta / Nt * U)))
    A = nsl.HubbardAction(lat_d, Nt, beta, U)
    phi = phi0.clone().to(DEV)
    reps = A.chain(phi, 100, 1.0, 10, 3000)
    Ao = ohmc.HubbardAction(lat_o, Nt, beta, U)
    outs = ohmc.hmc_chain(Ao, phi0, 100, 1.0, 10, 3000)
    # Per trajectory, the oracle's own sensitivity: its Delta H again with
    # the momenta perturbed at the rounding level (relative 1e-15, two
    # draws).  Most trajectories move by 1e-11 .. 1e-9 and must agree to
    # 1e-8.  Trajectories whose leapfrog kicks hard (|Delta H| of tens to
    # thousands) amplify rounding by up to ~1e10 (Delta H moves by up to
    # ~3e-5 under the perturbation, while cond(M) at their ends is only
    # 50-500: it is the trajectory, not log det M, that is ill-conditioned);
    # there the GPU must agree within 100x that self-disagreement.  The
    # accept/reject decision must agree everywhere.
    g = np.random.default_rng(17)
    worst = 0.0
    phi_k = phi0
    S_k = Ao.eval(phi_k)
    for k, (r, o) in enumerate(zip(reps, outs)):
        ref = o["dH"].real
        p = ohmc.momentum_draw(3000, k, (Nt, 18))
        sens = 0.0
        for _ in range(2):
            pp = p * (1 + 1e-15 * torch.as_tensor(g.standard_normal(p.shape)))
            sens = max(sens, abs(ohmc.hmc_trajectory(Ao, phi_k, S_k, pp, 0.5, 1.0, 10)["dH"].real - ref))
        tol = max(1e-8, 100.0 * sens)
        err = abs(r["dH_re"] - ref)
        assert err < tol, (k, r["dH_re"], ref, sens)
        worst = max(worst, err / tol)
        assert bool(r["accepted"]) == bool(o["accepted"]), k
        phi_k, S_k = o["phi"], o["S"]
    print(f"Delta H: worst error / per-trajectory tolerance {worst:.3g}")
    acc = sum(int(r["accepted"]) for r in reps)
    assert acc > 30, acc
    assert relerr(phi, outs[-1]["phi"]) < 1e-9
