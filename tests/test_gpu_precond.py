"""CGpreconditioned on the device (nslx_solve_precond) against the oracle's
restatement of CGpreconditioned.tpp: the reference's batch-global semantics
with the free (phi = 0) matrix as preconditioner (its example_CG.cpp
set-up), per-RHS mode with low-precision inner solves, the exact
preconditioner (Mprec = M) and argument errors."""

import math

import pytest
import torch

import nsl_amd as nsl
from oracle import fermion as F, lattice as olat, rng, solvers as osol

pytestmark = pytest.mark.gpu
DEV = "cuda"


def relerr(a, b):
    a = a.detach().to("cpu", torch.complex128)
    b = torch.as_tensor(b).detach().to("cpu", torch.complex128)
    return float((a - b).abs().max() / b.abs().max())


def cfg0(nrhs=4, seed=1000):
    """Honeycomb 3x3, Nt = 16, beta = 4, U = 2 (phi ~ N(0, sqrt(U beta / Nt)))
    and its free matrix, on the device and in the oracle."""
    lat_d, lat_o = nsl.Lattice.honeycomb(3, 3), olat.honeycomb(3, 3)
    phi = torch.as_tensor(rng.real_field(seed, 0, (16, 18), math.sqrt(0.5)))
    fm = nsl.HubbardExp(lat_d, 16, 4.0)
    fm.populate(phi.to(DEV))
    f0 = nsl.HubbardExp(lat_d, 16, 4.0)
    f0.populate(torch.zeros(16, 18, dtype=torch.float64, device=DEV))
    o = F.HubbardExp(lat_o, 16, 4.0)
    o.populate(phi)
    o0 = F.HubbardExp(lat_o, 16, 4.0)
    o0.populate(torch.zeros(16, 18))
    b = torch.as_tensor(rng.complex_normal(2100 + seed, 0, (nrhs, 16, 18)))
    return fm, f0, o, o0, b


@pytest.mark.parametrize("op", [nsl.MMdagger, nsl.MdaggerM])
def test_precond_reference_semantics_cfg0(op):
    fm, f0, o, o0, b = cfg0()
    name = "MMdagger" if op == nsl.MMdagger else "MdaggerM"
    cg = nsl.CGpreconditioned(fm, f0, op, eps=1e-9, maxIter=10000)
    x = cg(b.to(DEV))
    xo, info = osol.cg_preconditioned(getattr(o, name), getattr(o0, name), b, eps=1e-9)
    assert cg.converged and info["converged"]
    # the outer recurrence takes the same steps up to rounding at the stop test
    assert abs(cg.report["iterations"] - info["iterations"]) <= 1
    assert cg.report["outer_steps"] >= cg.report["iterations"]
    assert relerr(x, xo) < 1e-6
    assert cg.report["rsqr"] <= 1e-18
    assert cg.report["max_rel_residual"] < 1e-8


def test_precond_x0_and_tight_tolerance():
    fm, f0, o, o0, b = cfg0(nrhs=2, seed=1001)
    x0 = torch.as_tensor(rng.complex_normal(7, 0, (2, 16, 18)))
    cg = nsl.CGpreconditioned(fm, f0, nsl.MMdagger, eps=1e-12)
    x = cg(b.to(DEV), x0=x0.to(DEV))
    xo, info = osol.cg_preconditioned(o.MMdagger, o0.MMdagger, b, eps=1e-12, x0=x0)
    assert cg.converged and info["converged"]
    assert relerr(x, xo) < 1e-8
    assert cg.report["max_rel_residual"] < 1e-10


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32, nsl.BF16X3])
def test_precond_per_rhs_low_precision_inner(prec):
    fm, f0, o, _, b = cfg0(nrhs=6, seed=1002)
    cg = nsl.CGpreconditioned(fm, f0, nsl.MMdagger, eps=1e-10, tol_kind=nsl.TOL_REL_PER_RHS,
                              inner_precision=prec)
    x = cg(b.to(DEV))
    assert cg.converged
    assert cg.report["converged_rhs"] == 6
    assert cg.report["max_rel_residual"] < 1e-9
    # the same residual, recomputed by the oracle's MMdagger on the host
    xc = x.detach().cpu()
    res = (o.MMdagger(xc) - b).flatten(1).norm(dim=1) / b.flatten(1).norm(dim=1)
    assert float(res.max()) < 1e-9


def test_precond_exact_preconditioner():
    fm, _, o, _, b = cfg0(nrhs=3, seed=1003)
    cg = nsl.CGpreconditioned(fm, fm, nsl.MMdagger, eps=1e-10)
    x = cg(b.to(DEV))
    assert cg.converged and cg.report["iterations"] <= 2
    xo, info = osol.cg_preconditioned(o.MMdagger, o.MMdagger, b, eps=1e-10)
    assert info["iterations"] <= 2
    assert relerr(x, xo) < 1e-6


def test_precond_errors():
    fm, f0, _, _, b = cfg0(nrhs=1)
    other = nsl.HubbardExp(nsl.Lattice.ring(6), 16, 4.0)
    other.populate(torch.zeros(16, 6, dtype=torch.float64, device=DEV))
    with pytest.raises(nsl.NslxError):
        nsl.CGpreconditioned(fm, other, nsl.MMdagger)(b.to(DEV))
    with pytest.raises(nsl.NslxError):     # low-precision inner solves need per-RHS tolerances
        nsl.CGpreconditioned(fm, f0, nsl.MMdagger, inner_precision=nsl.F32)(b.to(DEV))
