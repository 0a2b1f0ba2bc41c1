"""The benchmark workload at full size: honeycomb 12x12 (Nx = 288), Nt = 64,
beta = 8, U = 3, 16 fields x 8 right-hand sides (128 RHS), on the same seeded
inputs as bench.py.  FP64 applications against the oracle on every RHS
(1e-13), the adjoint identity <a, M b> = <M^dag a, b> per RHS (1e-12, no
oracle involved), and CG on MMdagger to 1e-12 per RHS in FP64 and with the
BF16x3 recurrence, each solution's true residual recomputed with the oracle's
own operator."""

import math

import pytest
import torch

import nsl_amd as nsl
from tests import ddarith as dd
from oracle import fermion as F, lattice as olat, rng

pytestmark = pytest.mark.gpu
DEV = "cuda"
L, NT, BETA, U = 12, 64, 8.0, 3.0
NF, NR = 16, 8
OPS = ["M", "Mdagger", "MMdagger", "MdaggerM"]


@pytest.fixture(scope="module")
def workload():
    lat_d, lat_o = nsl.Lattice.honeycomb(L, L), olat.honeycomb(L, L)
    Nx = lat_d.sites()
    std = math.sqrt(BETA / NT * U)
    phis = [torch.as_tensor(rng.real_field(1000 + f, 0, (NT, Nx), std)) for f in range(NF)]
    b = torch.stack([torch.as_tensor(rng.complex_normal(2000 + 16 * f + k, 0, (NT, Nx)))
                     for f in range(NF) for k in range(NR)])
    fm = nsl.HubbardExp(lat_d, NT, BETA)
    fm.populate(torch.stack(phis).to(DEV))
    ofs = []
    for f in range(NF):
        o = F.HubbardExp(lat_o, NT, BETA)
        o.populate(phis[f])
        ofs.append(o)
    return fm, ofs, b


def oracle_apply(ofs, name, psi):
    return torch.stack([getattr(ofs[k // NR], name)(psi[k]) for k in range(psi.shape[0])])


def test_inputs_are_the_bench_inputs(workload):
    """bench.py draws the same Philox streams through the library (its
    Box-Muller differs from numpy's in the last bits only)."""
    import bench
    _, _, b = workload
    phi_b, b_b = bench.fields()
    assert phi_b.shape == (NF, NT, 2 * L * L) and b_b.shape == tuple(b.shape)
    assert float((torch.as_tensor(b_b) - b).abs().max()) < 1e-14


def test_fullsize_apply_parity(workload):
    fm, ofs, b = workload
    for name in OPS:
        got = getattr(fm, name)(b.to(DEV)).cpu()
        ref = oracle_apply(ofs, name, b)
        err = ((got - ref).abs().amax(dim=(1, 2)) / ref.abs().amax(dim=(1, 2))).max()
        assert float(err) < 1e-13, name


def test_fullsize_adjoint_identity(workload):
    fm, _, b = workload
    a = torch.as_tensor(rng.complex_normal(3000, 0, tuple(b.shape))).to(DEV)
    bd = b.to(DEV)
    lhs = (a.conj() * fm.M(bd)).sum(dim=(1, 2))
    rhs = (fm.Mdagger(a).conj() * bd).sum(dim=(1, 2))
    scale = a.abs().pow(2).sum(dim=(1, 2)).sqrt() * bd.abs().pow(2).sum(dim=(1, 2)).sqrt()
    assert float(((lhs - rhs).abs() / scale).max()) < 1e-12
    # MMdagger is M after Mdagger
    mm = fm.MMdagger(bd)
    comp = fm.M(fm.Mdagger(bd))
    assert float((mm - comp).abs().max() / comp.abs().max()) < 1e-13


@pytest.mark.parametrize("prec", [nsl.F64, nsl.BF16X3])
def test_fullsize_cg_true_residual(workload, prec):
    """The plain (unpreconditioned) CG on the bench workload with the
    solution in double-double: every RHS to 1e-12 true residual, where x
    rounded to FP64 alone stays at the FP64 floor (7e-11 .. 1e-7); one RHS
    rechecked on the CPU in double-double from the exported operator."""
    fm, ofs, b = workload
    cg = nsl.CG(fm, nsl.MMdagger, 1e-12, 200000, precision=prec, double_double=True)
    x = cg(b.to(DEV))
    assert cg.converged, cg.report
    assert cg.report["converged_rhs"] == NF * NR
    assert cg.report["max_rel_residual"] <= 1e-12
    E, P = fm.export()
    _, rel = dd.mmdag_residual(E, P[:1], b[:1].numpy(), x[:1].cpu().numpy(), cg.x_lo[:1].cpu().numpy(), rpf=1)
    assert rel.max() <= 1.05e-12, rel


def test_fullsize_reference_criterion(workload):
    """The reference's own CG (x0 = b, batch-global alpha/beta, stop when
    sum |r|^2 <= eps^2 on the recursive residual; CG.tpp) on the first field's
    8 RHS of the bench workload.  (a) 40 iterations against the oracle's 40:
    the same recurrence, so the iterates agree up to rounding (a complete
    oracle solve takes ~20k iterations, about a minute on the host).  (b) The
    complete solve at eps = 1e-12 stops on the recursive residual, and its true
    residual, recomputed with the oracle's operator, is the FP64 floor of these
    fields (cond(MM^dag) ~ 1e9): small, but above the recursive one."""
    from oracle import solvers as osol
    _, ofs, b = workload
    Nx = 2 * L * L
    phi0 = torch.as_tensor(rng.real_field(1000, 0, (NT, Nx), math.sqrt(BETA / NT * U)))
    fm = nsl.HubbardExp(nsl.Lattice.honeycomb(L, L), NT, BETA)
    fm.populate(phi0[None].to(DEV))
    b0 = b[:NR]
    cg = nsl.CG(fm, nsl.MMdagger, 1e-12, 40)
    x40 = cg(b0.to(DEV)).cpu()
    xo40, info = osol.cg(ofs[0].MMdagger, b0, 1e-12, 40)
    assert not cg.converged and not info["converged"]
    err40 = float(((x40 - xo40).abs() ** 2).sum().sqrt() / (xo40.abs() ** 2).sum().sqrt())
    print(f"reference CG, 40 iterations: |x_gpu - x_oracle| / |x_oracle| = {err40:.3e}")
    assert err40 < 1e-9
    cg = nsl.CG(fm, nsl.MMdagger, 1e-12, 100000, check_residual=True)
    x = cg(b0.to(DEV)).cpu()
    assert cg.converged, cg.report
    assert cg.report["rsqr"] <= 1e-24
    r = b0 - oracle_apply(ofs[:1], "MMdagger", x)
    rel = [float(r[k].norm() / b0[k].norm()) for k in range(NR)]
    print(f"reference CG complete: {cg.report['iterations']} iterations, true relative residuals "
          f"{min(rel):.2e} .. {max(rel):.2e}")
    # at the floor the true residual is itself mostly rounding: the device's
    # and the oracle's values agree in magnitude, not in digits
    assert max(rel) < 1e-5 and cg.report["max_rel_residual"] < 1e-5
