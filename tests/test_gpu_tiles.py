"""Every hop tile variant (the classic LDS-staged tiles and the E-slice-
resident tiles) against the oracle, in every precision and in the fused CG
modes, with the tile pinned through the context's tile table.  Tolerances:
1e-13 relative for FP64, 2e-6 for FP32, 4e-5 for BF16x3 applications; the
solvers to their FP64 true-residual target."""

import math

import pytest
import torch

import nsl_amd as nsl
from oracle import fermion as F, lattice as olat, rng

pytestmark = pytest.mark.gpu

DEV = "cuda"
TOL = {nsl.F64: 1e-13, nsl.F32: 2e-6, nsl.BF16X3: 4e-5}


def relerr(a, b):
    a = a.detach().to("cpu", torch.complex128)
    b = torch.as_tensor(b).detach().to("cpu", torch.complex128)
    return float((a - b).abs().max() / b.abs().max())


def setup(kind, size, Nt, beta, nfield, ctx, std=0.7):
    if kind == "honeycomb":
        lat_d, lat_o = nsl.Lattice.honeycomb(*size), olat.honeycomb(*size)
    elif kind == "ring":
        lat_d, lat_o = nsl.Lattice.ring(size), olat.ring(size)
    else:
        lat_d, lat_o = nsl.Lattice.armchair_nanotube(*size), olat.armchair_nanotube(*size)
    Nx = lat_d.sites()
    phis = [torch.as_tensor(rng.real_field(4000 + f, 0, (Nt, Nx), std)) for f in range(nfield)]
    fm = nsl.HubbardExp(lat_d, Nt, beta, 0.03, ctx=ctx)
    fm.populate(torch.stack(phis).to(DEV))
    ofs = []
    for p in phis:
        o = F.HubbardExp(lat_o, Nt, beta, 0.03)
        o.populate(p)
        ofs.append(o)
    return fm, ofs, Nx


def oracle_apply(ofs, name, psi):
    rpf = psi.shape[0] // len(ofs)
    return torch.stack([getattr(ofs[k // rpf], name)(psi[k]) for k in range(psi.shape[0])])


def fitting_tiles(ctx, prec, Nx, Nt, nb):
    out = []
    for t in range(nsl.Context.tile_count(prec)):
        try:
            ctx.set_tile(prec, Nx, Nt, nb, t)
        except nsl.NslxError:
            continue
        out.append(t)
    ctx.set_tile(prec, Nx, Nt, nb, None)
    return out


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32, nsl.BF16X3])
@pytest.mark.parametrize("kind,size,Nt,nfield,nrhs", [
    ("ring", 10, 14, 2, 2),                # Nx, Nt below one 16-row block
    ("honeycomb", (6, 6), 20, 1, 3),       # Nx = 72, Nt not a multiple of 16
    ("nanotube", (5, 3), 33, 1, 2),        # Nx = 60 (not a multiple of 16 or 32), odd Nt
    ("honeycomb", (12, 12), 64, 2, 2),     # Nx = 288, Nt = 64
])
def test_every_tile_apply(prec, kind, size, Nt, nfield, nrhs):
    ctx = nsl.Context(0)
    fm, ofs, Nx = setup(kind, size, Nt, 4.0, nfield, ctx)
    nb = nfield * nrhs
    psi = torch.as_tensor(rng.complex_normal(91, 0, (nb, Nt, Nx)))
    ref = {name: oracle_apply(ofs, name, psi) for name in ("M", "Mdagger")}
    tiles = fitting_tiles(ctx, prec, Nx, Nt, nb)
    assert len(tiles) >= 4
    for t in tiles:
        ctx.set_tile(prec, Nx, Nt, nb, t)
        assert ctx.get_tile(prec, Nx, Nt, nb) == t
        for name in ("M", "Mdagger"):
            got = getattr(fm, name)(psi.to(DEV), precision=prec)
            e = relerr(got, ref[name])
            assert e < TOL[prec], (t, name, e)


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32, nsl.BF16X3])
def test_every_tile_cg(prec):
    """Fused CG kernels (CGI/CGA/CGB) of every tile: reliable-update CG to the
    FP64 true residual, per RHS, and for F64 also the reference's batch-global
    recurrence (one alpha/beta for the batch)."""
    ctx = nsl.Context(0)
    Nt, beta = 16, 4.0
    fm, ofs, Nx = setup("honeycomb", (6, 6), Nt, beta, 2, ctx, std=math.sqrt(beta / Nt * 2.0))
    nb = 4
    b = torch.as_tensor(rng.complex_normal(92, 0, (nb, Nt, Nx)))
    for t in fitting_tiles(ctx, prec, Nx, Nt, nb):
        ctx.set_tile(prec, Nx, Nt, nb, t)
        cg = nsl.CG(fm, nsl.MMdagger, 1e-12, 100000, precision=prec, tol_kind=nsl.TOL_REL_PER_RHS_FLOOR)
        x = cg(b.to(DEV))
        assert cg.converged, (t, cg.report)
        target = max(1e-12, cg.report["max_floor_rel"])
        r = b - oracle_apply(ofs, "MMdagger", x.cpu())
        for k in range(nb):
            assert float(r[k].norm() / b[k].norm()) < 2 * target, (t, k)
        if prec == nsl.F64:
            cg = nsl.CG(fm, nsl.MMdagger, 1e-10, 100000)          # reference criterion
            x = cg(b.to(DEV))
            assert cg.converged, (t, cg.report)
            r = b - oracle_apply(ofs, "MMdagger", x.cpu())
            assert float(r.norm()) < 1e-8, t


def test_tune_records_choice():
    ctx = nsl.Context(0)
    fm, ofs, Nx = setup("honeycomb", (12, 12), 64, 8.0, 2, ctx)
    nb = 8
    for prec in (nsl.F64, nsl.F32, nsl.BF16X3):
        t, us = fm.tune(prec, nb, reps=3)
        assert 0 <= t < nsl.Context.tile_count(prec)
        assert us > 0
        assert ctx.get_tile(prec, Nx, 64, nb) == t
        # CG mode: the fused CGA + CGB pair, classic tiles 0-3
        t2, us2 = fm.tune(prec, nb, reps=3, cg=True)
        assert 0 <= t2 < 4 and us2 > 0
        assert ctx.get_tile(prec, Nx, 64, nb) == t2
