"""Complex fields and a complex chemical potential on the device, against the
oracle (the reference's M tests draw phi with phi.rand() on a complex tensor,
test_fermionMatrixHubbardExp.cpp:150-151, so Re phi and Im phi are both
random; its ctors take `const Type & beta, const Type & mu` with Type
complex, hubbardExp.hpp:48-76).

Tolerances: 1e-13 relative for FP64 applications, 2e-6 / 4e-5 for FP32 /
BF16x3, log det to 1e-10 relative (imaginary part mod 2 pi), its gradient to
1e-11, solutions through their FP64 true residual <= 1e-12 recomputed with
the oracle's operator.  Im phi ~ N(0, 0.25) makes |P| = e^{-s Im phi} differ
from 1 by up to ~2x, which the real-phi tests never exercise."""

import math

import numpy as np
import pytest
import torch

import nsl_amd as nsl
from oracle import fermion as F, lattice as olat, rng, solvers as osol

pytestmark = pytest.mark.gpu

DEV = "cuda"
OPS = ["M", "Mdagger", "MMdagger", "MdaggerM"]


def relerr(a, b):
    a = a.detach().to("cpu", torch.complex128)
    b = torch.as_tensor(b).detach().to("cpu", torch.complex128)
    return float((a - b).abs().max() / b.abs().max())


def lattices(kind, size):
    if kind == "honeycomb":
        return nsl.Lattice.honeycomb(*size), olat.honeycomb(*size)
    if kind == "ring":
        return nsl.Lattice.ring(size), olat.ring(size)
    return nsl.Lattice.armchair_nanotube(*size), olat.armchair_nanotube(*size)


def complex_field(seed, Nt, Nx, re_std=0.7, im_std=0.25):
    re = torch.as_tensor(rng.real_field(seed, 0, (Nt, Nx), re_std))
    im = torch.as_tensor(rng.real_field(seed + 7919, 0, (Nt, Nx), im_std))
    return (re + 1j * im).to(torch.complex128)


def make(kind, size, Nt, beta, nfield=1, seed=1300, mu=0.0, species=nsl.Particle, cls=nsl.HubbardExp,
         ocls=F.HubbardExp, im_std=0.25):
    lat_d, lat_o = lattices(kind, size)
    Nx = lat_d.sites()
    phis = [complex_field(seed + f, Nt, Nx, im_std=im_std) for f in range(nfield)]
    fm = cls(lat_d, Nt, beta, mu, species)
    fm.populate(torch.stack(phis).to(DEV))
    ofs = []
    for p in phis:
        o = ocls(lat_o, Nt, beta, mu, species)
        o.populate(p)
        ofs.append(o)
    return fm, ofs, lat_d


def oracle_apply(ofs, name, psi):
    rpf = psi.shape[0] // len(ofs)
    return torch.stack([getattr(ofs[k // rpf], name)(psi[k]) for k in range(psi.shape[0])])


APPLY_CASES = [
    ("honeycomb", (3, 3), 16, 1, 3, 0.0),
    ("ring", 10, 14, 2, 4, 0.05),
    ("nanotube", (6, 4), 32, 1, 2, 0.0),
    ("honeycomb", (12, 12), 64, 2, 4, 0.0),           # the Nx = 288, Nt = 64 shape
    ("honeycomb", (3, 3), 16, 1, 2, 0.2 + 0.4j),      # complex mu on top of complex phi
    ("ring", 6, 8, 1, 2, 0.7j),                       # purely imaginary mu
]


@pytest.mark.parametrize("case", APPLY_CASES, ids=[f"{c[0]}{c[1]}-Nt{c[2]}-mu{c[5]}" for c in APPLY_CASES])
@pytest.mark.parametrize("species", [nsl.Particle, nsl.Hole])
def test_apply_complex_phi_f64(case, species):
    kind, size, Nt, nfield, nrhs, mu = case
    fm, ofs, lat = make(kind, size, Nt, 4.0, nfield, mu=mu, species=species)
    psi = torch.as_tensor(rng.complex_normal(91, 0, (nrhs, Nt, lat.sites())))
    for name in OPS:
        got = getattr(fm, name)(psi.to(DEV))
        e = relerr(got, oracle_apply(ofs, name, psi))
        assert e < 1e-13, (name, e)


@pytest.mark.parametrize("prec,tol", [(nsl.F32, 2e-6), (nsl.BF16X3, 4e-5)])
def test_apply_complex_phi_low_precision(prec, tol):
    fm, ofs, lat = make("honeycomb", (6, 6), 32, 4.0, 1, mu=0.1 + 0.3j)
    psi = torch.as_tensor(rng.complex_normal(92, 0, (2, 32, lat.sites())))
    for name in OPS:
        got = getattr(fm, name)(psi.to(DEV), precision=prec)
        e = relerr(got, oracle_apply(ofs, name, psi))
        assert e < tol, (name, e)


def test_apply_complex_phi_diag():
    fm, ofs, lat = make("honeycomb", (3, 3), 16, 4.0, 2, mu=0.3j, cls=nsl.HubbardDiag, ocls=F.HubbardDiag)
    psi = torch.as_tensor(rng.complex_normal(93, 0, (4, 16, lat.sites())))
    for name in OPS:
        got = getattr(fm, name)(psi.to(DEV))
        e = relerr(got, oracle_apply(ofs, name, psi))
        assert e < 1e-13, (name, e)


@pytest.mark.parametrize("species", [nsl.Particle, nsl.Hole])
@pytest.mark.parametrize("mu", [0.0, 0.25 - 0.5j])
def test_logdet_grad_complex_phi(species, mu):
    fm, ofs, _ = make("honeycomb", (3, 3), 16, 4.0, 1, mu=mu, species=species)
    ld = fm.logDetM()
    ref = ofs[0].logDetM()
    assert abs(ld.real - ref.real) < 1e-10 * abs(ref.real)
    assert abs(math.remainder(ld.imag - ref.imag, 2 * math.pi)) < 1e-9
    g = fm.gradLogDetM()
    assert relerr(g, ofs[0].gradLogDetM()) < 1e-11


def test_logdet_complex_phi_diag():
    fm, ofs, _ = make("ring", 6, 12, 3.0, 1, cls=nsl.HubbardDiag, ocls=F.HubbardDiag)
    ld = fm.logDetM()
    ref = ofs[0].logDetM()
    assert abs(ld.real - ref.real) < 1e-10 * max(1.0, abs(ref.real))
    assert abs(math.remainder(ld.imag - ref.imag, 2 * math.pi)) < 1e-9
    assert relerr(fm.gradLogDetM(), ofs[0].gradLogDetM()) < 1e-11


@pytest.mark.parametrize("op", [nsl.MMdagger, nsl.MdaggerM])
@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32])
def test_cg_complex_phi_true_residual(op, prec):
    """Per-RHS CG (reliable updates for F32) on a complex field with complex
    mu: every solution's FP64 true residual, recomputed with the oracle's
    operator, is <= 1e-12 relative."""
    fm, ofs, lat = make("honeycomb", (3, 3), 16, 4.0, 2, mu=0.1 + 0.2j)
    b = torch.as_tensor(rng.complex_normal(94, 0, (4, 16, lat.sites())))
    cg = nsl.CG(fm, op, 1e-12, 20000, precision=prec, tol_kind=nsl.TOL_REL_PER_RHS)
    x = cg(b.to(DEV)).cpu()
    name = "MMdagger" if op == nsl.MMdagger else "MdaggerM"
    r = b - oracle_apply(ofs, name, x)
    rel = (r.abs() ** 2).sum((-1, -2)).sqrt() / (b.abs() ** 2).sum((-1, -2)).sqrt()
    assert cg.converged
    assert float(rel.max()) <= 1e-12, rel


def test_cg_reference_semantics_complex_phi():
    """The reference's CG (x0 = b, batch-global alpha/beta, |r|^2 <= eps^2)
    on a complex field: the solution matches the oracle's CG."""
    fm, ofs, lat = make("honeycomb", (3, 3), 16, 4.0, 1, mu=0.3j)
    b = torch.as_tensor(rng.complex_normal(95, 0, (2, 16, lat.sites())))
    cg = nsl.CG(fm, nsl.MMdagger, 1e-10, 10000)
    x = cg(b.to(DEV))
    xo, _ = osol.cg(ofs[0].MMdagger, b, 1e-10)
    assert cg.converged
    assert relerr(x, xo) < 1e-8


def test_bicgstab_complex_phi():
    """The reference's BiCGStab (x0 = b, batch-global scalars) on MMdagger of
    a complex field with complex mu: the oracle's BiCGStab converges in 342
    iterations here; the device's solution matches it and its true residual
    is <= 1e-10 relative.  (On M itself the reference's recurrence breaks
    down for such fields -- the oracle stalls or divides by zero -- so M is
    not a case.)"""
    fm, ofs, lat = make("honeycomb", (3, 3), 16, 4.0, 1, mu=0.2 - 0.1j)
    b = torch.as_tensor(rng.complex_normal(96, 0, (2, 16, lat.sites())))
    bs = nsl.BiCGStab(fm, nsl.MMdagger, 1e-10, 2000)
    x = bs(b.to(DEV)).cpu()
    xo, info = osol.bicgstab(ofs[0].MMdagger, b, 1e-10, 2000)
    assert bs.converged and info["converged"]
    assert relerr(x, xo) < 1e-8
    r = b - oracle_apply(ofs, "MMdagger", x)
    assert float(r.norm() / b.norm()) < 1e-10


def test_complex_mu_is_an_imaginary_field_shift():
    """mu~ = delta mu enters only the phases: P(phi, mu) = P(phi - i mu~, 0),
    so M at complex mu equals M at mu = 0 with phi shifted by -i delta mu --
    a check of the device's phase table that does not use the oracle."""
    lat = nsl.Lattice.honeycomb(3, 3)
    Nt, beta, mu = 16, 4.0, 0.4 - 0.3j
    phi = complex_field(1400, Nt, 18)
    psi = torch.as_tensor(rng.complex_normal(97, 0, (3, Nt, 18))).to(DEV)
    a = nsl.HubbardExp(lat, Nt, beta, mu)
    a.populate(phi.to(DEV))
    b = nsl.HubbardExp(lat, Nt, beta, 0.0)
    b.populate((phi - 1j * (beta / Nt) * mu).to(DEV))
    for name in OPS:
        assert relerr(getattr(a, name)(psi), getattr(b, name)(psi)) < 1e-13, name
    la, lb = a.logDetM(), b.logDetM()
    assert abs(la.real - lb.real) < 1e-11 * abs(lb.real)
    assert abs(math.remainder(la.imag - lb.imag, 2 * math.pi)) < 1e-9


def test_complex_beta_rejected():
    """A beta with a nonzero imaginary part is refused at construction
    (NSLX_E_ARG); a complex beta with zero imaginary part is taken."""
    lat = nsl.Lattice.honeycomb(3, 3)
    for cls in (nsl.HubbardExp, nsl.HubbardDiag):
        with pytest.raises(nsl.NslxError):
            cls(lat, 16, 4.0 + 0.5j)
        fm = cls(lat, 16, 4.0 + 0j, 0.1j)
        fm.populate(torch.zeros((16, 18), dtype=torch.complex128, device=DEV))
        assert fm.M(torch.ones((16, 18), dtype=torch.complex128, device=DEV)).shape == (16, 18)
