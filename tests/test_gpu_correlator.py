"""TwoPointCorrelator::measure on the device against the oracle
(twoPointCorrelationFunction.tpp:116-162).  Tolerances: the reference mode
solves each time source to sum|r|^2 <= 1e-24 like the oracle, so the two
correlators agree to ~|r| / lambda_min(MM^dag); 1e-9 relative is asked on
the CPU-runnable configuration.  Per-RHS and mixed-precision modes are held
to their solve tolerance against the FP64 reference-mode result."""

import math

import pytest
import torch

import nsl_amd as nsl
from oracle import correlator as ocorr, fermion as F, lattice as olat, rng

pytestmark = pytest.mark.gpu

DEV = "cuda"


def relerr(a, b):
    a = a.detach().to("cpu", torch.complex128)
    b = torch.as_tensor(b).detach().to("cpu", torch.complex128)
    return float((a - b).abs().max() / b.abs().max())


def setup(L=3, Nt=16, beta=4.0, U=2.0, species=nsl.Particle, zero=False, seed=1000):
    lat_d, lat_o = nsl.Lattice.honeycomb(L, L), olat.honeycomb(L, L)
    Nx = lat_d.sites()
    if zero:
        phi = torch.zeros((Nt, Nx), dtype=torch.complex128)
    else:
        phi = torch.as_tensor(rng.real_field(seed, 0, (Nt, Nx), math.sqrt(beta / Nt * U)))
    fm = nsl.HubbardExp(lat_d, Nt, beta, 0.0, species)
    fm.populate(phi.to(DEV))
    ofm = F.HubbardExp(lat_o, Nt, beta, 0.0, species)
    ofm.populate(phi)
    return fm, ofm, Nt, Nx


@pytest.mark.parametrize("species", [nsl.Particle, nsl.Hole])
@pytest.mark.parametrize("ntsrc", [1, 4, 3, 16])      # 3 does not divide Nt = 16 (sources 0, 5, 10, 15)
def test_two_point_reference_mode_cfg0(species, ntsrc):
    fm, ofm, Nt, Nx = setup(species=species)
    corr = nsl.TwoPointCorrelator(fm).measure(ntsrc)
    ref, _ = ocorr.two_point(ofm, Nt, Nx, ntsrc)
    assert corr.shape == (Nt, Nx, Nx)
    assert relerr(corr, ref) < 1e-9


def test_two_point_free_field():
    """phi = 0: the oracle's correlator is pinned to the free propagator
    (Matsubara sum); the device must match it."""
    fm, ofm, Nt, Nx = setup(zero=True)
    corr = nsl.TwoPointCorrelator(fm).measure(Nt)
    ref, _ = ocorr.two_point(ofm, Nt, Nx, Nt)
    assert relerr(corr, ref) < 1e-10


@pytest.mark.parametrize("prec,tol", [(nsl.F64, 1e-9), (nsl.F32, 1e-8), (nsl.BF16X3, 1e-8)])
def test_two_point_per_rhs_and_mixed(prec, tol):
    """Per-RHS relative stop at 1e-11 (FP64 true residual for the low
    precisions), all time sources in one batch and one source per batch:
    each against the reference-mode FP64 correlator."""
    fm, ofm, Nt, Nx = setup(L=6, Nt=16, beta=4.0)
    ref = nsl.TwoPointCorrelator(fm).measure(4)
    for max_rhs in (0, Nx):
        tp = nsl.TwoPointCorrelator(fm, eps=1e-11, maxIter=100000, tol_kind=nsl.TOL_REL_PER_RHS, precision=prec,
                                    max_rhs=max_rhs)
        corr = tp.measure(4)
        assert tp.converged, tp.report
        assert relerr(corr, ref) < tol, (max_rhs, relerr(corr, ref))


def test_two_point_errors():
    fm, ofm, Nt, Nx = setup()
    with pytest.raises(nsl.NslxError):
        nsl.TwoPointCorrelator(fm).measure(Nt + 1)
    with pytest.raises(nsl.NslxError):                       # low precision needs the per-RHS criterion
        nsl.TwoPointCorrelator(fm, precision=nsl.F32).measure(2)
    lat = nsl.Lattice.honeycomb(3, 3)
    fm2 = nsl.HubbardExp(lat, Nt, 4.0)
    fm2.populate(torch.zeros((2, Nt, Nx), dtype=torch.complex128, device=DEV))   # two fields
    with pytest.raises(nsl.NslxError):
        nsl.TwoPointCorrelator(fm2).measure(2)
