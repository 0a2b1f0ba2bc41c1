"""Wall-source correlators on the device against the oracle:
TwoPointCorrelator::measureK(k, N) (nslx_correlator_two_point_wall) and
TwoBodyCorrelator::measure(N) (nslx_correlator_two_body, twelve channels,
momentum quadruples from nslx_two_body_quads)."""

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


def walls(L, ks):
    """w[k, s, x] = exp(i k.r_x) / L on sublattice s; k = (m C + n D) / L."""
    lat = olat.honeycomb(L, L)
    pos = torch.as_tensor(lat.sites_xyz[:, :2], dtype=torch.float64)
    C = torch.tensor([4 * math.pi / 3, 0.0], dtype=torch.float64)
    D = torch.tensor([-2 * math.pi / 3, 2 * math.pi / math.sqrt(3.0)], dtype=torch.float64)
    mom = torch.stack([(m * C + n * D) / L for (m, n) in ks])
    Nx = lat.sites()
    w = torch.zeros((len(ks), 2, Nx), dtype=torch.complex128)
    for s in range(2):
        sel = torch.arange(Nx) % 2 == s
        w[:, s, sel] = torch.exp(1j * (mom @ pos[sel].T)) / L
    return lat, w, mom


def setup(L, Nt, beta, U, seed, cls=nsl.HubbardExp, ocls=F.HubbardExp):
    lat_d = nsl.Lattice.honeycomb(L, L)
    lat_o = olat.honeycomb(L, L)
    phi = torch.as_tensor(rng.real_field(seed, 0, (Nt, lat_d.sites()), math.sqrt(beta / Nt * U)))
    fm = cls(lat_d, Nt, beta)
    fm.populate(phi.to(DEV))
    o = ocls(lat_o, Nt, beta)
    o.populate(phi)
    return fm, o, lat_o, phi


@pytest.mark.parametrize("cls,ocls", [(nsl.HubbardExp, F.HubbardExp), (nsl.HubbardDiag, F.HubbardDiag)])
def test_measureK_vs_oracle(cls, ocls):
    L, Nt = 3, 16
    fm, o, _, _ = setup(L, Nt, 4.0, 2.0, 1000, cls, ocls)
    _, w, _ = walls(L, [(0, 0), (1, 0), (1, 2)])
    tpc = nsl.TwoPointCorrelator(fm, eps=1e-12)
    for k in range(w.shape[0]):
        corrK, block = tpc.measureK(w[k], 2)
        rK, rB = ocorr.two_point_wall(o, Nt, 2, w[k])
        assert tpc.converged
        assert relerr(corrK, rK) < 1e-9, k
        assert relerr(block, rB) < 1e-9, k


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32, nsl.BF16X3])
def test_measureK_per_rhs_precision(prec):
    L, Nt = 3, 16
    fm, o, _, _ = setup(L, Nt, 4.0, 2.0, 1001)
    _, w, _ = walls(L, [(2, 1)])
    tpc = nsl.TwoPointCorrelator(fm, eps=1e-11, tol_kind=nsl.TOL_REL_PER_RHS, precision=prec)
    corrK, block = tpc.measureK(w[0], 4)
    rK, rB = ocorr.two_point_wall(o, Nt, 4, w[0])
    assert tpc.converged
    assert relerr(corrK, rK) < 1e-8
    assert relerr(block, rB) < 1e-8


def test_two_body_quads_match_oracle():
    _, _, mom = walls(3, [(m, n) for m in range(3) for n in range(3)])
    q = nsl.two_body_quads(mom.numpy())
    ref = ocorr.two_body_quads(mom)
    assert [tuple(int(v) for v in r) for r in q] == ref
    assert len(ref) == 9 ** 3


def test_two_body_vs_oracle():
    """Honeycomb 3x3, Nt = 8, beta = 2, U = 2; momenta (0,0), (1,0), (2,0)
    (and so every conserving quadruple among them), two bands, two time
    sources: all twelve channels against the oracle."""
    L, Nt, beta, U = 3, 8, 2.0, 2.0
    fm, o, lat_o, phi = setup(L, Nt, beta, U, 1002)
    _, w, mom = walls(L, [(0, 0), (1, 0), (2, 0)])
    tb = nsl.TwoBodyCorrelator(fm, w, mom.numpy(), eps=1e-12)
    res = tb.measure(2)
    assert tb.converged
    species_before = fm.species()

    def make(species):
        h = F.HubbardExp(lat_o, Nt, beta, species=species)
        h.populate(phi)
        return h

    ref = ocorr.two_body(make, phi, w, mom, 2)
    assert set(res) == set(ref)
    for name in ref:
        assert set(res[name]) == set(ref[name]), name
        for q, v in ref[name].items():
            scale = max(float(v.abs().max()), 1e-300)
            assert float((res[name][q].cpu() - v).abs().max()) <= 1e-9 * scale, (name, q)
    # the device matrix is left as it was populated (particle)
    assert fm.species() == species_before
    x = torch.as_tensor(rng.complex_normal(5, 0, (1, Nt, 18)))
    assert relerr(fm.apply(nsl.M, x.to(DEV)), o.M(x)) < 1e-13
