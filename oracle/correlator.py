"""TwoPointCorrelator::measure(N) / measureK(k, N) and
TwoBodyCorrelator::measure(N) arithmetic (CPU oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows src/NSL/Measurements/Hubbard/twoPointCorrelationFunction.tpp:116-162:
for tsrc = 0, Nt/N, ...: point sources S[x, tsrc, x] = 1 (:136), X = CG on
MM^dagger (:139, CG defaults eps=1e-12), G = M^dagger X (:143),
G.shift(-tsrc, -2, -1) (antiperiodic re-centring, shift.tpp:63-68),
corr[t, y, x] += G[x, t, y] (:152), corr /= N (:160).

Parity pin (tests/test_oracle_core.py::test_correlator_free_propagator): at phi = 0 the correlator is
the free propagator, computed independently as (M_dense)^{-1} and as the
Matsubara sum over the eigenbasis of K.  measureK and the two-body pools
(tests/test_oracle_core.py::test_wall_correlators_linearity): by linearity
of the solves, the wall-source propagators are the point-source correlator
contracted with the walls, corrK[s, t, y] = sum_x wall[s, x] corr[t, y, x],
and the two-body pools are its projections (for both species and, for the
M^dagger M pools, the transposed propagator).  The twelve contraction
channels (TWO_BODY_TABLE) are restated from the reference and read only
those pools; their exchange structure is pinned by Fermi antisymmetry
(tests/test_oracle_core.py::test_two_body_pauli: for any field the eight
I = 1, S = 1 channels vanish where the two sink or the two source labels
coincide, the others do not), which a wrong exchange sign, a transposed
exchange index or a swapped pool pair fails.  The prefactors alpha are
pinned pairwise by particle-hole symmetry
(tests/test_oracle_core.py::test_two_body_particle_hole_pairs: at phi = 0,
mu = 0 on the bipartite honeycomb each Iz = +1 channel equals its Iz = -1
partner times the sublattice signs of its four band labels, and the Iz = 0
channels with Sz = +1 and -1 are equal; the six pairs cover all twelve
channels, so a wrong alpha in any one channel fails).  A factor common to
both channels of a pair remains parity unpinned (no independent value of a
channel exists without running the reference).
"""

import math

import torch

from . import solvers
from .fermion import _shift


def two_point(hfm, Nt, Nx, n_tsrc, eps=1e-12, maxIter=10000, solve=None):
    corr = torch.zeros((Nt, Nx, Nx), dtype=torch.complex128)
    step = Nt // n_tsrc
    iters = []
    for tsrc in range(0, Nt, step):
        src = torch.zeros((Nx, Nt, Nx), dtype=torch.complex128)
        src[torch.arange(Nx), tsrc, torch.arange(Nx)] = 1.0
        if solve is None:
            X, info = solvers.cg(hfm.MMdagger, src, eps, maxIter)
            iters.append(info["iterations"])
        else:
            X = solve(src)
        G = hfm.Mdagger(X)
        G = _shift(G, -tsrc, -2, -1.0)
        corr += G.permute(1, 2, 0)        # corr[t,y,x] = G[x,t,y]
    corr /= n_tsrc
    return corr, iters


def source_slice(Nt, Nx, tsrc):
    """The Nx point sources of one time source, shape (Nx, Nt, Nx)."""
    src = torch.zeros((Nx, Nt, Nx), dtype=torch.complex128)
    src[torch.arange(Nx), tsrc, torch.arange(Nx)] = 1.0
    return src


def recentre(G, tsrc):
    """G (Nx,Nt,Nx) -> contribution corr[t,y,x] for one time source."""
    return _shift(G, -tsrc, -2, -1.0).permute(1, 2, 0)


def two_point_wall(hfm, Nt, n_tsrc, wall, eps=1e-12, maxIter=10000):
    """TwoPointCorrelator::measureK(k, N) (twoPointCorrelationFunction.tpp:
    169-233) for wall = wallSources[k] (uDim, Nx): srcVecK(:, tsrc, :) = wall
    (:188), CG on MM^dagger (:191), M^dagger (:195), shift(-tsrc, -2, -1)
    (:200), corrK += .; corrK /= N (:205); corrKblock[t, s1, s2] =
    sum_x conj(wall[s1, x]) corrK[s2, t, x] (:208-214)."""
    wall = torch.as_tensor(wall, dtype=torch.complex128)
    nw, Nx = wall.shape
    corrK = torch.zeros((nw, Nt, Nx), dtype=torch.complex128)
    step = Nt // n_tsrc
    for tsrc in range(0, Nt, step):
        src = torch.zeros((nw, Nt, Nx), dtype=torch.complex128)
        src[:, tsrc, :] = wall
        X, _ = solvers.cg(hfm.MMdagger, src, eps, maxIter)
        corrK += _shift(hfm.Mdagger(X), -tsrc, -2, -1.0)
    corrK /= n_tsrc
    block = torch.einsum("ax,btx->tab", torch.conj(wall), corrK)
    return corrK, block


def modbz_equal(k, q, eps=2e-12):
    """eq_modBZ_ (twoBodyCorrelationFunction.tpp:121-135)."""
    C = (4 * math.pi / 3, 0.0)
    D = (-2 * math.pi / 3, 2 * math.pi / math.sqrt(3.0))
    dx, dy = k[0] - q[0], k[1] - q[1]
    den = C[1] * D[0] - C[0] * D[1]
    m = (dy * D[0] - dx * D[1]) / den
    n = (dx * C[1] - dy * C[0]) / den
    return abs(round(m) - m) < eps and abs(round(n) - n) < eps


def two_body_quads(momenta):
    """(w, x, z, y) keys of measure(N) (:139-151), loop order w, x, y, z."""
    k = [tuple(float(v) for v in row) for row in momenta]
    out = []
    for w in range(len(k)):
        for x in range(len(k)):
            for y in range(len(k)):
                for z in range(len(k)):
                    if modbz_equal((k[w][0] + k[x][0], k[w][1] + k[x][1]), (k[z][0] + k[y][0], k[z][1] + k[y][1])):
                        out.append((w, x, z, y))
    return out


# (channel, alpha, terms): a term is (X, Y, exchange, sign) over the pools
# "CP", "CH" (corrPool_ particle / hole) and "DP", "DH" (corrPoolDag_); the
# direct pattern is X(w,z,t,i,l) Y(x,y,t,j,k), the exchange pattern
# X(w,y,t,i,k) Y(x,z,t,j,l) (twoBodyCorrelationFunction.tpp:481-1000).
TWO_BODY_TABLE = (
    ("I1S1Iz1Sz1", 1.0, (("DP", "DP", 0, 1), ("DP", "DP", 1, -1))),
    ("I1S1Iz1Sz0", -0.5, (("DP", "CH", 0, 1), ("DP", "CH", 1, -1), ("CH", "DP", 1, -1), ("CH", "DP", 0, 1))),
    ("I1S1Iz1Szn1", 1.0, (("CH", "CH", 0, 1), ("CH", "CH", 1, -1))),
    ("I1S1Iz0Sz1", 0.5, (("DP", "DH", 0, 1), ("DP", "DH", 1, -1), ("DH", "DP", 1, -1), ("DH", "DP", 0, 1))),
    ("I1S1Iz0Szn1", 0.5, (("CP", "CH", 0, 1), ("CP", "CH", 1, -1), ("CH", "CP", 1, -1), ("CH", "CP", 0, 1))),
    ("I1S1Izn1Sz1", 1.0, (("DH", "DH", 0, 1), ("DH", "DH", 1, -1))),
    ("I1S1Izn1Sz0", -0.5, (("CP", "DH", 0, 1), ("CP", "DH", 1, -1), ("DH", "CP", 1, -1), ("DH", "CP", 0, 1))),
    ("I1S1Izn1Szn1", 1.0, (("CP", "CP", 0, 1), ("CP", "CP", 1, -1))),
    ("I0S1Iz0Sz1", 0.5, (("DP", "DH", 0, 1), ("DP", "DH", 1, 1), ("DH", "DP", 1, 1), ("DH", "DP", 0, 1))),
    ("I0S1Iz0Szn1", 0.5, (("CP", "CH", 0, 1), ("CP", "CH", 1, 1), ("CH", "CP", 1, 1), ("CH", "CP", 0, 1))),
    ("I1S0Iz1Sz0", -0.5, (("DP", "CH", 0, 1), ("DP", "CH", 1, 1), ("CH", "DP", 1, 1), ("CH", "DP", 0, 1))),
    ("I1S0Izn1Sz0", -0.5, (("CP", "DH", 0, 1), ("CP", "DH", 1, 1), ("DH", "CP", 1, 1), ("DH", "CP", 0, 1))),
)


def two_body_pools(make_fm, phi, wall, tsrc, eps=1e-12, maxIter=10000):
    """corrPool_ and corrPoolDag_ of one time source (:256-283): per species,
    CG on MM^dagger of the wall sources and on MdaggerM of their conjugates,
    M^dagger / M back-multiplication, shift(-tsrc, -2, -1), projection on the
    walls.  make_fm(species) returns a populated oracle fermion matrix.
    Pools have shape (kSink, kSrc, Nt, sSink, sSrc)."""
    wall = torch.as_tensor(wall, dtype=torch.complex128)
    kdim, bdim, Nx = wall.shape
    Nt = phi.shape[0]
    src = torch.zeros((kdim, bdim, Nt, Nx), dtype=torch.complex128)
    src[:, :, tsrc, :] = wall
    pools = {}
    from .fermion import PARTICLE, HOLE
    for name, species in (("P", PARTICLE), ("H", HOLE)):
        h = make_fm(species)
        X, _ = solvers.cg(h.MMdagger, src, eps, maxIter)
        Xd, _ = solvers.cg(h.MdaggerM, torch.conj(src), eps, maxIter)
        G = _shift(h.Mdagger(X), -tsrc, -2, -1.0)
        Gd = _shift(h.M(Xd), -tsrc, -2, -1.0)
        # corrKPool(kS, kR, t, sS, sR) = sum_x conj(w[kS,sS,x]) G[kR,sR,t,x]
        pools["C" + name] = torch.einsum("aix,bjtx->abtij", torch.conj(wall), G)
        # corrPoolDag = conj(sum_x w[kS,sS,x] Gd[kR,sR,t,x])
        pools["D" + name] = torch.conj(torch.einsum("aix,bjtx->abtij", wall, Gd))
    return pools


def two_body(make_fm, phi, wall, momenta, n_tsrc, eps=1e-12, maxIter=10000):
    """TwoBodyCorrelator::measure(N): {channel: {(w, x, z, y): (Nt, b^2, b^2)}}."""
    wall = torch.as_tensor(wall, dtype=torch.complex128)
    kdim, bdim, Nx = wall.shape
    Nt = phi.shape[0]
    quads = two_body_quads(momenta)
    out = {name: {q: torch.zeros((Nt, bdim * bdim, bdim * bdim), dtype=torch.complex128) for q in quads}
           for name, _, _ in TWO_BODY_TABLE}
    step = Nt // n_tsrc
    for tsrc in range(0, Nt, step):
        P = two_body_pools(make_fm, phi, wall, tsrc, eps, maxIter)
        for name, alpha, terms in TWO_BODY_TABLE:
            for (w, x, z, y) in quads:
                c = out[name][(w, x, z, y)]
                for i in range(bdim):
                    for j in range(bdim):
                        for k in range(bdim):
                            for l in range(bdim):
                                acc = 0
                                for (X, Y, exch, sign) in terms:
                                    if exch:
                                        acc = acc + sign * P[X][w, y, :, i, k] * P[Y][x, z, :, j, l]
                                    else:
                                        acc = acc + sign * P[X][w, z, :, i, l] * P[Y][x, y, :, j, k]
                                c[:, i * bdim + j, l * bdim + k] += alpha * acc / n_tsrc
    return out
