"""TwoPointCorrelator::measure(N) arithmetic (CPU oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows src/NSL/Measurements/Hubbard/twoPointCorrelationFunction.tpp:116-162:
for tsrc = 0, Nt/N, ...: point sources S[x, tsrc, x] = 1 (:136), X = CG on
MM^dagger (:139, CG defaults eps=1e-12), G = M^dagger X (:143),
G.shift(-tsrc, -2, -1) (antiperiodic re-centring, shift.tpp:63-68),
corr[t, y, x] += G[x, t, y] (:152), corr /= N (:160).

Parity pin (tests/test_oracle_core.py::test_correlator_free_propagator): at phi = 0 the correlator is
the free propagator, computed independently as (M_dense)^{-1} and as the
Matsubara sum over the eigenbasis of K.
"""

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
