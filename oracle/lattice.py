"""Spatial lattices and the memoised exponential hopping matrix (CPU oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows:
* Honeycomb: src/NSL/Lattice/Implementations/honeycomb.tpp:47-69 (periodic
  minimum-image distance) and :72-117 (site coordinates, cell index
  i = L1 + L[0]*L2, A = 2i at +r/2, B = 2i+1 at -r/2, hop kappa when the
  distance^2 is within 1e-4 of 1).
* Ring: src/NSL/Lattice/Implementations/ring.tpp:8-42 (including the
  overwrite order of the periodic corner entries).
* hopping_matrix / exp_hopping_matrix: src/NSL/Lattice/lattice.tpp:46-58, with
  the memo keyed by the (complex) exponent as in lattice.hpp:127.
* ArmchairNanotube: NOT in the reference (it has no nanotube lattice; its only
  route would be a YAML adjacency through Lattice::Generic).  Project-defined:
  graphene with the same basis as the honeycomb (a0=(3/2, sqrt3/2),
  a1=(3/2,-sqrt3/2), bond r=(1,0)) rolled along C = n(a0+a1) = (3n,0), periodic
  in x with period 3n, open along y over `cells` translation periods of
  length sqrt3; 4n sites per cell.

Parity: pinned by closed forms in tests/test_oracle_core.py (honeycomb:
valence 3, spectrum in [-3,3], Bloch spectrum |1 + e^{ik.a0} + e^{ik.a1}|,
4 zero modes iff L0 and L1 are multiples of 3; E vs eigendecomposition;
E(a)E(-a) = I).  Nanotube: site count, bipartiteness, valence (3 inside,
2 at the open ends) and the infinite-tube limit of its spectrum.
"""

import math
import numpy as np
import torch

SQRT3_2 = 0.8660254037844386


class SpatialLattice:
    """Hopping matrix K (Nx x Nx, complex128) + memoised E(a) = exp(a K)."""

    def __init__(self, name, hops, sites):
        self.name = name
        self.hops = torch.as_tensor(hops, dtype=torch.complex128)
        self.sites_xyz = np.asarray(sites, dtype=np.float64)
        self._exp_memo = {}

    def sites(self):
        return self.hops.shape[0]

    def hopping_matrix(self, delta=1.0):
        # lattice.tpp:46-49
        return self.hops if delta == 1.0 else self.hops * delta

    def exp_hopping_matrix(self, delta):
        # lattice.tpp:51-58: memo keyed by the complex exponent
        key = complex(delta)
        if key not in self._exp_memo:
            self._exp_memo[key] = torch.linalg.matrix_exp(self.hopping_matrix(key))
        return self._exp_memo[key]


def _honeycomb_sites(L0, L1):
    a0 = np.array([1.5, SQRT3_2])
    a1 = np.array([1.5, -SQRT3_2])
    r = np.array([1.0, 0.0])
    n = 2 * L0 * L1
    sites = np.zeros((n, 2))
    for l1 in range(L0):
        for l2 in range(L1):
            i = l1 + L0 * l2
            sites[2 * i] = l1 * a0 + l2 * a1 + r / 2
            sites[2 * i + 1] = l1 * a0 + l2 * a1 - r / 2
    return sites, a0, a1


def honeycomb(L0, L1=None, kappa=1.0):
    """Periodic honeycomb of L0 x L1 unit cells (honeycomb.tpp:72-138)."""
    if L1 is None:
        L1 = L0
    sites, a0, a1 = _honeycomb_sites(L0, L1)
    P0 = L0 * a0
    P1 = L1 * a1
    n = sites.shape[0]
    # minimum-image distance^2 over the 3x3 periodic images (honeycomb.tpp:47-69)
    d = sites[:, None, :] - sites[None, :, :]
    best = np.full((n, n), float(np.dot(P0 + P1, P0 + P1)))
    for s0 in (-1, 0, 1):
        for s1 in (-1, 0, 1):
            v = d + s0 * P0 + s1 * P1
            best = np.minimum(best, np.einsum("ijk,ijk->ij", v, v))
    K = np.zeros((n, n), dtype=np.complex128)
    K[np.abs(best - 1.0) < 1e-4] = kappa
    return SpatialLattice(f"Honeycomb({L0},{L1})", K, sites)


def ring(n, kappa=1.0, radius=1.0):
    """Ring of n sites (ring.tpp:8-42), including its corner overwrite order."""
    K = np.zeros((n, n), dtype=np.complex128)
    theta = 2 * math.pi / n
    sites = np.array([[radius * math.cos(i * theta), radius * math.sin(i * theta), 0.0]
                      for i in range(n)])
    for i in range(n - 1):
        K[i, i + 1] = kappa
    for i in range(1, n):
        K[i, i - 1] = np.conj(kappa)
    if n > 1:
        K[0, n - 1] = np.conj(kappa)
        K[n - 1, 0] = kappa
    else:
        K[0, 0] = 0
    return SpatialLattice(f"Ring({n})", K, sites)


def armchair_nanotube_sites(n, cells):
    """Site coordinates (x along the circumference in [0,3n), y along the
    axis in [0, cells*sqrt3)), sorted by the integer keys
    (round(1e6*y), round(1e6*x)).  Project definition; the C++ host builder
    uses the same rule."""
    a0 = np.array([1.5, SQRT3_2])
    a1 = np.array([1.5, -SQRT3_2])
    r = np.array([1.0, 0.0])
    circ = 3.0 * n
    height = cells * 2 * SQRT3_2
    pts = {}
    R = 2 * (n + cells) + 4
    for l0 in range(-R, R + 1):
        for l1 in range(-R, R + 1):
            base = l0 * a0 + l1 * a1
            for sgn in (+1, -1):
                p = base + sgn * r / 2
                y = p[1]
                if y < -1e-9 or y >= height - 1e-9:
                    continue
                x = p[0] % circ
                if x > circ - 1e-9:
                    x -= circ
                key = (int(round(y * 1e6)), int(round(x * 1e6)))
                pts[key] = (x, y)
    keys = sorted(pts)
    return np.array([pts[k] for k in keys])


def armchair_nanotube(n, cells, kappa=1.0):
    """(n,n) armchair nanotube of `cells` translation periods, open ends."""
    sites = armchair_nanotube_sites(n, cells)
    circ = 3.0 * n
    dx = sites[:, None, 0] - sites[None, :, 0]
    dx = dx - circ * np.round(dx / circ)
    dy = sites[:, None, 1] - sites[None, :, 1]
    d2 = dx * dx + dy * dy
    K = np.zeros(d2.shape, dtype=np.complex128)
    K[np.abs(d2 - 1.0) < 1e-4] = kappa
    return SpatialLattice(f"ArmchairNanotube({n},{n};{cells})", K, sites)
