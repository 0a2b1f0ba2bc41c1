"""Philox-4x32-10 counter-based RNG and derived samplers (numpy, vectorised).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference draws random numbers from torch's global generator
(src/NSL/Tensor/Impl/factory.tpp:38-41 for complex randn with std sqrt(2), i.e.
Re and Im each N(0,1); src/NSL/MarkovChain/HMC.tpp:189-195 for momenta,
HMC.tpp:223 for the Metropolis uniform).  That stream is not reproducible
between a CPU oracle and a GPU engine, so the project generates every input
from Philox-4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as
easy as 1, 2, 3", SC'11) and injects identical numbers into both paths.  The
C++ host library implements the same generator (nslx_rng_* in include/nslx.h).

Pinned by: the Random123 known-answer vectors for philox4x32_10
(tests/test_oracle_core.py) and by equality with the C++ implementation.

Stream layout (shared with the C++ side): block i of stream (seed, offset) is
philox(counter=(i, offset & 0xffffffff, offset >> 32, 0),
       key=(seed & 0xffffffff, seed >> 32)).
Each block yields two doubles u = (a>>5 * 2^26 + b>>6) * 2^-53 from the word
pairs (w0,w1), (w2,w3); normals use Box-Muller on (1-u0, u1):
z0 = r cos(2 pi u1), z1 = r sin(2 pi u1), r = sqrt(-2 log(1-u0)).
"""

import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """Philox-4x32 with 10 rounds.  ctr: (..., 4) uint32, key: (..., 2) uint32."""
    c = [np.asarray(ctr[..., i], dtype=np.uint32) for i in range(4)]
    k0 = np.asarray(key[..., 0], dtype=np.uint32).copy()
    k1 = np.asarray(key[..., 1], dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        for r in range(10):
            p0 = _M0 * c[0].astype(np.uint64)
            p1 = _M1 * c[2].astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & _MASK32).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & _MASK32).astype(np.uint32)
            c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
            if r < 9:
                k0 = k0 + _W0
                k1 = k1 + _W1
    return np.stack(c, axis=-1)


def _blocks(seed, offset, nblocks):
    i = np.arange(nblocks, dtype=np.uint64)
    ctr = np.empty((nblocks, 4), dtype=np.uint32)
    ctr[:, 0] = (i & _MASK32).astype(np.uint32)
    ctr[:, 1] = np.uint32(offset & 0xFFFFFFFF)
    ctr[:, 2] = np.uint32((offset >> 32) & 0xFFFFFFFF)
    ctr[:, 3] = 0
    key = np.empty((nblocks, 2), dtype=np.uint32)
    key[:, 0] = np.uint32(seed & 0xFFFFFFFF)
    key[:, 1] = np.uint32((seed >> 32) & 0xFFFFFFFF)
    return philox4x32_10(ctr, key)


def _u53(a, b):
    return ((a >> np.uint32(5)).astype(np.float64) * 67108864.0
            + (b >> np.uint32(6)).astype(np.float64)) * (1.0 / 9007199254740992.0)


def uniform(seed, offset, n):
    """n doubles in [0,1): two per Philox block."""
    nb = (n + 1) // 2
    w = _blocks(seed, offset, nb)
    u = np.empty((nb, 2))
    u[:, 0] = _u53(w[:, 0], w[:, 1])
    u[:, 1] = _u53(w[:, 2], w[:, 3])
    return u.reshape(-1)[:n]


def normal(seed, offset, n):
    """n standard normal doubles (Box-Muller, two per Philox block)."""
    nb = (n + 1) // 2
    w = _blocks(seed, offset, nb)
    u0 = _u53(w[:, 0], w[:, 1])
    u1 = _u53(w[:, 2], w[:, 3])
    r = np.sqrt(-2.0 * np.log(1.0 - u0))
    z = np.empty((nb, 2))
    z[:, 0] = r * np.cos(2.0 * np.pi * u1)
    z[:, 1] = r * np.sin(2.0 * np.pi * u1)
    return z.reshape(-1)[:n]


def complex_normal(seed, offset, shape):
    """Complex Gaussian with Re, Im each N(0,1) (the reference's complex randn
    convention, factory.tpp:38-41: torch normal_ with std sqrt(2) on a complex
    tensor gives variance 1 per component).  Layout: interleaved re/im."""
    n = int(np.prod(shape))
    z = normal(seed, offset, 2 * n)
    return (z[0::2] + 1j * z[1::2]).reshape(shape)


def real_field(seed, offset, shape, std):
    """Real auxiliary field phi ~ N(0, std), stored as complex with Im = 0."""
    n = int(np.prod(shape))
    return (std * normal(seed, offset, n)).reshape(shape).astype(np.complex128)
