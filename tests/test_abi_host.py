"""CPU-side checks of the C ABI library (no GPU needed): it loads, exports
every function include/nslx.h declares, and its host-only routines (Philox
RNG, lattice builders) agree with the oracle bit for bit / exactly."""

import ctypes
import os
import re

import numpy as np
import pytest

from oracle import lattice, rng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "nslx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nslx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(libpath):
    L = ctypes.CDLL(libpath)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_abi(libpath):
    from nsl_amd import _lib
    assert set(declared_functions()) == set(_lib.PROTOS)


def test_status_strings(libpath):
    import nsl_amd
    L = nsl_amd._lib.lib()
    assert L.nslx_status_string(0) == b"ok"
    assert L.nslx_status_string(5) == b"not converged"
    assert L.nslx_version() >= 10000


def test_philox_matches_oracle(libpath):
    import nsl_amd
    for seed, off in ((0, 0), (2**40 + 5, 3), (123, 2**33 + 1)):
        got = nsl_amd.rng_philox(seed, off, 64)
        ctr = np.zeros((64, 4), np.uint32)
        ctr[:, 0] = np.arange(64)
        ctr[:, 1] = off & 0xFFFFFFFF
        ctr[:, 2] = off >> 32
        key = np.zeros((64, 2), np.uint32)
        key[:, 0] = seed & 0xFFFFFFFF
        key[:, 1] = seed >> 32
        assert (got == rng.philox4x32_10(ctr, key)).all()
        assert (nsl_amd.rng_uniform(seed, off, 101) == rng.uniform(seed, off, 101)).all()
        assert np.abs(nsl_amd.rng_normal(seed, off, 101) - rng.normal(seed, off, 101)).max() < 1e-14


@pytest.mark.parametrize("L", [(3, 3), (12, 12), (2, 5)])
def test_honeycomb_builder_matches_oracle(libpath, L):
    import nsl_amd
    K = nsl_amd.Lattice.honeycomb(*L).K
    assert (K == lattice.honeycomb(*L).hops.real.numpy()).all()


@pytest.mark.parametrize("n,cells", [(6, 4), (6, 32), (3, 3)])
def test_nanotube_builder_matches_oracle(libpath, n, cells):
    import nsl_amd
    K = nsl_amd.Lattice.armchair_nanotube(n, cells).K
    assert (K == lattice.armchair_nanotube(n, cells).hops.real.numpy()).all()


def test_ring_builder(libpath):
    import nsl_amd
    for n in (2, 3, 7):
        assert (nsl_amd.Lattice.ring(n).K == lattice.ring(n).hops.real.numpy()).all()


def test_no_cpu_fallback_without_library(tmp_path, monkeypatch):
    """The product path must fail loudly when the HIP library is missing."""
    from nsl_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libnslx.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.lib()


def test_one_hip_runtime_in_process(libpath):
    """The binding loads torch before libnslx.so, so the library binds to
    torch's HIP/HSA runtime (same soname) instead of bringing up the system
    one beside it: exactly one of each is mapped, in a fresh interpreter that
    loads the library first."""
    import subprocess
    import sys
    code = ("import nsl_amd._lib as L; L.lib(); import torch; "
            "m = open('/proc/self/maps').read().splitlines(); "
            "print(len({l.split()[-1] for l in m if 'libamdhip64' in l}), "
            "len({l.split()[-1] for l in m if 'libhsa-runtime64' in l}))")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["1", "1"], out.stdout
