"""CPU oracle for the HubbardExp fermion-matrix / Krylov-solver path.

TEST INFRASTRUCTURE ONLY.  This package restates, op for op on libtorch's CPU
backend in complex128, the arithmetic of the reference NSL library
(Marcel-Rodekamp/NSL) for the fermion-matrix + Krylov-solver hot path.  It is
the checker that the HIP product path is compared against and the timed CPU
baseline.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
``nsl_amd`` never does, and nothing in the product falls back to it.

Modules (each function cites the reference file:line it follows):

* ``rng``        Philox-4x32-10 counter RNG (Salmon et al., SC'11) + Box-Muller;
                 the project's seeded input generator (the reference uses
                 torch's global RNG, which cannot be reproduced across devices).
* ``lattice``    Honeycomb / Ring hopping matrices (src/NSL/Lattice/...), the
                 project-defined (n,n) armchair nanotube, and the memoised
                 E(a) = matrix_exp(a K).
* ``fermion``    HubbardExp: populate, M, Mdagger, MMdagger, MdaggerM, logDetM,
                 gradLogDetM, M_dense, plus a stabilised log det for large beta.
* ``solvers``    CG and BiCGStab with the reference's semantics (x0 = b,
                 batch-global alpha/beta, absolute |r|^2 <= eps^2 stop).
* ``hmc``        Gauge/fermion actions, leapfrog with real force, one HMC
                 trajectory with injected momenta and Metropolis uniform.
* ``correlator`` TwoPointCorrelator::measure(N) arithmetic.

Parity status of every function is recorded in the header of its module.
"""

from . import rng, lattice, fermion, solvers, hmc, correlator  # noqa: F401
