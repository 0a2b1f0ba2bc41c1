"""CPU pins of the double-double residual checker (tests/ddarith.py) used by
the GPU tests of the time-slice solver: it must equal the exact rational
residual (Python fractions) on tiny inputs, resolve a perturbation of x far
below FP64's resolution, and reproduce the FP64 floor that makes it
necessary."""

import numpy as np
import scipy.linalg as sla

from oracle import lattice as olat, rng
from tests import ddarith as dd


def setup(Nt=4, L=3, beta=2.0, seed=5, std=0.7, im=0.3):
    K = olat.honeycomb(L, L).hopping_matrix().numpy().real
    Nx = K.shape[0]
    E = sla.expm(beta / Nt * K)
    phi = rng.real_field(seed, 0, (Nt, Nx), std) + im * 1j * rng.real_field(seed + 1, 0, (Nt, Nx), std)
    return E, np.exp(1j * phi)[None], Nt, Nx


def dense_mmdag(E, P, Nt, Nx):
    I = np.eye(Nt * Nx, dtype=complex).reshape(-1, Nt, Nx)
    s = np.ones(Nt); s[0] = -1
    sp = np.ones(Nt); sp[-1] = -1
    Md = I - np.conj(P) * sp[None, :, None] * (np.roll(I, -1, axis=1) @ E.T)
    M = Md - s[None, :, None] * (np.roll(Md * P, 1, axis=1) @ E.T)
    return M.reshape(Nt * Nx, -1).T


def test_dd_equals_exact_rational():
    E, P, Nt, Nx = setup()
    g = np.random.default_rng(1)
    b = g.standard_normal((1, Nt, Nx)) + 1j * g.standard_normal((1, Nt, Nx))
    A = dense_mmdag(E, P[0], Nt, Nx)
    x = np.linalg.solve(A, b.reshape(-1)).reshape(1, Nt, Nx)
    xl = (g.standard_normal((1, Nt, Nx)) + 1j * g.standard_normal((1, Nt, Nx))) * np.abs(x).max() * 1e-18
    r, _ = dd.mmdag_residual(E, P, b, x, xl)
    re = dd.exact_mmdag_residual(E, P[0], b[0], x[0], xl[0])
    assert np.abs(r[0] - re).max() <= 1e-15 * np.abs(re).max()


def test_dd_equals_exact_rational_mdagm():
    """The M^dag M residual (op "MdaggerM") against exact rational arithmetic,
    at the solution of the dense M^dag M system (a residual of pure
    cancellation), and against the dense matrix built independently:
    M^dag M = (M M^dag with M and M^dag swapped), checked through the dense
    M of dense_mmdag's construction."""
    E, P, Nt, Nx = setup(seed=11)
    g = np.random.default_rng(3)
    b = g.standard_normal((1, Nt, Nx)) + 1j * g.standard_normal((1, Nt, Nx))
    I = np.eye(Nt * Nx, dtype=complex).reshape(-1, Nt, Nx)
    s = np.ones(Nt); s[0] = -1
    M = (I - s[None, :, None] * (np.roll(I * P[0], 1, axis=1) @ E.T)).reshape(Nt * Nx, -1).T
    A = M.conj().T @ M
    x = np.linalg.solve(A, b.reshape(-1)).reshape(1, Nt, Nx)
    xl = (g.standard_normal((1, Nt, Nx)) + 1j * g.standard_normal((1, Nt, Nx))) * np.abs(x).max() * 1e-18
    r, _ = dd.mmdag_residual(E, P, b, x, xl, op="MdaggerM")
    re = dd.exact_mmdag_residual(E, P[0], b[0], x[0], xl[0], op="MdaggerM")
    assert np.abs(r[0] - re).max() <= 1e-15 * np.abs(re).max()
    # at a random x the residual is b - A x to FP64 accuracy (a swapped
    # operator order or a dropped conjugate fails this)
    y = g.standard_normal((1, Nt, Nx)) + 1j * g.standard_normal((1, Nt, Nx))
    r2, _ = dd.mmdag_residual(E, P, b, y, np.zeros_like(y), op="MdaggerM")
    want = b.reshape(-1) - A @ y.reshape(-1)
    assert np.abs(r2.reshape(-1) - want).max() <= 1e-12 * np.abs(want).max()
    r3, _ = dd.mmdag_residual(E, P, b, y, np.zeros_like(y))
    assert np.abs(r3.reshape(-1) - want).max() > 1e-6 * np.abs(want).max()


def test_dd_resolves_below_fp64():
    """x_lo of size 1e-20 |x| changes the double-double residual by the
    amount A x_lo (the FP64 sum x + x_lo would not change at all)."""
    E, P, Nt, Nx = setup()
    g = np.random.default_rng(2)
    b = g.standard_normal((1, Nt, Nx)) + 1j * g.standard_normal((1, Nt, Nx))
    A = dense_mmdag(E, P[0], Nt, Nx)
    x = np.linalg.solve(A, b.reshape(-1)).reshape(1, Nt, Nx)
    d = (g.standard_normal((1, Nt, Nx)) + 1j * g.standard_normal((1, Nt, Nx))) * np.abs(x).max() * 1e-20
    assert np.array_equal(x + d, x)
    r0, _ = dd.mmdag_residual(E, P, b, x, np.zeros_like(x))
    r1, _ = dd.mmdag_residual(E, P, b, x, d)
    want = (A @ d.reshape(-1)).reshape(1, Nt, Nx)
    assert np.abs((r0 - r1) - want).max() <= 1e-6 * np.abs(want).max()


def test_fp64_floor_of_an_illconditioned_system():
    """At large beta with random fields the FP64-rounded solution of MM^dag
    has a residual far above 1e-12 (the floor eps ||A|| |x|), while the
    double-double refinement of it reaches below 1e-14: the reason the
    solver keeps x in double-double."""
    Nt, L, beta, U = 16, 3, 8.0, 3.0
    # seed 26: cond(MM^dag) ~ 1e8 for this lattice and field distribution
    E, P, Nt, Nx = setup(Nt=Nt, L=L, beta=beta, seed=26, std=np.sqrt(beta / Nt * U), im=0.0)
    A = dense_mmdag(E, P[0], Nt, Nx)
    cond = np.linalg.cond(A)
    g = np.random.default_rng(3)
    b = g.standard_normal((1, Nt, Nx)) + 1j * g.standard_normal((1, Nt, Nx))
    lu = sla.lu_factor(A)
    x = sla.lu_solve(lu, b.reshape(-1)).reshape(1, Nt, Nx)
    xl = np.zeros_like(x)
    r, rel0 = dd.mmdag_residual(E, P, b, x, xl)
    for _ in range(3):                       # refinement with exact residuals
        dx = sla.lu_solve(lu, r.reshape(-1)).reshape(1, Nt, Nx)
        s, e = dd.two_sum(x.real, dx.real)
        si, ei = dd.two_sum(x.imag, dx.imag)
        hr, lr = dd.quick_two_sum(s, xl.real + e)
        hi, li = dd.quick_two_sum(si, xl.imag + ei)
        x, xl = hr + 1j * hi, lr + 1j * li
        r, rel1 = dd.mmdag_residual(E, P, b, x, xl)
    assert cond > 1e7
    assert rel0[0] > 1e-12          # the FP64 solution's residual is above 1e-12
    assert rel1[0] < 1e-14
