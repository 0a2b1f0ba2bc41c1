"""CG and BiCGStab with the reference's semantics (CPU oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* CG: src/NSL/LinAlg/Solver/Impl/CG.tpp:14-113, CG.hpp:31-40 (errSq = eps^2,
  maxIter default 10000).  x0 = b (CG.tpp:97-103) unless x0 is given; one
  alpha/beta for the whole batch because inner_product sums every element
  (src/NSL/LinAlg/inner_product.tpp:31-33); stop on the absolute recursive
  residual |r|^2 <= eps^2; non-convergence returns the current x.
* BiCGStab: src/NSL/LinAlg/Solver/Impl/BiCGStab.tpp:14-150 (same x0 and stop).
* CGpreconditioned: src/NSL/LinAlg/Solver/Impl/CGpreconditioned.tpp:14-118
  (outer CG, preconditioner = inner CG on a second operator).

Each solver returns (x, info) with info = {"iterations", "rsqr", "converged"}.

Parity pins (tests/test_oracle_core.py): random HPD systems against a
direct solve; CGpreconditioned with the identity as preconditioner equals
plain CG iterate for iterate, with A itself converges in one or two outer
steps; the fermion-matrix systems through the FP64 true residual
|b - A x| computed independently; the iteration recurrences are the
reference's (not pinned beyond that: the reference cannot be run here).
"""

import torch


def inner_product(a, b):
    # inner_product.tpp:31-33: sum over all elements of conj(a)*b
    return complex((torch.conj(a) * b).sum())


def cg(A, b, eps=1e-12, maxIter=10000, x0=None):
    b = torch.as_tensor(b, dtype=torch.complex128)
    x = (b if x0 is None else torch.as_tensor(x0, dtype=torch.complex128)).clone()
    errSq = eps * eps
    t = A(x)
    r = b - t
    rsqr_prev = inner_product(r, r).real
    if rsqr_prev <= errSq:
        return x, {"iterations": 0, "rsqr": rsqr_prev, "converged": True}
    p = r.clone()
    for count in range(1, maxIter + 1):
        t = A(p)
        alpha = rsqr_prev / inner_product(p, t)
        x = x + alpha * p
        r = r - alpha * t
        rsqr_curr = inner_product(r, r).real
        if rsqr_curr <= errSq:
            return x, {"iterations": count, "rsqr": rsqr_curr, "converged": True}
        beta = rsqr_curr / rsqr_prev
        p = r + beta * p
        rsqr_prev = rsqr_curr
    return x, {"iterations": maxIter, "rsqr": rsqr_prev, "converged": False}


def bicgstab(A, b, eps=1e-12, maxIter=10000, x0=None):
    # BiCGStab.tpp:14-125
    b = torch.as_tensor(b, dtype=torch.complex128)
    x = (b if x0 is None else torch.as_tensor(x0, dtype=torch.complex128)).clone()
    errSq = eps * eps
    t = A(x)
    r = b - t
    rsqr_prev = inner_product(r, r).real
    if rsqr_prev <= errSq:
        return x, {"iterations": 0, "rsqr": rsqr_prev, "converged": True}
    y = r.clone()
    s = r.clone()
    As = A(s)
    delta = inner_product(y, r)
    phi = inner_product(y, As) / delta
    for count in range(1, maxIter + 1):
        omega = 1.0 / phi
        w = r - omega * As
        Aw = A(w)
        chi = inner_product(Aw, w) / inner_product(Aw, Aw).real
        r = w - chi * Aw
        x = x + omega * s
        x = x + chi * w
        delta_p1 = inner_product(y, r)
        psi = -(omega * delta_p1) / (delta * chi)
        t = s
        s = r - psi * t
        s = s + psi * chi * As
        As = A(s)
        phi = inner_product(y, As) / delta_p1
        delta = delta_p1
        rsqr_curr = inner_product(r, r).real
        if rsqr_curr <= errSq:
            return x, {"iterations": count, "rsqr": rsqr_curr, "converged": True}
        rsqr_prev = rsqr_curr
    return x, {"iterations": maxIter, "rsqr": rsqr_prev, "converged": False}


def cg_preconditioned(A, Aprec, b, eps=1e-12, maxIter=10000, x0=None):
    """CGpreconditioned.tpp:14-100 (solve_), :103-118 (x0 = b unless given):
    outer CG on A whose preconditioner step z = innerCG_(r) is the
    reference's CG on Aprec with the same eps and maxIter, started from r
    (CG.tpp:97-103, CGpreconditioned.hpp:34-41).  Returns (x, info) with
    info["inner_iterations"] summed over the inner solves.  The reference
    computes z before its first stop test; the returned x is the same."""
    b = torch.as_tensor(b, dtype=torch.complex128)
    x = (b if x0 is None else torch.as_tensor(x0, dtype=torch.complex128)).clone()
    errSq = eps * eps
    inner = 0

    def precond(r):
        nonlocal inner
        z, info = cg(Aprec, r, eps=eps, maxIter=maxIter)
        inner += info["iterations"]
        return z

    t = A(x)
    r = b - t
    rsqr_curr = inner_product(r, r).real
    if rsqr_curr <= errSq:
        return x, {"iterations": 0, "rsqr": rsqr_curr, "converged": True, "inner_iterations": 0}
    z = precond(r)
    p = z.clone()
    rz_prev = inner_product(r, z).real
    for count in range(1, maxIter + 1):
        t = A(p)
        alpha = rz_prev / inner_product(p, t)
        x = x + alpha * p
        r = r - alpha * t
        rsqr_curr = inner_product(r, r).real
        if rsqr_curr <= errSq:
            return x, {"iterations": count, "rsqr": rsqr_curr, "converged": True, "inner_iterations": inner}
        z = precond(r)
        rz_curr = inner_product(r, z).real
        beta = rz_curr / rz_prev
        p = z + beta * p
        rz_prev = rz_curr
    return x, {"iterations": maxIter, "rsqr": rsqr_curr, "converged": False, "inner_iterations": inner}
