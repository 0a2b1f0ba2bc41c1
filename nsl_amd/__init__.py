"""nsl_amd: MI355X-native HubbardExp / HubbardDiag fermion matrices and Krylov solvers.

Host-side mirror of the reference's interfaces for this path (Marcel-Rodekamp/
NSL): ``HubbardExp`` (src/NSL/FermionMatrix/Impl/hubbardExp.hpp:33-233),
``HubbardDiag`` (src/NSL/FermionMatrix/Impl/hubbardDiag.hpp),
``CG`` / ``BiCGStab`` (src/NSL/LinAlg/Solver/Impl/CG.hpp, BiCGStab.hpp) with
``MatrixCombination`` (applicationType.tpp:18), the Hubbard action and HMC
step (Action/, Integrator/, MarkovChain/HMC.tpp).  All arithmetic runs in
libnslx.so (hand-written HIP kernels for gfx950) through the C ABI of
include/nslx.h; torch tensors are used only to hold device memory.  There is
no CPU fallback: without the built library or a GPU every call raises.
"""

import ctypes
import logging

import numpy as np

from . import _lib
from ._lib import (NslxError, SolveReport, HmcReport, F64, F32, BF16X3, TOL_ABS_BATCH,  # noqa: F401
                   TOL_REL_PER_RHS, TOL_REL_PER_RHS_FLOOR)

log = logging.getLogger("nsl_amd")

Particle = _lib.PARTICLE
Hole = _lib.HOLE


class MatrixCombination:
    """NSL::FermionMatrix::MatrixCombination (applicationType.tpp:18)."""
    M = _lib.OP_M
    Mdagger = _lib.OP_MDAGGER
    MMdagger = _lib.OP_MMDAGGER
    MdaggerM = _lib.OP_MDAGGERM


M, Mdagger, MMdagger, MdaggerM = (MatrixCombination.M, MatrixCombination.Mdagger,
                                  MatrixCombination.MMdagger, MatrixCombination.MdaggerM)


def _torch():
    import torch
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(device=None):
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


# ----------------------------------------------------------------- context
class Context:
    """Device context (nslx_ctx)."""

    _default = {}

    def __init__(self, device=0):
        self.device = int(device)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().nslx_ctx_create(self.device, ctypes.byref(h)))
        self._h = h

    @classmethod
    def default(cls, device=0):
        if device not in cls._default:
            cls._default[device] = Context(device)
        return cls._default[device]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.nslx_ctx_destroy(h)
            self._h = None

    # -- hop tile table (nslx_ctx_set_tile / nslx_ctx_get_tile) -------------
    @staticmethod
    def tile_count(precision=F64):
        n = ctypes.c_int()
        _lib.check(_lib.lib().nslx_tile_count(int(precision), ctypes.byref(n)))
        return n.value

    def set_tile(self, precision, nx, nt, nbatch, tile):
        """Pin the hop tile for (precision, nx, nt, nbatch class); tile=None clears."""
        _lib.check(_lib.lib().nslx_ctx_set_tile(self._h, int(precision), int(nx), int(nt), int(nbatch),
                                                -1 if tile is None else int(tile)))

    def get_tile(self, precision, nx, nt, nbatch):
        t = ctypes.c_int()
        _lib.check(_lib.lib().nslx_ctx_get_tile(self._h, int(precision), int(nx), int(nt), int(nbatch),
                                                ctypes.byref(t)))
        return t.value


# ----------------------------------------------------------------- lattices
class Lattice:
    """Spatial lattice: hopping matrix K (numpy float64) built on the host by
    the library (honeycomb.tpp:72-138, ring.tpp:8-42, project nanotube)."""

    def __init__(self, name, K, sites=None):
        self.name = name
        self.K = np.ascontiguousarray(K, dtype=np.float64)
        self.site_coords = sites

    def sites(self):
        return self.K.shape[0]

    def hopping_matrix(self):
        return self.K

    @staticmethod
    def _build(fn, nx, sdim, *args):
        K = np.zeros((nx, nx))
        S = np.zeros((nx, sdim))
        _lib.check(fn(*args, K.ctypes.data_as(_lib.P_double), S.ctypes.data_as(_lib.P_double)))
        return K, S

    @classmethod
    def honeycomb(cls, L0, L1=None, kappa=1.0):
        L1 = L0 if L1 is None else L1
        K, S = cls._build(_lib.lib().nslx_lattice_honeycomb, 2 * L0 * L1, 2, L0, L1, kappa)
        return cls(f"Honeycomb({L0},{L1})", K, S)

    @classmethod
    def ring(cls, n, kappa=1.0):
        K, S = cls._build(_lib.lib().nslx_lattice_ring, n, 3, n, kappa)
        return cls(f"Ring({n})", K, S)

    @classmethod
    def armchair_nanotube(cls, n, cells, kappa=1.0):
        K, S = cls._build(_lib.lib().nslx_lattice_armchair_nanotube, 4 * n * cells, 3, n, cells, kappa)
        return cls(f"ArmchairNanotube({n},{n};{cells})", K, S)


# ------------------------------------------------------------ fermion matrix
class HubbardExp:
    """NSL::FermionMatrix::HubbardExp on the device.

    ctor (lat, Nt, beta, mu) / (species, lat, Nt, beta, mu): hubbardExp.hpp:48-76;
    mu may be complex, beta must be real (a complex beta raises NslxError).
    populate(phi[, species]): hubbardExp.hpp:103-128;
    phi is a complex128 CUDA tensor (Nt, Nx), or (nfield, Nt, Nx) to hold
    several fields sharing the lattice (RHS k of a batch of nbatch uses field
    k // (nbatch // nfield)).
    """

    _create = "nslx_fm_create"

    def __init__(self, lat, Nt, beta=1.0, mu=0.0, species=Particle, ctx=None):
        self.lat = lat
        self.Nt = int(Nt)
        self.Nx = lat.sites()
        beta, mu = complex(beta), complex(mu)
        self.beta = beta.real if beta.imag == 0 else beta
        self.mu = mu.real if mu.imag == 0 else mu
        self._species = int(species)
        self.ctx = ctx or Context.default()
        h = ctypes.c_void_p()
        _lib.check(getattr(_lib.lib(), self._create)(self.ctx._h, self.Nx, self.Nt,
                                                     lat.K.ctypes.data_as(_lib.P_double), beta.real, beta.imag,
                                                     mu.real, mu.imag, self._species, ctypes.byref(h)))
        self._h = h
        self.nfield = 0
        self._phi = None

    @classmethod
    def with_species(cls, species, lat, Nt, beta=1.0, mu=0.0, ctx=None):
        """The (species, lat, Nt, beta, mu) constructor (hubbardExp.hpp:64-76)."""
        return cls(lat, Nt, beta, mu, species, ctx)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.nslx_fm_destroy(h)
            self._h = None

    def species(self):
        return self._species

    def populate(self, phi, species=None):
        torch = _torch()
        phi = phi.to(dtype=torch.complex128).contiguous()
        if phi.dim() == 2:
            nf = 1
        else:
            nf = int(np.prod(phi.shape[:-2]))
        if tuple(phi.shape[-2:]) != (self.Nt, self.Nx):
            raise ValueError(f"phi must end in (Nt, Nx) = ({self.Nt}, {self.Nx}), got {tuple(phi.shape)}")
        if species is not None:
            self._species = int(species)
        _lib.check(_lib.lib().nslx_fm_populate(self._h, _ptr(phi), nf,
                                               -1 if species is None else int(species), _stream()))
        self.nfield = nf
        self._phi = phi
        return self

    def tune(self, precision=F64, nbatch=None, reps=10, cg=False):
        """Autotune the hop tile for this lattice and nbatch (nslx_fm_tune):
        plain M + Mdagger applies, or with cg=True one fused CG iteration on
        MMdagger (the CGA and CGB launches, classic tiles); returns (tile,
        microseconds per pair)."""
        nb = int(nbatch or self.nfield)
        t, us = ctypes.c_int(), ctypes.c_double()
        _lib.check(_lib.lib().nslx_fm_tune(self._h, int(precision), nb, int(reps), 1 if cg else 0, ctypes.byref(t),
                                           ctypes.byref(us), _stream()))
        return t.value, us.value

    # -- applications ------------------------------------------------------
    def apply(self, op, psi, out=None, precision=F64):
        torch = _torch()
        dt = torch.complex128 if precision == F64 else torch.complex64
        if psi.dtype != dt or not psi.is_contiguous():
            psi = psi.to(dtype=dt).contiguous()
        if tuple(psi.shape[-2:]) != (self.Nt, self.Nx):
            raise ValueError(f"psi must end in (Nt, Nx) = ({self.Nt}, {self.Nx})")
        nb = int(np.prod(psi.shape[:-2])) if psi.dim() > 2 else 1
        if out is None:
            out = torch.empty_like(psi)
        _lib.check(_lib.lib().nslx_fm_apply(self._h, int(op), _ptr(psi), _ptr(out), nb, int(precision),
                                            _stream()))
        return out

    def M(self, psi, precision=F64):
        return self.apply(MatrixCombination.M, psi, precision=precision)

    def Mdagger(self, psi, precision=F64):
        return self.apply(MatrixCombination.Mdagger, psi, precision=precision)

    def MMdagger(self, psi, precision=F64):
        return self.apply(MatrixCombination.MMdagger, psi, precision=precision)

    def MdaggerM(self, psi, precision=F64):
        return self.apply(MatrixCombination.MdaggerM, psi, precision=precision)

    # -- determinant and force --------------------------------------------
    def logDetM(self, field=0, nchunks=0):
        out = (ctypes.c_double * 2)()
        _lib.check(_lib.lib().nslx_fm_logdet(self._h, int(field), int(nchunks),
                                             ctypes.cast(out, _lib.P_double), _stream()))
        return complex(out[0], out[1])

    def gradLogDetM(self, field=0):
        torch = _torch()
        out = torch.empty((self.Nt, self.Nx), dtype=torch.complex128, device="cuda")
        _lib.check(_lib.lib().nslx_fm_grad_logdet(self._h, int(field), _ptr(out), _stream()))
        return out

    # -- time-slice direct solve (no counterpart in the reference) ----------
    def solve_direct(self, op, y, out=None, precision=F64, nchunks=0):
        """M^{-1} y (op M), M^{-dag} y (Mdagger), (M^dag M)^{-1} y (MdaggerM) or (M M^dag)^{-1} y
        (MMdagger) by the time-slice solve (nslx_fm_solve_direct); y is
        complex128 for F64, complex64 for F32."""
        torch = _torch()
        dt = torch.complex128 if precision == F64 else torch.complex64
        if y.dtype != dt or not y.is_contiguous():
            y = y.to(dtype=dt).contiguous()
        nb = int(np.prod(y.shape[:-2])) if y.dim() > 2 else 1
        if out is None:
            out = torch.empty_like(y)
        _lib.check(_lib.lib().nslx_fm_solve_direct(self._h, int(op), _ptr(y), _ptr(out), nb, int(precision),
                                                   int(nchunks), _stream()))
        return out

    def export(self):
        """(E, P): the FP64 exp(s delta K) (Nx, Nx) and phase table
        (nfield, Nt, Nx) the engine applies (nslx_fm_export)."""
        E = np.empty((self.Nx, self.Nx), dtype=np.float64)
        P = np.empty((self.nfield, self.Nt, self.Nx), dtype=np.complex128)
        _lib.check(_lib.lib().nslx_fm_export(self._h, E.ctypes.data_as(_lib.P_double), P.ctypes.data, _stream()))
        return E, P


class HubbardDiag(HubbardExp):
    """NSL::FermionMatrix::HubbardDiag on the device (hubbardDiag.hpp,
    hubbardDiag.tpp): (M psi)_t = psi_t - s delta K psi_t - B_t P_{t-1} psi_{t-1}.
    Same constructors, populate, M/Mdagger/MMdagger/MdaggerM, logDetM and
    gradLogDetM as HubbardExp; solvers, the correlator and the HMC entries
    take it too, and batch_solve and Server accept it (MMdagger systems)."""

    _create = "nslx_fm_create_diag"


# ------------------------------------------------------------------ solvers
class _Solver:
    method = _lib.CG_METHOD

    def __init__(self, fm, matrixCombination=MatrixCombination.M, eps=1e-12, maxIter=10000, precision=F64,
                 tol_kind=None, check_residual=False, double_double=False):
        self.fm = fm
        self.op = int(matrixCombination)
        self.eps = float(eps)
        self.maxIter = int(maxIter)
        self.precision = int(precision)
        if tol_kind is None:
            tol_kind = TOL_ABS_BATCH if precision == F64 else TOL_REL_PER_RHS
        self.tol_kind = int(tol_kind)
        self.check = bool(check_residual)
        # per-RHS relative tolerance with the solution in double-double
        # (nslx_solve_dd, CG only); .x_lo holds the low part after a call
        self.dd = bool(double_double)
        self.x_lo = None
        self.report = None
        self.converged = None

    def __call__(self, b, x0=None, out=None):
        torch = _torch()
        b = b.to(dtype=torch.complex128).contiguous()
        nb = int(np.prod(b.shape[:-2])) if b.dim() > 2 else 1
        x = torch.empty_like(b) if out is None else out
        if x0 is not None:
            x0 = x0.to(dtype=torch.complex128).contiguous()
        rep = SolveReport()
        if self.dd:
            if self.method != _lib.CG_METHOD:
                raise ValueError("double_double needs CG")
            if self.x_lo is None or self.x_lo.shape != b.shape:
                self.x_lo = torch.empty_like(b)
            st = _lib.lib().nslx_solve_dd(self.fm._h, self.op, _ptr(b), _ptr(x), _ptr(self.x_lo),
                                          _ptr(x0) if x0 is not None else None, nb, self.eps, self.maxIter,
                                          self.precision, ctypes.byref(rep), _stream())
        else:
            st = _lib.lib().nslx_solve(self.fm._h, self.op, self.method, _ptr(b), _ptr(x),
                                       _ptr(x0) if x0 is not None else None, nb, self.eps, self.tol_kind,
                                       self.maxIter, self.precision, int(self.check), ctypes.byref(rep), _stream())
        _lib.check(st, allow=(_lib.OK, _lib.NOT_CONVERGED))
        self.report = rep.as_dict()
        self.converged = st == _lib.OK
        if not self.converged:
            # the reference logs and returns the current x (CG.tpp:90-93)
            log.error("%s did not converge within %d iterations", type(self).__name__, self.maxIter)
        return x


class CG(_Solver):
    """NSL::LinAlg::CG (CG.hpp:31-186, CG.tpp:14-113): x0 = b by default,
    batch-global alpha/beta and |r|^2 <= eps^2 stop (TOL_ABS_BATCH, the
    default for FP64); TOL_REL_PER_RHS solves each RHS to |r|/|b| <= eps.
    precision=F32/BF16X3 runs the reliable-update solver (low-precision
    recurrence, FP64 true residual)."""
    method = _lib.CG_METHOD


class BiCGStab(_Solver):
    """NSL::LinAlg::BiCGStab (BiCGStab.hpp:31-200, BiCGStab.tpp:14-150)."""
    method = _lib.BICGSTAB_METHOD


class CGpreconditioned:
    """NSL::LinAlg::CGpreconditioned (CGpreconditioned.hpp:83-91, .tpp:14-118):
    CG on matrixCombination of M, each preconditioner step z = P^-1 r an
    inner CG on the same combination of Mprec started from r, with the same
    eps and maxIter.  The reference's example preconditions with the
    non-interacting matrix (phi = 0).  tol_kind TOL_ABS_BATCH (default) is
    the reference's batch-global criterion at both levels; inner_precision
    F32/BF16X3 runs the inner solves in low precision (TOL_REL_PER_RHS,
    MMdagger/MdaggerM only).  Mprec may be M itself."""

    def __init__(self, M, Mprec, matrixCombination=MatrixCombination.M, eps=1e-12, maxIter=10000,
                 tol_kind=TOL_ABS_BATCH, inner_precision=F64):
        self.fm, self.fp = M, Mprec
        self.op = int(matrixCombination)
        self.eps = float(eps)
        self.maxIter = int(maxIter)
        self.tol_kind = int(tol_kind)
        self.inner_precision = int(inner_precision)
        self.report = None
        self.converged = None

    def __call__(self, b, x0=None, out=None):
        torch = _torch()
        b = b.to(dtype=torch.complex128).contiguous()
        nb = int(np.prod(b.shape[:-2])) if b.dim() > 2 else 1
        x = torch.empty_like(b) if out is None else out
        if x0 is not None:
            x0 = x0.to(dtype=torch.complex128).contiguous()
        rep = SolveReport()
        st = _lib.lib().nslx_solve_precond(self.fm._h, self.fp._h, self.op, _ptr(b), _ptr(x),
                                           _ptr(x0) if x0 is not None else None, nb, self.eps, self.tol_kind,
                                           self.maxIter, self.inner_precision, ctypes.byref(rep), _stream())
        _lib.check(st, allow=(_lib.OK, _lib.NOT_CONVERGED))
        self.report = rep.as_dict()
        self.converged = st == _lib.OK
        if not self.converged:
            log.error("CGpreconditioned did not converge within %d iterations", self.maxIter)
        return x


class CGTimeSlice:
    """CG on MMdagger (or, with matrixCombination=MdaggerM, on M^dag M)
    preconditioned by the time-slice direct solve, in a double-double defect
    correction to |b - A x| <= eps |b| per RHS (nslx_solve_tss; no
    counterpart in the reference, whose CG on MMdagger is CG.tpp:14-113).  precision: the preconditioner's (F32: factorisation
    and applies on FP32 matrix cores, the CG recurrence in FP64; F64).
    After a call, .x_lo holds the low part of the double-double solution
    (x + x_lo) and .report the solver report.  nchunks 0: the chunk count
    nslx_solve_tss picks for the precision (wider chunks in F64)."""

    def __init__(self, fm, eps=1e-12, maxIter=10000, precision=F32, nchunks=0,
                 matrixCombination=MatrixCombination.MMdagger):
        self.fm = fm
        self.op = int(matrixCombination)
        self.eps = float(eps)
        self.maxIter = int(maxIter)
        self.precision = int(precision)
        self.nchunks = int(nchunks)
        self.report = None
        self.converged = None
        self.x_lo = None

    def __call__(self, b, x0=None, out=None, x_lo=None):
        torch = _torch()
        b = b.to(dtype=torch.complex128).contiguous()
        nb = int(np.prod(b.shape[:-2])) if b.dim() > 2 else 1
        x = torch.empty_like(b) if out is None else out
        lo = torch.empty_like(b) if x_lo is None else x_lo
        if x0 is not None:
            x0 = x0.to(dtype=torch.complex128).contiguous()
        rep = SolveReport()
        st = _lib.lib().nslx_solve_tss(self.fm._h, self.op, _ptr(b), _ptr(x), _ptr(lo),
                                       _ptr(x0) if x0 is not None else None, nb, self.eps, self.maxIter,
                                       self.precision, self.nchunks, ctypes.byref(rep), _stream())
        _lib.check(st, allow=(_lib.OK, _lib.NOT_CONVERGED))
        self.report = rep.as_dict()
        self.converged = st == _lib.OK
        self.x_lo = lo
        if not self.converged:
            log.error("CGTimeSlice did not converge within %d iterations", self.maxIter)
        return x


def dd_residual(fm, b, x, x_lo=None, op=MatrixCombination.MMdagger):
    """Per-RHS |b - A (x + x_lo)| / |b| with the residual in double-double
    (nslx_dd_residual); numpy array of nbatch values."""
    torch = _torch()
    b = b.to(dtype=torch.complex128).contiguous()
    x = x.to(dtype=torch.complex128).contiguous()
    nb = int(np.prod(b.shape[:-2])) if b.dim() > 2 else 1
    rel = np.empty(nb, dtype=np.float64)
    _lib.check(_lib.lib().nslx_dd_residual(fm._h, int(op), _ptr(b), _ptr(x),
                                           _ptr(x_lo.contiguous()) if x_lo is not None else None, None,
                                           rel.ctypes.data_as(_lib.P_double), nb, _stream()))
    return rel


def batch_solve(problems, tol=1e-12, precision=F64, maxIter=100000, use_graph=True, ctx=None):
    """Solve MMdagger x = b (or MdaggerM x = b) for many fermion matrices of
    different Nx, Nt and fields at once (nslx_batch_solve): problems is a
    list of (fm, b) or (fm, b, op) with b (nbatch, Nt, Nx) complex128 on the
    device and op MMdagger (default) or MdaggerM.  Returns (list of x,
    report dict); raises on errors, reports non-convergence in the dict."""
    torch = _torch()
    if not problems:
        raise ValueError("batch_solve: no problems")
    ctx = ctx or problems[0][0].ctx
    items = (_lib.BatchItem * len(problems))()
    keep, xs = [], []
    for i, pr in enumerate(problems):
        fm, b = pr[0], pr[1]
        b = b.to(dtype=torch.complex128).contiguous()
        nb = int(np.prod(b.shape[:-2])) if b.dim() > 2 else 1
        x = torch.empty_like(b)
        keep.append(b)
        xs.append(x)
        items[i].fm = fm._h
        items[i].b = _ptr(b)
        items[i].x = _ptr(x)
        items[i].nbatch = nb
        items[i].op = int(pr[2]) if len(pr) > 2 else int(MatrixCombination.MMdagger)
    rep = _lib.BatchReport()
    st = _lib.lib().nslx_batch_solve(ctx._h, len(problems), items, float(tol), int(precision), int(maxIter),
                                     int(bool(use_graph)), ctypes.byref(rep), _stream())
    _lib.check(st, allow=(_lib.OK, _lib.NOT_CONVERGED))
    d = rep.as_dict()
    d["converged"] = st == _lib.OK
    return xs, d


def batch_solve_tss(problems, tol=1e-12, precision=F64, maxIter=2000, nstreams=8, ctx=None, return_lo=False):
    """Solve MMdagger x = b (or MdaggerM x = b) for many HubbardExp matrices
    of different Nx, Nt, beta and fields, each by CG preconditioned with its
    own time-slice solve and double-double correction (nslx_batch_solve_tss;
    no counterpart in the reference), the items spread over nstreams worker
    streams.  problems: a list of (fm, b) or (fm, b, op), every fm at most
    once, b (nbatch, Nt, Nx) complex128 on the device, op MMdagger (default)
    or MdaggerM.  precision: the preconditioner's (F64 or F32).  Returns (list
    of x, report dict), or (list of x, list of x_lo, report) with return_lo
    (x + x_lo is the double-double solution)."""
    torch = _torch()
    if not problems:
        raise ValueError("batch_solve_tss: no problems")
    ctx = ctx or problems[0][0].ctx
    items = (_lib.BatchItem * len(problems))()
    keep, xs = [], []
    for i, pr in enumerate(problems):
        fm, b = pr[0], pr[1]
        b = b.to(dtype=torch.complex128).contiguous()
        nb = int(np.prod(b.shape[:-2])) if b.dim() > 2 else 1
        x = torch.empty_like(b)
        keep.append(b)
        xs.append(x)
        items[i].fm = fm._h
        items[i].b = _ptr(b)
        items[i].x = _ptr(x)
        items[i].nbatch = nb
        items[i].op = int(pr[2]) if len(pr) > 2 else int(MatrixCombination.MMdagger)
    los, lo_arr = None, None
    if return_lo:
        los = [torch.empty_like(x) for x in xs]
        lo_arr = (ctypes.c_void_p * len(problems))(*[_ptr(t) for t in los])
    rep = _lib.BatchReport()
    st = _lib.lib().nslx_batch_solve_tss(ctx._h, len(problems), items, lo_arr, float(tol), int(precision),
                                         int(maxIter), int(nstreams), ctypes.byref(rep))
    _lib.check(st, allow=(_lib.OK, _lib.NOT_CONVERGED))
    d = rep.as_dict()
    d["converged"] = st == _lib.OK
    return (xs, los, d) if return_lo else (xs, d)


class Server:
    """Correlator server with continuous batching (nslx_server_*): submit
    requests (lattice, Nt, beta, mu, species, phi, time sources, tol,
    precision), wait for their two-point correlators; admission by device
    footprint, requests share every Krylov iteration with what is running."""

    def __init__(self, ctx=None, max_rhs=0, hbm_budget=0.0, use_graph=True, iters_per_replay=32, precond_workers=0):
        """precond_workers > 0: serve with the time-slice-preconditioned
        solver on that many worker streams (a request's precision is then
        the preconditioner's, F64 or F32) instead of the Krylov pool."""
        self.ctx = ctx or Context.default()
        cfg = _lib.ServerConfig(int(max_rhs), float(hbm_budget), int(bool(use_graph)), int(iters_per_replay),
                                int(precond_workers))
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().nslx_server_create(self.ctx._h, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._shapes = {}

    def submit(self, lat, Nt, beta, phi, n_tsrc=1, mu=0.0, species=Particle, tol=1e-10, precision=F64,
               maxIter=100000, fermion_matrix=None):
        """Queue TwoPointCorrelator::measure(n_tsrc) on the fermion matrix
        fermion_matrix (HubbardExp by default, or HubbardDiag) of lat, Nt,
        beta, mu, species populated with phi; returns the request id."""
        phi = np.ascontiguousarray(np.asarray(phi.cpu() if hasattr(phi, "cpu") else phi, dtype=np.complex128))
        Nx = lat.sites()
        if phi.shape != (Nt, Nx):
            raise ValueError(f"phi must be (Nt, Nx) = ({Nt}, {Nx})")
        fmcls = fermion_matrix or HubbardExp
        if fmcls not in (HubbardExp, HubbardDiag):
            raise ValueError("fermion_matrix must be HubbardExp or HubbardDiag")
        disc = _lib.DISC_DIAG if fmcls is HubbardDiag else _lib.DISC_EXP
        req = _lib.Request(Nx, int(Nt), lat.K.ctypes.data_as(_lib.P_double), float(beta), float(mu), int(species),
                           phi.ctypes.data, int(n_tsrc), float(tol), int(precision), int(maxIter), disc)
        rid = ctypes.c_int64()
        _lib.check(_lib.lib().nslx_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid)))
        self._shapes[rid.value] = (int(Nt), Nx)
        return rid.value

    def wait(self, rid, timeout_ms=-1.0):
        """(corr (Nt, Nx, Nx) complex128 on the device, result dict)."""
        torch = _torch()
        Nt, Nx = self._shapes.pop(rid)
        corr = torch.empty((Nt, Nx, Nx), dtype=torch.complex128, device="cuda")
        res = _lib.Result()
        st = _lib.lib().nslx_server_wait(self._h, int(rid), float(timeout_ms), _ptr(corr), ctypes.byref(res))
        _lib.check(st, allow=(_lib.OK, _lib.NOT_CONVERGED))
        return corr, res.as_dict()

    def stats(self):
        st = _lib.ServerStats()
        _lib.check(_lib.lib().nslx_server_stats_get(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.nslx_server_destroy(h)
            self._h = None

    def __del__(self):
        self.close()


# ------------------------------------------------------------- measurements
class TwoPointCorrelator:
    """NSL::Measure::Hubbard::TwoPointCorrelator::measure(N)
    (twoPointCorrelationFunction.tpp:116-162) on a populated HubbardExp with
    one field.  measure() returns corr (Nt, Nx, Nx) complex128 on the device,
    corr[t, y, x].  The default (TOL_ABS_BATCH, eps 1e-12) is the reference's
    CG per time source; TOL_REL_PER_RHS batches time sources and allows
    precision=F32/BF16X3."""

    def __init__(self, fm, eps=1e-12, maxIter=10000, tol_kind=TOL_ABS_BATCH, precision=F64, max_rhs=0):
        self.fm = fm
        self.eps = float(eps)
        self.maxIter = int(maxIter)
        self.tol_kind = int(tol_kind)
        self.precision = int(precision)
        self.max_rhs = int(max_rhs)
        self.report = None
        self.converged = None

    def measure(self, NumberTimeSources=None, out=None):
        torch = _torch()
        n = int(NumberTimeSources or self.fm.Nt)
        corr = out if out is not None else torch.empty((self.fm.Nt, self.fm.Nx, self.fm.Nx),
                                                       dtype=torch.complex128, device="cuda")
        rep = SolveReport()
        st = _lib.lib().nslx_correlator_two_point(self.fm._h, n, self.eps, self.maxIter, self.tol_kind,
                                                  self.precision, self.max_rhs, _ptr(corr), ctypes.byref(rep),
                                                  _stream())
        _lib.check(st, allow=(_lib.OK, _lib.NOT_CONVERGED))
        self.report = rep.as_dict()
        self.converged = st == _lib.OK
        if not self.converged:
            log.error("TwoPointCorrelator: CG did not converge within %d iterations", self.maxIter)
        return corr

    def measureK(self, wall, NumberTimeSources=None):
        """measureK(k, N) (twoPointCorrelationFunction.tpp:169-233) for wall =
        wallSources[k], a complex (uDim, Nx) tensor: returns (corrK, corrKblock)
        with shapes (uDim, Nt, Nx) and (Nt, uDim, uDim)."""
        torch = _torch()
        wall = torch.as_tensor(wall).to(device="cuda", dtype=torch.complex128).contiguous()
        nw = wall.shape[0]
        n = int(NumberTimeSources or self.fm.Nt)
        corrK = torch.empty((nw, self.fm.Nt, self.fm.Nx), dtype=torch.complex128, device="cuda")
        block = torch.empty((self.fm.Nt, nw, nw), dtype=torch.complex128, device="cuda")
        rep = SolveReport()
        st = _lib.lib().nslx_correlator_two_point_wall(self.fm._h, _ptr(wall), nw, n, self.eps, self.maxIter,
                                                       self.tol_kind, self.precision, _ptr(corrK), _ptr(block),
                                                       ctypes.byref(rep), _stream())
        _lib.check(st, allow=(_lib.OK, _lib.NOT_CONVERGED))
        self.report = rep.as_dict()
        self.converged = st == _lib.OK
        return corrK, block


TWO_BODY_CHANNELS = ("I1S1Iz1Sz1", "I1S1Iz1Sz0", "I1S1Iz1Szn1", "I1S1Iz0Sz1", "I1S1Iz0Szn1", "I1S1Izn1Sz1",
                     "I1S1Izn1Sz0", "I1S1Izn1Szn1", "I0S1Iz0Sz1", "I0S1Iz0Szn1", "I1S0Iz1Sz0", "I1S0Izn1Sz0")


def two_body_quads(momenta):
    """Momentum quadruples (w, x, z, y) with k_w + k_x = k_z + k_y modulo the
    honeycomb reciprocal lattice, in the reference's loop order
    (twoBodyCorrelationFunction.tpp:139-151)."""
    m = np.ascontiguousarray(np.asarray(momenta, dtype=np.float64))
    kdim = m.shape[0]
    n = ctypes.c_int64(0)
    L = _lib.lib()
    _lib.check(L.nslx_two_body_quads(m.ctypes.data_as(_lib.P_double), kdim, ctypes.byref(n), None))
    q = np.zeros((max(n.value, 1), 4), dtype=np.int32)
    _lib.check(L.nslx_two_body_quads(m.ctypes.data_as(_lib.P_double), kdim, ctypes.byref(n),
                                     q.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return q[:n.value]


class TwoBodyCorrelator:
    """NSL::Measure::Hubbard::TwoBodyCorrelator::measure(N)
    (twoBodyCorrelationFunction.tpp:244-300, channels :481-1000) on a
    populated fermion matrix: wallSources (kDim, bDim, Nx) and momenta
    (kDim, 2).  measure(N) returns {channel name: {(w, x, z, y): tensor
    (Nt, bDim^2, bDim^2)}} like the reference's maps, from one device array
    (12, nquad, Nt, bDim^2, bDim^2) kept as `.out`.  CG to eps with the
    reference's batch-global criterion by default (cg_, cgDag_ built with
    eps = 1e-12)."""

    def __init__(self, fm, wallSources, momenta, eps=1e-12, maxIter=10000, tol_kind=TOL_ABS_BATCH, precision=F64):
        torch = _torch()
        self.fm = fm
        self.wall = torch.as_tensor(wallSources).to(device="cuda", dtype=torch.complex128).contiguous()
        self.kdim, self.bdim = int(self.wall.shape[0]), int(self.wall.shape[1])
        self.quads = two_body_quads(momenta)
        self.eps, self.maxIter = float(eps), int(maxIter)
        self.tol_kind, self.precision = int(tol_kind), int(precision)
        self.report = None
        self.converged = None
        self.out = None

    def measure(self, NumberTimeSources=None):
        torch = _torch()
        n = int(NumberTimeSources or self.fm.Nt)
        nq = len(self.quads)
        b2 = self.bdim * self.bdim
        out = torch.empty((12, nq, self.fm.Nt, b2, b2), dtype=torch.complex128, device="cuda")
        q = np.ascontiguousarray(self.quads, dtype=np.int32)
        rep = SolveReport()
        st = _lib.lib().nslx_correlator_two_body(self.fm._h, _ptr(self.wall), self.kdim, self.bdim, nq,
                                                 q.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n, self.eps,
                                                 self.maxIter, self.tol_kind, self.precision, _ptr(out),
                                                 ctypes.byref(rep), _stream())
        _lib.check(st, allow=(_lib.OK, _lib.NOT_CONVERGED))
        self.report = rep.as_dict()
        self.converged = st == _lib.OK
        self.out = out
        return {name: {tuple(int(v) for v in self.quads[k]): out[c, k] for k in range(nq)}
                for c, name in enumerate(TWO_BODY_CHANNELS)}


# ------------------------------------------------------------------ HMC
class HubbardAction:
    """Gauge + fermion Hubbard action and the HMC step on one real field
    (hubbardGaugeAction.tpp, hubbardFermiAction.tpp, leapfrogRealForce.tpp,
    HMC.tpp:185-241)."""

    def __init__(self, lat, Nt, beta, U, mu=0.0, ctx=None, nchunks=0, fermion_matrix=None):
        # HubbardFermionAction<..., FermionMatrixType> (hubbardFermiAction.tpp:23-60):
        # HubbardExp by default, HubbardDiag if passed
        self.fm = (fermion_matrix or HubbardExp)(lat, Nt, beta, mu, Particle, ctx)
        self.U = float(U)
        self.nchunks = int(nchunks)

    def eval(self, phi):
        phi = phi.to(dtype=_torch().complex128).contiguous()
        out = (ctypes.c_double * 2)()
        _lib.check(_lib.lib().nslx_hmc_action(self.fm._h, _ptr(phi), self.U, self.nchunks,
                                              ctypes.cast(out, _lib.P_double), _stream()))
        return complex(out[0], out[1])

    def force(self, phi):
        torch = _torch()
        phi = phi.to(dtype=torch.complex128).contiguous()
        out = torch.empty_like(phi)
        _lib.check(_lib.lib().nslx_hmc_force(self.fm._h, _ptr(phi), self.U, _ptr(out), _stream()))
        return out

    def trajectory(self, phi, momentum, S0, traj_length, nmd, uniform):
        """One trajectory; phi is updated in place on accept.  Returns
        (report dict, new S0)."""
        S = (ctypes.c_double * 2)(S0.real, S0.imag)
        rep = HmcReport()
        mom = momentum.to(dtype=_torch().complex128).contiguous()
        _lib.check(_lib.lib().nslx_hmc_trajectory(self.fm._h, _ptr(phi), _ptr(mom), ctypes.cast(S, _lib.P_double),
                                                  self.U, float(traj_length), int(nmd), float(uniform),
                                                  self.nchunks, ctypes.byref(rep), _stream()))
        return rep.as_dict(), complex(S[0], S[1])

    def chain(self, phi, ntraj, traj_length, nmd, seed):
        reps = (HmcReport * int(ntraj))()
        _lib.check(_lib.lib().nslx_hmc_chain(self.fm._h, _ptr(phi), self.U, float(traj_length), int(nmd),
                                             int(ntraj), int(seed), self.nchunks, reps, _stream()))
        return [r.as_dict() for r in reps]

    def generate(self, phi, nconf, traj_length, nmd, seed, save_frequency=1, markov_time=0, all_states=True,
                 keep_configs=True):
        """HMC::generate<AllStates | LastState>(state, Nconf, saveFrequency)
        (HMC.tpp:57-155, MarkovState markovState.tpp) from the state (phi,
        markov_time), without the file I/O.  phi (device, in/out) ends as the
        last state's configuration.  Returns (states, configs): a list of
        MarkovState dicts and, if keep_configs, a device tensor (n, Nt, Nx)
        of the saved configurations.  Resuming from any saved state (its
        configuration and markov_time) continues the chain exactly."""
        torch = _torch()
        nstates = int(nconf) if all_states else 1
        reps = (_lib.MarkovState * nstates)()
        cfg = (torch.empty((nstates, self.fm.Nt, self.fm.Nx), dtype=torch.complex128, device="cuda")
               if keep_configs else None)
        _lib.check(_lib.lib().nslx_hmc_generate(self.fm._h, _ptr(phi), int(markov_time), self.U, float(traj_length),
                                                int(nmd), int(nconf), int(save_frequency),
                                                _lib.CHAIN_ALL if all_states else _lib.CHAIN_LAST, int(seed),
                                                self.nchunks, _ptr(cfg) if cfg is not None else None, reps,
                                                _stream()))
        return [r.as_dict() for r in reps], cfg


# ------------------------------------------------------------------ RNG
def rng_normal(seed, offset, n):
    out = np.empty(int(n))
    _lib.check(_lib.lib().nslx_rng_normal(int(seed), int(offset), int(n), out.ctypes.data_as(_lib.P_double)))
    return out


def rng_uniform(seed, offset, n):
    out = np.empty(int(n))
    _lib.check(_lib.lib().nslx_rng_uniform(int(seed), int(offset), int(n), out.ctypes.data_as(_lib.P_double)))
    return out


def real_field(seed, shape, std, offset=0):
    """Real auxiliary field phi ~ N(0, std) (complex128 with Im = 0), numpy."""
    n = int(np.prod(shape))
    return (std * rng_normal(seed, offset, n)).reshape(shape).astype(np.complex128)


def complex_normal(seed, shape, offset=0):
    """Complex Gaussian with Re, Im each N(0,1) (the reference's complex randn,
    src/NSL/Tensor/Impl/factory.tpp:38-41), numpy complex128."""
    n = int(np.prod(shape))
    z = rng_normal(seed, offset, 2 * n)
    return (z[0::2] + 1j * z[1::2]).reshape(shape)


def rng_philox(seed, offset, nblocks):
    out = np.empty((int(nblocks), 4), dtype=np.uint32)
    _lib.check(_lib.lib().nslx_rng_philox4x32(int(seed), int(offset), int(nblocks),
                                              out.ctypes.data_as(_lib.P_uint32)))
    return out
