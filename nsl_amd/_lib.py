"""ctypes binding of libnslx.so (the C ABI declared in include/nslx.h).

Loading fails loudly: there is no CPU fallback anywhere in nsl_amd.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnslx.so")

c_int, c_int64, c_uint64, c_double, c_void_p = (ctypes.c_int, ctypes.c_int64, ctypes.c_uint64,
                                                ctypes.c_double, ctypes.c_void_p)
P_double = ctypes.POINTER(ctypes.c_double)
P_uint32 = ctypes.POINTER(ctypes.c_uint32)

OK, E_ARG, E_DEVICE, E_OOM, E_STATE, NOT_CONVERGED, E_INTERNAL = range(7)
OP_M, OP_MDAGGER, OP_MMDAGGER, OP_MDAGGERM = range(4)
PARTICLE, HOLE = 0, 1
F64, F32, BF16X3 = 0, 1, 2
CG_METHOD, BICGSTAB_METHOD = 0, 1
TOL_ABS_BATCH, TOL_REL_PER_RHS, TOL_REL_PER_RHS_FLOOR = 0, 1, 2


class SolveReport(ctypes.Structure):
    _fields_ = [("iterations", c_int64), ("iterations_total", c_int64), ("outer_steps", c_int64),
                ("converged_rhs", c_int64), ("rsqr", c_double), ("max_rel_residual", c_double),
                ("seconds", c_double), ("floor_stops", c_int64), ("max_floor_rel", c_double),
                ("rhs_iters_low", c_int64), ("rhs_iters_f64", c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchItem(ctypes.Structure):
    _fields_ = [("fm", c_void_p), ("b", c_void_p), ("x", c_void_p), ("nbatch", c_int64), ("op", c_int)]


class BatchReport(ctypes.Structure):
    _fields_ = [("iterations", c_int64), ("replays", c_int64), ("restarts", c_int64), ("escalations", c_int64),
                ("converged_items", c_int64), ("floor_stops", c_int64), ("max_rel_residual", c_double),
                ("max_floor_rel", c_double), ("seconds", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_rhs", c_int64), ("hbm_budget", c_double), ("use_graph", c_int), ("iters_per_replay", c_int),
                ("precond_workers", c_int)]


class Request(ctypes.Structure):
    _fields_ = [("nx", c_int), ("nt", c_int), ("K", P_double), ("beta", c_double), ("mu", c_double),
                ("species", c_int), ("phi", c_void_p), ("n_tsrc", c_int64), ("tol", c_double),
                ("precision", c_int), ("maxit", c_int64), ("discretisation", c_int)]


class Result(ctypes.Structure):
    _fields_ = [("status", c_int), ("iterations", c_int64), ("escalated", c_int), ("max_rel_residual", c_double),
                ("queue_ms", c_double), ("latency_ms", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", c_int64), ("completed", c_int64), ("rhs_solved", c_int64), ("admitted_peak", c_int64),
                ("p50_ms", c_double), ("p99_ms", c_double), ("max_ms", c_double), ("busy_seconds", c_double),
                ("bytes_peak", c_double), ("iterations", c_int64), ("replays", c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class HmcReport(ctypes.Structure):
    _fields_ = [("dH_re", c_double), ("dH_im", c_double), ("acc_prob", c_double), ("accepted", c_int),
                ("S_re", c_double), ("S_im", c_double), ("seconds", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MarkovState(ctypes.Structure):
    _fields_ = [("S_re", c_double), ("S_im", c_double), ("acc_prob", c_double), ("markov_time", c_int64),
                ("accepted", c_int), ("dH_re", c_double), ("dH_im", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


CHAIN_ALL, CHAIN_LAST = 0, 1
DISC_EXP, DISC_DIAG = 0, 1

# name -> (restype, argtypes)
PROTOS = {
    "nslx_version": (c_int, []),
    "nslx_status_string": (ctypes.c_char_p, [c_int]),
    "nslx_last_error": (ctypes.c_char_p, []),
    "nslx_ctx_create": (c_int, [c_int, ctypes.POINTER(c_void_p)]),
    "nslx_ctx_destroy": (c_int, [c_void_p]),
    "nslx_rng_philox4x32": (c_int, [c_uint64, c_uint64, c_int64, P_uint32]),
    "nslx_rng_uniform": (c_int, [c_uint64, c_uint64, c_int64, P_double]),
    "nslx_rng_normal": (c_int, [c_uint64, c_uint64, c_int64, P_double]),
    "nslx_lattice_honeycomb": (c_int, [c_int, c_int, c_double, P_double, P_double]),
    "nslx_lattice_ring": (c_int, [c_int, c_double, P_double, P_double]),
    "nslx_lattice_armchair_nanotube": (c_int, [c_int, c_int, c_double, P_double, P_double]),
    "nslx_fm_create": (c_int, [c_void_p, c_int, c_int, P_double, c_double, c_double, c_double, c_double,
                               c_int, ctypes.POINTER(c_void_p)]),
    "nslx_fm_create_diag": (c_int, [c_void_p, c_int, c_int, P_double, c_double, c_double, c_double, c_double,
                                    c_int, ctypes.POINTER(c_void_p)]),
    "nslx_fm_destroy": (c_int, [c_void_p]),
    "nslx_fm_populate": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_void_p]),
    "nslx_fm_apply": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int64, c_int, c_void_p]),
    "nslx_fm_apply_noalloc": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int64, c_int, c_void_p]),
    "nslx_fm_logdet": (c_int, [c_void_p, c_int64, c_int, P_double, c_void_p]),
    "nslx_fm_grad_logdet": (c_int, [c_void_p, c_int64, c_void_p, c_void_p]),
    "nslx_solve": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int64, c_double, c_int,
                           c_int64, c_int, c_int, ctypes.POINTER(SolveReport), c_void_p]),
    "nslx_solve_dd": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_double, c_int64,
                              c_int, ctypes.POINTER(SolveReport), c_void_p]),
    "nslx_solve_precond": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int64, c_double,
                                   c_int, c_int64, c_int, ctypes.POINTER(SolveReport), c_void_p]),
    "nslx_fm_solve_direct": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int64, c_int, c_int, c_void_p]),
    "nslx_solve_tss": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_double, c_int64,
                               c_int, c_int, ctypes.POINTER(SolveReport), c_void_p]),
    "nslx_dd_residual": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, P_double, c_int64,
                                 c_void_p]),
    "nslx_fm_export": (c_int, [c_void_p, P_double, c_void_p, c_void_p]),
    "nslx_batch_solve": (c_int, [c_void_p, c_int64, ctypes.POINTER(BatchItem), c_double, c_int, c_int64, c_int,
                                 ctypes.POINTER(BatchReport), c_void_p]),
    "nslx_batch_solve_tss": (c_int, [c_void_p, c_int64, ctypes.POINTER(BatchItem), ctypes.POINTER(c_void_p),
                                     c_double, c_int, c_int64, c_int, ctypes.POINTER(BatchReport)]),
    "nslx_server_create": (c_int, [c_void_p, ctypes.POINTER(ServerConfig), ctypes.POINTER(c_void_p)]),
    "nslx_server_destroy": (c_int, [c_void_p]),
    "nslx_server_submit": (c_int, [c_void_p, ctypes.POINTER(Request), ctypes.POINTER(c_int64)]),
    "nslx_server_wait": (c_int, [c_void_p, c_int64, c_double, c_void_p, ctypes.POINTER(Result)]),
    "nslx_server_stats_get": (c_int, [c_void_p, ctypes.POINTER(ServerStats)]),
    "nslx_correlator_two_point": (c_int, [c_void_p, c_int64, c_double, c_int64, c_int, c_int, c_int64, c_void_p,
                                          ctypes.POINTER(SolveReport), c_void_p]),
    "nslx_correlator_two_point_wall": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_double, c_int64, c_int, c_int,
                                               c_void_p, c_void_p, ctypes.POINTER(SolveReport), c_void_p]),
    "nslx_two_body_quads": (c_int, [P_double, c_int, ctypes.POINTER(c_int64), ctypes.POINTER(ctypes.c_int)]),
    "nslx_correlator_two_body": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int64, ctypes.POINTER(ctypes.c_int),
                                         c_int64, c_double, c_int64, c_int, c_int, c_void_p,
                                         ctypes.POINTER(SolveReport), c_void_p]),
    "nslx_hmc_action": (c_int, [c_void_p, c_void_p, c_double, c_int, P_double, c_void_p]),
    "nslx_hmc_force": (c_int, [c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    "nslx_hmc_trajectory": (c_int, [c_void_p, c_void_p, c_void_p, P_double, c_double, c_double, c_int,
                                    c_double, c_int, ctypes.POINTER(HmcReport), c_void_p]),
    "nslx_hmc_chain": (c_int, [c_void_p, c_void_p, c_double, c_double, c_int, c_int64, c_uint64, c_int,
                               ctypes.POINTER(HmcReport), c_void_p]),
    "nslx_hmc_generate": (c_int, [c_void_p, c_void_p, c_int64, c_double, c_double, c_int, c_int64, c_int64, c_int,
                                  c_uint64, c_int, c_void_p, ctypes.POINTER(MarkovState), c_void_p]),
    "nslx_tile_count": (c_int, [c_int, ctypes.POINTER(c_int)]),
    "nslx_ctx_set_tile": (c_int, [c_void_p, c_int, c_int, c_int, c_int64, c_int]),
    "nslx_ctx_get_tile": (c_int, [c_void_p, c_int, c_int, c_int, c_int64, ctypes.POINTER(c_int)]),
    "nslx_fm_tune": (c_int, [c_void_p, c_int, c_int64, c_int, c_int, ctypes.POINTER(c_int), P_double, c_void_p]),
    "nslx_correlator_two_point_fields": (c_int, [c_void_p, c_int64, c_double, c_int64, c_int, c_int64, c_void_p,
                                                 ctypes.POINTER(SolveReport), c_void_p]),
    "nslx_prof_enable": (c_int, [c_int]),
    "nslx_prof_reset": (c_int, []),
    "nslx_prof_read": (c_int, [ctypes.c_char_p, c_int64]),
}

_lib = None


class NslxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"nslx status {status}: {msg}")
        self.status = status


def lib():
    """Load libnslx.so (built by nsl_amd/build.py).  Raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m nsl_amd.build`")
        # torch ships its own HIP/HSA runtime.  Loaded first, it is the one
        # libnslx.so binds to (same soname), so torch tensors, torch streams and
        # the library's kernels share one runtime and one default stream;
        # loaded after libnslx.so, a second runtime would come up beside the
        # system one and the two would neither order their work nor share
        # stream handles.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, allow=(OK,)):
    if status not in allow:
        msg = lib().nslx_last_error().decode(errors="replace")
        raise NslxError(status, msg)
    return status
