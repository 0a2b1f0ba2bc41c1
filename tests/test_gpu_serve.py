"""Correlator server (continuous batching): results against the oracle's
TwoPointCorrelator, identical alone and in a mix of concurrent requests
(bit for bit: a request's recurrence does not depend on the batch), mixed
precision to its FP64 true-residual target, admission by memory budget, and
latency statistics."""

import math
import time

import numpy as np
import pytest
import torch

import nsl_amd as nsl
from oracle import correlator as ocorr, fermion as F, lattice as olat, rng

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = a.detach().to("cpu", torch.complex128)
    b = torch.as_tensor(b).detach().to("cpu", torch.complex128)
    return float((a - b).abs().max() / b.abs().max())


def field(lat, Nt, beta, U, seed):
    return torch.as_tensor(rng.real_field(seed, 0, (Nt, lat.sites()), math.sqrt(beta / Nt * U)))


def test_server_vs_oracle_cfg0():
    """Honeycomb 3x3, Nt = 16, beta = 4, U = 2; 2 time sources; per-RHS
    1e-12: against the oracle's reference-mode correlator to 1e-9."""
    lat = nsl.Lattice.honeycomb(3, 3)
    phi = field(lat, 16, 4.0, 2.0, 1000)
    srv = nsl.Server()
    try:
        rid = srv.submit(lat, 16, 4.0, phi, n_tsrc=2, tol=1e-12)
        corr, res = srv.wait(rid)
    finally:
        srv.close()
    assert res["status"] == 0, res
    o = F.HubbardExp(olat.honeycomb(3, 3), 16, 4.0)
    o.populate(phi)
    ref, _ = ocorr.two_point(o, 16, 18, 2)
    assert relerr(corr, ref) < 1e-9


def test_server_hubbard_diag_vs_oracle():
    """A HubbardDiag request (served in the same pool as a HubbardExp one):
    against the oracle's correlator on HubbardDiag to 1e-9, and the
    HubbardExp request against its own oracle."""
    lat = nsl.Lattice.honeycomb(3, 3)
    phi = field(lat, 16, 4.0, 2.0, 1001)
    srv = nsl.Server()
    try:
        rd = srv.submit(lat, 16, 4.0, phi, n_tsrc=2, tol=1e-12, fermion_matrix=nsl.HubbardDiag)
        re_ = srv.submit(lat, 16, 4.0, phi, n_tsrc=2, tol=1e-12)
        corr_d, res_d = srv.wait(rd)
        corr_e, res_e = srv.wait(re_)
    finally:
        srv.close()
    assert res_d["status"] == 0 and res_e["status"] == 0, (res_d, res_e)
    od = F.HubbardDiag(olat.honeycomb(3, 3), 16, 4.0)
    od.populate(phi)
    ref_d, _ = ocorr.two_point(od, 16, 18, 2)
    assert relerr(corr_d, ref_d) < 1e-9
    oe = F.HubbardExp(olat.honeycomb(3, 3), 16, 4.0)
    oe.populate(phi)
    ref_e, _ = ocorr.two_point(oe, 16, 18, 2)
    assert relerr(corr_e, ref_e) < 1e-9
    assert relerr(corr_d, corr_e) > 1e-6          # the two discretisations differ


def test_server_alone_equals_in_mix():
    lats = [nsl.Lattice.honeycomb(3, 3), nsl.Lattice.honeycomb(6, 6), nsl.Lattice.armchair_nanotube(3, 4),
            nsl.Lattice.ring(10)]
    reqs = [(lats[0], 16, 4.0, 2.0, 1), (lats[1], 32, 6.0, 2.0, 2), (lats[2], 20, 3.0, 1.5, 1),
            (lats[3], 12, 2.0, 3.0, 3), (lats[1], 16, 3.0, 1.0, 1)]
    mine = 1
    lat, Nt, beta, U, nts = reqs[mine]
    phi_mine = field(lat, Nt, beta, U, 77)
    srv = nsl.Server()
    try:
        corr_alone, res_alone = srv.wait(srv.submit(lat, Nt, beta, phi_mine, n_tsrc=nts, tol=1e-11))
    finally:
        srv.close()
    srv = nsl.Server()            # same replay length: reported iterations count whole replays
    try:
        ids = {}
        for k, (lt, nt, bt, u, n) in enumerate(reqs):
            phi = phi_mine if k == mine else field(lt, nt, bt, u, 100 + k)
            ids[k] = srv.submit(lt, nt, bt, phi, n_tsrc=n, tol=1e-11)
            time.sleep(0.002)        # arrivals while earlier requests are on the GPU
        out = {k: srv.wait(i) for k, i in ids.items()}
        st = srv.stats()
    finally:
        srv.close()
    assert torch.equal(out[mine][0], corr_alone)
    assert out[mine][1]["iterations"] == res_alone["iterations"]
    assert all(r[1]["status"] == 0 for r in out.values())
    assert st["completed"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0


@pytest.mark.parametrize("prec", [nsl.F32, nsl.BF16X3])
def test_server_mixed_precision(prec):
    lat = nsl.Lattice.honeycomb(6, 6)
    phi = field(lat, 16, 4.0, 2.0, 5)
    srv = nsl.Server()
    try:
        c64, r64 = srv.wait(srv.submit(lat, 16, 4.0, phi, n_tsrc=2, tol=1e-10))
        cm, rm = srv.wait(srv.submit(lat, 16, 4.0, phi, n_tsrc=2, tol=1e-10, precision=prec))
    finally:
        srv.close()
    assert r64["status"] == 0 and rm["status"] == 0, (r64, rm)
    assert rm["max_rel_residual"] <= 1e-10 or rm["escalated"]
    assert relerr(cm, c64) < 1e-7


def test_server_admission_budget():
    """A budget that holds about two requests: all complete, never more than
    the budget admitted at once."""
    lat = nsl.Lattice.honeycomb(6, 6)
    Nt = 16
    # the footprint the server charges for one request (serve.cpp footprint())
    one = 11 * 72 * Nt * 72 * 16 + 72 * 72 * 16 + 192 * 128 * 16 + Nt * 72 * 40 + Nt * 72 * 72 * 16 + (64 << 20)
    srv = nsl.Server(hbm_budget=2.5 * one)
    try:
        ids = [srv.submit(lat, Nt, 4.0, field(lat, Nt, 4.0, 2.0, 300 + k), tol=1e-10) for k in range(6)]
        res = [srv.wait(i)[1] for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    assert all(r["status"] == 0 for r in res)
    assert 1 <= st["admitted_peak"] <= 2
    assert st["bytes_peak"] <= 2.5 * one
    assert max(r["queue_ms"] for r in res) > 0


def test_server_errors():
    lat = nsl.Lattice.honeycomb(3, 3)
    srv = nsl.Server(hbm_budget=1e6)
    try:
        with pytest.raises(nsl.NslxError):                  # more time sources than Nt
            srv.submit(lat, 16, 4.0, np.zeros((16, 18)), n_tsrc=17)
        with pytest.raises(nsl.NslxError):                  # larger than the memory budget
            srv.submit(lat, 16, 4.0, np.zeros((16, 18)))
        with pytest.raises(ValueError):                     # not a fermion matrix class
            srv.submit(lat, 16, 4.0, np.zeros((16, 18)), fermion_matrix=dict)
    finally:
        srv.close()
    import ctypes
    from nsl_amd import _lib
    srv = nsl.Server()
    try:
        phi = np.zeros((16, 18), dtype=np.complex128)
        req = _lib.Request(18, 16, lat.K.ctypes.data_as(_lib.P_double), 4.0, 0.0, 0, phi.ctypes.data, 1, 1e-10,
                           nsl.F64, 1000, 7)                # no such discretisation
        rid = ctypes.c_int64()
        assert _lib.lib().nslx_server_submit(srv._h, ctypes.byref(req), ctypes.byref(rid)) == _lib.E_ARG
    finally:
        srv.close()


def test_server_admission_waits_for_contiguous_slots():
    """A pool of 100 RHS slots and requests of 72 and 18 RHS (honeycomb
    6x6 / 3x3, one time source each): as requests finish out of order the
    free slots fragment; a request that fits by count but not in one run
    must wait at the queue front, never fail.  Every result must be OK and
    equal to the same request served alone."""
    big, small = nsl.Lattice.honeycomb(6, 6), nsl.Lattice.honeycomb(3, 3)
    spec = [(small, 16, 2.0, 1.0, 3000), (big, 16, 3.0, 2.0, 3001), (small, 16, 6.0, 3.0, 3002),
            (small, 16, 2.0, 1.0, 3003), (big, 16, 2.0, 1.0, 3004), (small, 16, 5.0, 3.0, 3005),
            (small, 16, 2.0, 1.0, 3006), (big, 16, 4.0, 2.0, 3007)]
    srv = nsl.Server(max_rhs=100, iters_per_replay=8)
    try:
        ids = [srv.submit(lat, Nt, beta, field(lat, Nt, beta, U, seed), tol=1e-10) for lat, Nt, beta, U, seed in spec]
        out = [srv.wait(i) for i in ids]
    finally:
        srv.close()
    for (lat, Nt, beta, U, seed), (corr, res) in zip(spec, out):
        assert res["status"] == 0, res
    lat, Nt, beta, U, seed = spec[4]
    alone = nsl.Server(max_rhs=100, iters_per_replay=8)
    try:
        c1, r1 = alone.wait(alone.submit(lat, Nt, beta, field(lat, Nt, beta, U, seed), tol=1e-10))
    finally:
        alone.close()
    assert r1["status"] == 0
    assert torch.equal(c1, out[4][0])


# ---- time-slice-preconditioned serving (precond_workers > 0) ----------------

def test_tss_server_vs_oracle_cfg0():
    """The preconditioned server on the CPU-runnable lattice (honeycomb 3x3,
    Nt = 16, beta = 4, U = 2; 2 time sources; per-RHS 1e-12) against the
    oracle's reference-mode correlator to 1e-9, with 3 worker streams."""
    lat = nsl.Lattice.honeycomb(3, 3)
    phi = field(lat, 16, 4.0, 2.0, 1000)
    srv = nsl.Server(precond_workers=3)
    try:
        rid = srv.submit(lat, 16, 4.0, phi, n_tsrc=2, tol=1e-12)
        corr, res = srv.wait(rid)
    finally:
        srv.close()
    assert res["status"] == 0, res
    assert res["max_rel_residual"] <= 1e-12
    o = F.HubbardExp(olat.honeycomb(3, 3), 16, 4.0)
    o.populate(phi)
    ref, _ = ocorr.two_point(o, 16, 18, 2)
    assert relerr(corr, ref) < 1e-9


def _tss_stream():
    """Six requests of the serving stream's kind: honeycomb Nx 18 / 72 and a
    (6,6) nanotube of 4 cells, Nt 16-64, beta 2-12, U 1-4."""
    shapes = [(nsl.Lattice.honeycomb(3, 3), 16, 11.5, 3.7), (nsl.Lattice.honeycomb(6, 6), 32, 6.0, 2.0),
              (nsl.Lattice.armchair_nanotube(6, 4), 16, 3.0, 1.2), (nsl.Lattice.honeycomb(3, 3), 64, 8.0, 3.0),
              (nsl.Lattice.honeycomb(6, 6), 16, 2.5, 4.0), (nsl.Lattice.armchair_nanotube(6, 4), 32, 12.0, 2.5)]
    return [(lat, Nt, beta, field(lat, Nt, beta, U, 3100 + k)) for k, (lat, Nt, beta, U) in enumerate(shapes)]


@pytest.mark.parametrize("prec", [nsl.F64, nsl.F32])
def test_tss_server_alone_equals_in_mix(prec):
    """Continuous batching on worker streams: six requests submitted at once
    (their factorisations and solves overlap on the device) give the same
    correlators bit for bit as each request alone on the same server, and
    every request reaches its 1e-10 per-RHS target."""
    reqs = _tss_stream()
    srv = nsl.Server(precond_workers=4, hbm_budget=64e9)
    try:
        ids = [srv.submit(lat, Nt, beta, phi, tol=1e-10, precision=prec) for lat, Nt, beta, phi in reqs]
        mix = [srv.wait(i) for i in ids]
        alone = []
        for lat, Nt, beta, phi in reqs:
            alone.append(srv.wait(srv.submit(lat, Nt, beta, phi, tol=1e-10, precision=prec)))
        st = srv.stats()
    finally:
        srv.close()
    for (cm, rm), (ca, ra) in zip(mix, alone):
        assert rm["status"] == 0 and ra["status"] == 0, (rm, ra)
        assert rm["max_rel_residual"] <= 1e-10, rm
        assert torch.equal(cm, ca)
    assert st["completed"] == 2 * len(reqs)
    assert st["admitted_peak"] >= 2          # the mix really ran side by side


def test_tss_server_matches_krylov_server():
    """The same request (honeycomb 6x6, Nt 32, beta 6, U 2) through the
    preconditioned server and the Krylov pool, FP64, per-RHS 1e-12: the
    correlators agree to 1e-8 (cond(M) < 1e4 here)."""
    lat, Nt, beta, phi = _tss_stream()[1]
    out = []
    for workers in (2, 0):
        srv = nsl.Server(precond_workers=workers)
        try:
            out.append(srv.wait(srv.submit(lat, Nt, beta, phi, tol=1e-12)))
        finally:
            srv.close()
    (ct, rt), (ck, rk) = out
    assert rt["status"] == 0 and rk["status"] == 0
    assert relerr(ct, ck) < 1e-8


def test_tss_server_errors():
    lat = nsl.Lattice.honeycomb(3, 3)
    phi = field(lat, 16, 4.0, 2.0, 1000)
    srv = nsl.Server(precond_workers=2)
    try:
        with pytest.raises(nsl.NslxError):      # the preconditioner runs in F64 or F32 only
            srv.submit(lat, 16, 4.0, phi, precision=nsl.BF16X3)
        with pytest.raises(nsl.NslxError):      # HubbardExp only
            srv.submit(lat, 16, 4.0, phi, fermion_matrix=nsl.HubbardDiag)
    finally:
        srv.close()
