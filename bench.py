"""Benchmark: CG solves/s to a 1e-12 residual on MM^dag, preconditioned with
double-double correction against the plain FP64 GPU path, on the 12x12 periodic honeycomb (Nx = 288),
Nt = 64, beta = 8, U = 3: 16 seeded phi fields x 8 Gaussian right-hand
sides = 128 solves per step, all resident in HBM before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--fp64-steps K64]
                    [--precond f64|f32] [--no-cpu] [--no-fp64] [--no-refcrit]
                    [--no-batch] [--no-serve] [--prof-every N]

A step solves all 128 systems to |b - MM^dag x| <= 1e-12 |b| per right-hand
side, that residual evaluated in double-double on the FP64 operator (the
fields are so ill-conditioned -- cond(MM^dag) up to ~1e10 -- that an FP64
vector cannot represent a solution with a 1e-12 residual: x is returned as a
double-double pair x + x_lo).  Every step starts with populate(phi), so the
preconditioned forms refactorise inside the timed region.

  value:          CG on MM^dag preconditioned by the time-slice direct solve
                  of M and M^dag (factorised and applied on the matrix cores,
                  FP64 by default with 2 chunks of 32 slices; --precond f32:
                  FP32 with 8 chunks of 8), FP64 CG recurrence,
                  double-double defect correction
  fp64:           the plain FP64 path -- unpreconditioned CG (the reference's
                  algorithm, CG.tpp:14-113) per RHS with FP64 reliable updates
                  and the solution in double-double
  f32_tss:        the preconditioned solver with the time-slice solve in the
  (or f64_tss)    other precision
  refcrit:        the reference's own criterion (x0 = b, one alpha/beta per
                  field's 8-RHS batch, sum |r|^2 <= 1e-24 on the recursive
                  residual, CG.tpp:70) on a few fields, FP64

Prints one JSON line.  Profiling helpers (no JSON contract line):
    python bench.py --hops N        N plain M and M^dag applies per precision
    python bench.py --tss N         N populate + preconditioned solves (in the
                                    --precond precision), for rocprofv3 runs
    python bench.py --copy-calib M  three device copies of M MiB (scales the
                                    FETCH_SIZE / WRITE_SIZE counters)
    python bench.py --dgemm-ceiling S  the FP64 matrix-core rate the platform
                                    library reaches (torch.matmul, float64) on
                                    S = N (N x N x N) or MxNxK: the practical
                                    ceiling beside the published 78.6 TFLOP/s
"""

import argparse
import ctypes
import json
import math
import os
import platform
import statistics
import sys
import time

import numpy as np

METRIC = "CG solves/s to 1e-12 resid, Nx=288 Nt=64: mixed vs FP64; p99 request latency"
L, NT, BETA, U = 12, 64, 8.0, 3.0
NFIELD, NRHS = 16, 8
TOL = 1e-12
PROF_EVERY = 2        # every 2nd launch of a kernel class is timed with HIP events
# MI355X datasheet peaks: HBM3E GB/s; dense matrix TFLOP/s per precision;
# FP64 vector (VALU) TFLOP/s
PEAK = {"hbm": 8000.0, "f64": 78.6, "f32": 157.3, "bf16": 2500.0, "valu64": 78.6}


def fields():
    """Seeded inputs from the library's Philox generator: phi_f ~ N(0, sqrt(U~))
    with seed 1000+f, b_{f,k} complex Gaussian with seed 2000+16f+k."""
    import nsl_amd as nsl
    Nx = 2 * L * L
    Ut = BETA / NT * U
    phi = np.stack([nsl.real_field(1000 + f, (NT, Nx), math.sqrt(Ut)) for f in range(NFIELD)])
    b = np.stack([nsl.complex_normal(2000 + 16 * f + k, (NT, Nx))
                  for f in range(NFIELD) for k in range(NRHS)])
    return phi, b


def kernel_bytes_flops(name, nb, Nx, Nt):
    """Algorithmic bytes / flops of one hop launch over nb right-hand sides
    (E and the phase table are < 2% and excluded)."""
    prec = name.split("_")[2]
    mode = name.split("_")[3]
    w = 16 if prec == "f64" else 8
    vec = {"plain": 2, "cga": 4, "cgb": 6, "cgi": 4}[mode]
    bytes_ = vec * Nt * Nx * w * nb
    flops = 4.0 * Nx * Nx * Nt * nb * (3 if prec == "bf16x3" else 1)
    return bytes_, flops, prec


def parse_prof(text):
    """Rows (name, launches, total_ms, timed, flops, bytes) of nslx_prof_read
    (flops and bytes: summed over the timed launches; 0 for hop classes)."""
    rows = []
    for line in text.strip().splitlines():
        f = line.split()
        if len(f) == 4:
            f += ["0", "0"]
        rows.append((f[0], int(f[1]), float(f[2]), int(f[3]), float(f[4]), float(f[5])))
    return rows


def class_roofline(name, flops, bytes_, avg_s):
    """Bound, achieved and peak of a kernel class from its algorithmic work per
    launch: the GEMM-shaped classes (batched GEMM, chunk products, persistent
    recurrences) on the matrix cores of their precision, the double-double
    hop on the FP64 vector units, everything else on HBM."""
    if name.startswith(("bgemm_", "chunkprod_", "chain_")):
        prec = name.split("_")[1]
        return "mfma", flops / avg_s / 1e12, PEAK[prec], "TFLOP/s"
    if name.startswith("dd_hop"):
        return "valu", flops / avg_s / 1e12, PEAK["valu64"], "TFLOP/s"
    return "hbm", bytes_ / avg_s / 1e9, PEAK["hbm"], "GB/s"


def roofline_from_prof(text, nb, Nx, Nt, rep, traffic=None):
    """Roofline of the kernel class with the largest total time (timed
    average x launches).  Hop classes: algorithmic work from the shape, the
    fused CG kernels counted with the average number of active right-hand
    sides (the solver's active-RHS x iteration count of that precision over
    the launches of the class); other classes report their own work."""
    rows = parse_prof(text)
    if not rows:
        return None, []
    name, n, ms, timed, fl_t, by_t = max(rows, key=lambda r: r[2] / r[3] * r[1])
    avg_s = ms / timed / 1e3
    if name.startswith("hop_"):
        prec, mode = name.split("_")[2], name.split("_")[3]
        active = nb
        if mode in ("cga", "cgb"):
            key = "rhs_iters_f64" if prec == "f64" else "rhs_iters_low"
            if rep.get(key):
                active = rep[key] / n
        by, fl, prec = kernel_bytes_flops(name, active, Nx, Nt)
        if prec == "bf16x3":
            bound, achieved, peak, unit = "hbm", by / avg_s / 1e9, PEAK["hbm"], "GB/s"
        else:
            bound, achieved, peak, unit = "mfma", fl / avg_s / 1e12, PEAK[prec], "TFLOP/s"
    else:
        fl, by = fl_t / timed, by_t / timed
        bound, achieved, peak, unit = class_roofline(name, fl, by, avg_s)
    rf = {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
          "frac": round(achieved / peak, 4), "traffic": traffic, "kernel": name,
          "avg_us": round(avg_s * 1e6, 2), "launches": n, "timed_launches": timed,
          "algorithmic_bytes": by, "algorithmic_flops": fl}
    return rf, rows


def kernel_table(rows, top=8):
    """The classes by estimated total time, with each one's own roofline."""
    out = []
    tot = sum(r[2] / r[3] * r[1] for r in rows) or 1.0
    for name, n, ms, timed, fl_t, by_t in sorted(rows, key=lambda r: -r[2] / r[3] * r[1])[:top]:
        avg = ms / timed / 1e3
        d = {"kernel": name, "launches": n, "avg_us": round(avg * 1e6, 2),
             "share": round(ms / timed * n / tot, 4)}
        if not name.startswith("hop_") and timed:
            b, a, p, u = class_roofline(name, fl_t / timed, by_t / timed, avg)
            d.update({"bound": b, "achieved": round(a, 3), "unit": u, "frac": round(a / p, 4)})
        out.append(d)
    return out


def read_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hop_pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(kernel)
    except (OSError, ValueError):
        return None


def host_info(cores):
    """CPU model, thread counts and ATen's parallel configuration."""
    import torch
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    par = torch.__config__.parallel_info().strip().splitlines()
    return {"cpu_model": model, "nproc": os.cpu_count(), "threads": cores,
            "torch_num_threads": torch.get_num_threads(),
            "parallel_info": [p.strip() for p in par[:3]]}


def cpu_baseline(phi, b, cores, iters_needed, chunk_s=6.0):
    """The oracle (op-for-op libtorch complex128 restatement of the reference
    CG on HubbardExp::MMdagger) on one field x 8 RHS -- the reference's own
    batched call -- with all host threads.  A complete solve with the
    reference's batch-global stop takes iters_needed iterations (measured on
    the GPU with the same recurrence on the same field), minutes on the host,
    so the per-iteration time is taken as the median of three timed chunks of
    iterations (~chunk_s each, after a warm-up) and scaled to iters_needed."""
    import torch
    from oracle import fermion as F, lattice as olat, solvers as osol
    torch.set_num_threads(cores)
    lat = olat.honeycomb(L, L)
    h = F.HubbardExp(lat, NT, BETA)
    h.populate(torch.as_tensor(phi[0]))
    bb = torch.as_tensor(b[:NRHS])
    osol.cg(h.MMdagger, bb, eps=0.0, maxIter=3)     # warm-up (E memo, threads)
    t0 = time.perf_counter()
    osol.cg(h.MMdagger, bb, eps=0.0, maxIter=5)
    per_it = (time.perf_counter() - t0) / 6
    n = max(10, int(chunk_s / per_it))
    per = []
    for _ in range(3):
        t0 = time.perf_counter()
        osol.cg(h.MMdagger, bb, eps=0.0, maxIter=n)
        per.append((time.perf_counter() - t0) / (n + 1))
    per_it = statistics.median(per)
    solve_s = per_it * iters_needed
    info = host_info(cores)
    return {"value": round(NRHS / solve_s, 6), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"3 x {n} CG iterations on 1 field x {NRHS} RHS (median {per_it * 1e3:.2f} ms/iteration), "
                      f"scaled to the {iters_needed} iterations of the reference's batch-global CG "
                      f"(x0 = b, sum|r|^2 <= 1e-24) on that field; a complete solve would take "
                      f"{solve_s:.0f} s", **info}


def batch_mix_specs(nprob=64, seed=7000):
    """The batching form's mix: nprob independent problems, half honeycomb
    flakes with Nx in {18, 72, 288, 1152}, half (6,6) armchair nanotubes of
    4-32 unit cells (Nx 96-768); Nt in {16, 32, 64, 128}, beta in [2, 12],
    U in [1, 4], 1-16 Gaussian RHS, one Philox field each (all draws from
    the library's Philox stream at seed)."""
    import nsl_amd as nsl
    u = nsl.rng_uniform(seed, 0, nprob * 6).reshape(nprob, 6)
    specs = []
    for k in range(nprob):
        if u[k, 0] < 0.5:
            lat = ("honeycomb", [3, 6, 12, 24][int(u[k, 5] * 4)])
        else:
            lat = ("nanotube", 4 + int(u[k, 5] * 29))
        specs.append({"lattice": lat, "Nt": [16, 32, 64, 128][int(u[k, 1] * 4)], "beta": 2.0 + 10.0 * u[k, 2],
                      "U": 1.0 + 3.0 * u[k, 3], "nrhs": 1 + int(u[k, 4] * 16), "field_seed": seed + 1 + k,
                      "rhs_seed": seed + 500 + k})
    return specs


def mix_lattice(key, cache):
    import nsl_amd as nsl
    if key not in cache:
        cache[key] = (nsl.Lattice.honeycomb(key[1], key[1]) if key[0] == "honeycomb"
                      else nsl.Lattice.armchair_nanotube(6, key[1]))
    return cache[key]


def build_mix_problems(specs):
    """(fm, b, phi) per spec on the device."""
    import torch
    import nsl_amd as nsl
    lats, out = {}, []
    for sp in specs:
        lat = mix_lattice(sp["lattice"], lats)
        Nx = lat.sites()
        fm = nsl.HubbardExp(lat, sp["Nt"], sp["beta"])
        phi = torch.as_tensor(nsl.real_field(sp["field_seed"], (sp["Nt"], Nx),
                                             math.sqrt(sp["beta"] / sp["Nt"] * sp["U"]))).cuda()
        fm.populate(phi)
        b = torch.as_tensor(nsl.complex_normal(sp["rhs_seed"], (sp["nrhs"], sp["Nt"], Nx))).cuda()
        out.append((fm, b, phi))
    return out


def batching_leg(nprob=64, seed=7000, tol=1e-10, krylov=True, krylov_maxit=5000):
    """The batching form on its full mix (batch_mix_specs): every step
    re-populates all 64 fields (so each solve factorises afresh) and solves
    every RHS to |b - MM^dag x| <= 1e-10 |b|:
      (a) time-slice-preconditioned CG with double-double correction, the
          problems spread over 8 worker streams (nslx_batch_solve_tss);
      (b) the same on one stream (one problem after another);
      (c) the Krylov pool: unpreconditioned CG for all problems in one
          grouped launch per iteration, HIP-graph replayed, FP32 recurrence
          with FP64 reliable updates, capped at krylov_maxit iterations.
    Returns solves/s (RHS / wall time) and convergence of each."""
    import torch
    import nsl_amd as nsl
    specs = batch_mix_specs(nprob, seed)
    probs = build_mix_problems(specs)
    pb = [(fm, b) for fm, b, _ in probs]
    nsolves = sum(b.shape[0] for b in (p[1] for p in probs))
    sizes = sorted({(p[0].Nx, p[0].Nt) for p in probs})
    out = {"problems": nprob, "rhs": nsolves, "tol": tol,
           "workload": "64 fields: honeycomb Nx 18/72/288/1152 and (6,6) nanotubes of 4-32 cells (Nx 96-768), "
                       "Nt 16/32/64/128, beta 2-12, U 1-4, 1-16 Gaussian RHS each; fields re-populated per run",
           "distinct_shapes": len(sizes), "max_Nx": max(n for n, _ in sizes), "max_Nt": max(t for _, t in sizes)}
    for name, ns in (("tss_streams8", 8), ("tss_one_stream", 1)):
        nsl.batch_solve_tss(pb, tol, nsl.F64, nstreams=ns)          # warm-up: plans and buffers
        for fm, _, phi in probs:
            fm.populate(phi)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, rep = nsl.batch_solve_tss(pb, tol, nsl.F64, nstreams=ns)
        dt = time.perf_counter() - t0
        out[name] = {"solves_per_s": round(nsolves / dt, 1), "seconds": round(dt, 3),
                     "converged_items": rep["converged_items"], "max_rel_residual": rep["max_rel_residual"],
                     "max_iterations": rep["iterations"], "preconditioner_applies": rep["replays"]}
    out["streams_speedup"] = round(out["tss_streams8"]["solves_per_s"] / out["tss_one_stream"]["solves_per_s"], 2)
    if krylov:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, rep = nsl.batch_solve(pb, tol=tol, precision=nsl.F32, maxIter=krylov_maxit, use_graph=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out["krylov_pool_f32"] = {"solves_per_s": round(nsolves / dt, 1), "seconds": round(dt, 3),
                                  "converged_items": rep["converged_items"], "iterations": rep["iterations"],
                                  "max_rel_residual": rep["max_rel_residual"], "maxit": krylov_maxit,
                                  "note": "solves/s counts every RHS, including those of the items the "
                                          "unpreconditioned CG left unconverged at maxit"}
        out["tss_over_krylov"] = round(out["tss_streams8"]["solves_per_s"] / out["krylov_pool_f32"]["solves_per_s"], 1)
    return out


SERVE_RATE = 10.0      # requests / s (Poisson), about 60 % of what the server sustains on this stream
SERVE_N = 32           # requests per serving run


def serve_requests(seed=9000, n=SERVE_N, rate=SERVE_RATE, subset=None):
    """The serving stream: n requests drawn like the batching mix
    (batch_mix_specs at seed: honeycomb Nx 18-1152, (6,6) nanotubes Nx
    96-768, Nt 16-128, beta 2-12, U 1-4), each a Philox-seeded phi ~
    N(0, sqrt(U beta / Nt)) asking for the two-point correlator with Nx
    point sources at t = 0; Poisson arrivals at rate.  subset(Nx, Nt)
    keeps only the requests it accepts."""
    import nsl_amd as nsl
    specs = batch_mix_specs(n, seed)
    lats, out = {}, []
    for sp in specs:
        lat = mix_lattice(sp["lattice"], lats)
        if subset and not subset(lat.sites(), sp["Nt"]):
            continue
        phi = nsl.real_field(sp["field_seed"], (sp["Nt"], lat.sites()), math.sqrt(sp["beta"] / sp["Nt"] * sp["U"]))
        out.append((lat, sp["Nt"], sp["beta"], phi))
    arrivals = -np.log(1.0 - nsl.rng_uniform(seed + 999, 0, len(out))) / rate
    return out, arrivals


def serving_leg(prec, tol=1e-10, workers=0, rate=SERVE_RATE, subset=None, maxit=100000):
    """One Poisson request stream through the correlator server: workers > 0
    serves with the time-slice-preconditioned solver on that many streams
    (prec = the preconditioner's precision), workers = 0 with the Krylov
    pool (prec = the CG recurrence's).  Returns latency percentiles and
    sustained solves/s (RHS solved / wall time from the first submit)."""
    import nsl_amd as nsl
    reqs, gaps = serve_requests(rate=rate, subset=subset)
    srv = nsl.Server(hbm_budget=200e9, precond_workers=workers)
    try:
        warm = srv.submit(reqs[0][0], reqs[0][1], reqs[0][2], reqs[0][3], tol=tol, precision=prec, maxIter=maxit)
        srv.wait(warm)
        t0 = time.perf_counter()
        ids = []
        for (lat, Nt, beta, phi), gap in zip(reqs, gaps):
            time.sleep(float(gap))
            ids.append(srv.submit(lat, Nt, beta, phi, tol=tol, precision=prec, maxIter=maxit))
        res = [srv.wait(i)[1] for i in ids]
        wall = time.perf_counter() - t0
    finally:
        srv.close()
    lat_ms = np.array([r["latency_ms"] for r in res])
    rhs = sum(lat.sites() for lat, _, _, _ in reqs)
    return {"p50_ms": round(float(np.percentile(lat_ms, 50)), 3), "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "solves_per_s": round(rhs / wall, 1), "requests": len(res), "rhs": rhs, "rate_per_s": rate,
            "workers": workers, "ok": sum(int(r["status"] == 0) for r in res),
            "escalated": sum(r["escalated"] for r in res), "max_iterations": max(r["iterations"] for r in res),
            "max_rel_residual": max(r["max_rel_residual"] for r in res)}


def hmc_leg(ntraj=100, nforce=3):
    """The HMC side of the path: (a) the CPU-runnable configuration (honeycomb
    3x3, Nt 16, beta 4, U 2): ntraj leapfrog trajectories of length 1 with 10
    steps (each 11 force evaluations of both species, log det at both ends,
    Metropolis in FP64) from a Philox field, time per trajectory; (b) one
    force evaluation (gradLogDetM of both species) and one log det at the
    headline shapes (Nx 288, Nt 64, beta 8, U 3), median of nforce, with the
    kernel classes timed."""
    import torch
    import nsl_amd as nsl
    out = {}
    A = nsl.HubbardAction(nsl.Lattice.honeycomb(3, 3), 16, 4.0, 2.0)
    phi = torch.as_tensor(nsl.real_field(1000, (16, 18), math.sqrt(4.0 / 16 * 2.0))).cuda()
    A.chain(phi, 3, 1.0, 10, 2999)                       # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = A.chain(phi, ntraj, 1.0, 10, 3000)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out["cfg0"] = {"trajectories": ntraj, "ms_per_trajectory": round(dt / ntraj * 1e3, 3),
                   "acceptance": round(sum(int(r["accepted"]) for r in reps) / ntraj, 3),
                   "workload": "honeycomb 3x3 (Nx 18), Nt 16, beta 4, U 2, trajectory length 1, 10 leapfrog steps"}
    A = nsl.HubbardAction(nsl.Lattice.honeycomb(L, L), NT, BETA, U)
    phi = torch.as_tensor(nsl.real_field(1000, (NT, 2 * L * L), math.sqrt(BETA / NT * U))).cuda()
    A.force(phi)                                          # warm-up (E memo)
    L_ = nsl._lib.lib()
    ft, lt = [], []
    L_.nslx_prof_reset()
    L_.nslx_prof_enable(1)
    for _ in range(nforce):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        A.force(phi)
        torch.cuda.synchronize()
        ft.append(time.perf_counter() - t0)
    L_.nslx_prof_enable(0)
    text = prof_text_of(L_)
    A.fm.populate(phi)
    for _ in range(nforce):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        A.fm.logDetM()
        torch.cuda.synchronize()
        lt.append(time.perf_counter() - t0)
    out["nx288"] = {"force_ms": round(statistics.median(ft) * 1e3, 3),
                    "logdet_ms": round(statistics.median(lt) * 1e3, 3),
                    "force_kernels": kernel_table(parse_prof(text), top=5),
                    "workload": "honeycomb 12x12 (Nx 288), Nt 64, beta 8, U 3; force = gradLogDetM of both species "
                                "+ gauge force; log det of one species"}
    return out


def bicgstab_leg(nsolves=100, cpu=True):
    """The reference's BiCGStab example (Executables/Examples/
    example_BiCGStab.cpp:86-104) on the headline lattice: phi = 0,
    BiCGStab(M, MMdagger, 1e-10) on a point source b at (t, x) = (0, 0),
    nsolves solves timed; the device solver keeps every scalar on the GPU.
    Beside it the oracle's BiCGStab (the reference's recurrence on libtorch
    CPU, all host threads) on the same system, one solve."""
    import torch
    import nsl_amd as nsl
    Nx = 2 * L * L
    fm = nsl.HubbardExp(nsl.Lattice.honeycomb(L, L), NT, BETA)
    fm.populate(torch.zeros((NT, Nx), dtype=torch.complex128, device="cuda"))
    b = torch.zeros((NT, Nx), dtype=torch.complex128, device="cuda")
    b[0, 0] = 1.0
    bi = nsl.BiCGStab(fm, nsl.MMdagger, 1e-10, 10000)
    x = bi(b)                                              # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsolves):
        x = bi(b)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / nsolves
    r = b - fm.MMdagger(x)
    out = {"workload": "honeycomb 12x12 (Nx 288), Nt 64, beta 8, phi = 0, point source, BiCGStab on MMdagger, "
                       "eps 1e-10 (the reference's |r|^2 <= eps^2 stop)",
           "ms_per_solve": round(dt * 1e3, 3), "iterations": bi.report["iterations"], "converged": bool(bi.converged),
           "true_residual": float(r.abs().pow(2).sum().sqrt())}
    if cpu:
        from oracle import fermion as F, lattice as olat, solvers as osol
        torch.set_num_threads(os.cpu_count() or 1)
        o = F.HubbardExp(olat.honeycomb(L, L), NT, BETA)
        o.populate(torch.zeros((NT, Nx), dtype=torch.complex128))
        bc = b.cpu()
        t0 = time.perf_counter()
        xo, info = osol.bicgstab(o.MMdagger, bc, 1e-10, 10000)
        out["oracle_cpu_ms_per_solve"] = round((time.perf_counter() - t0) * 1e3, 1)
        out["oracle_iterations"] = info["iterations"]
        out["max_abs_diff_vs_oracle"] = float((x.cpu() - xo).abs().max())
    return out


def prof_text_of(L_):
    buf = ctypes.create_string_buffer(1 << 16)
    L_.nslx_prof_read(buf, len(buf))
    return buf.value.decode()


def hop_profile(n):
    """n plain M and M^dag applies at each precision on the bench block;
    prints the per-kernel average duration from HIP events."""
    import torch
    import nsl_amd as nsl
    phi_h, b_h = fields()
    fm = nsl.HubbardExp(nsl.Lattice.honeycomb(L, L), NT, BETA)
    fm.populate(torch.as_tensor(phi_h).cuda())
    b = torch.as_tensor(b_h).cuda()
    L_ = nsl._lib.lib()
    for prec in (nsl.F64, nsl.F32, nsl.BF16X3):
        x = b.to(torch.complex128 if prec == nsl.F64 else torch.complex64)
        y = torch.empty_like(x)
        fm.apply(nsl.M, x, y, precision=prec)
        torch.cuda.synchronize()
        L_.nslx_prof_reset()
        L_.nslx_prof_enable(1)
        for _ in range(n):
            fm.apply(nsl.M, x, y, precision=prec)
            fm.apply(nsl.Mdagger, x, y, precision=prec)
        torch.cuda.synchronize()
        L_.nslx_prof_enable(0)
        for name, k, ms, timed, _, _ in parse_prof(prof_text_of(L_)):
            print(f"{name} launches={k} avg_us={ms / timed * 1e3:.2f}", flush=True)


def tss_profile(n, precond):
    """n populate + preconditioned (time-slice solve in precond's precision)
    solves of the bench workload, for profiler runs; prints each step's
    time."""
    import torch
    import nsl_amd as nsl
    phi_h, b_h = fields()
    fm = nsl.HubbardExp(nsl.Lattice.honeycomb(L, L), NT, BETA)
    phi = torch.as_tensor(phi_h).cuda()
    b = torch.as_tensor(b_h).cuda()
    x = torch.empty_like(b)
    cg = nsl.CGTimeSlice(fm, TOL, 2000, precision=nsl.F64 if precond == "f64" else nsl.F32)
    for k in range(n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fm.populate(phi)
        cg(b, out=x)
        torch.cuda.synchronize()
        print(f"step {k}: {(time.perf_counter() - t0) * 1e3:.1f} ms, converged {cg.converged}, "
              f"max_rel_residual {cg.report['max_rel_residual']:.3e}", flush=True)


def copy_calib(mib):
    import torch
    n = mib << 20
    x = torch.ones(n, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)


def dgemm_ceiling(shape):
    """FP64 GEMM rate of the platform library (rocBLAS / hipBLASLt through
    torch.matmul) on M x N x K (shape "N" or "MxNxK"): 2 warm-up products,
    then 5 timed."""
    import torch
    dims = [int(v) for v in shape.lower().split("x")]
    m, n, k = dims * 3 if len(dims) == 1 else dims
    a = torch.randn(m, k, dtype=torch.float64, device="cuda")
    b = torch.randn(k, n, dtype=torch.float64, device="cuda")
    for _ in range(2):
        c = a @ b
    torch.cuda.synchronize()
    reps = 5 if m * n * k >= 1 << 33 else 50
    t0 = time.perf_counter()
    for _ in range(reps):
        c = a @ b
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"dgemm": [m, n, k], "us": round(dt * 1e6, 1), "tflops": round(2.0 * m * n * k / dt / 1e12, 2),
                      "checksum": float(c[0, 0])}), flush=True)


def refcrit_leg(phi_h, b_h, nfields):
    """The reference's criterion in FP64: per field, CG on that field's 8 RHS
    with x0 = b, one alpha/beta for the batch and sum |r|^2 <= 1e-24 on the
    recursive residual (CG.tpp:14-113, CG.hpp:34) -- the reference solves one
    HubbardExp at a time.  Reports the iterations per field, solves/s over the
    fields run, and the true residual the recursive stop leaves (FP64 and
    double-double)."""
    import torch
    import nsl_amd as nsl
    lat = nsl.Lattice.honeycomb(L, L)
    out = {"fields": nfields, "iterations": [], "true_rel_residual_max": 0.0}
    t_tot = 0.0
    for f in range(nfields):
        fm = nsl.HubbardExp(lat, NT, BETA)
        fm.populate(torch.as_tensor(phi_h[f]).cuda())
        b = torch.as_tensor(b_h[f * NRHS:(f + 1) * NRHS]).cuda()
        cg = nsl.CG(fm, nsl.MMdagger, TOL, 200000)           # TOL_ABS_BATCH: the reference's stop
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = cg(b)
        torch.cuda.synchronize()
        t_tot += time.perf_counter() - t0
        out["iterations"].append(cg.report["iterations"])
        rel = nsl.dd_residual(fm, b, x)
        out["true_rel_residual_max"] = max(out["true_rel_residual_max"], float(rel.max()))
        out.setdefault("converged", []).append(bool(cg.converged))
    out["solves_per_s"] = round(nfields * NRHS / t_tot, 3)
    out["note"] = ("recursive sum|r|^2 <= 1e-24 per 8-RHS batch; the true residual it leaves is the FP64 floor, "
                   "not 1e-12")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--refcrit-fields", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-refcrit", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-hmc", action="store_true")
    ap.add_argument("--hops", type=int, default=0)
    ap.add_argument("--tss", type=int, default=0)
    ap.add_argument("--precond", choices=["f64", "f32"], default="f64",
                    help="precision of the time-slice preconditioner of the headline (and of --tss)")
    ap.add_argument("--copy-calib", type=int, default=0)
    ap.add_argument("--dgemm-ceiling", default="")
    ap.add_argument("--prof-every", type=int, default=PROF_EVERY,
                    help="time every n-th launch of a kernel class with HIP events (0: none)")
    args = ap.parse_args()
    if args.copy_calib:
        return copy_calib(args.copy_calib)
    if args.dgemm_ceiling:
        return dgemm_ceiling(args.dgemm_ceiling)
    if args.hops:
        return hop_profile(args.hops)
    if args.tss:
        return tss_profile(args.tss, args.precond)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the path would shard by whole requests/fields with no communication",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import nsl_amd as nsl
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    phi_h, b_h = fields()
    Nx = 2 * L * L
    nb = NFIELD * NRHS
    lat = nsl.Lattice.honeycomb(L, L)
    fm = nsl.HubbardExp(lat, NT, BETA)
    phi = torch.as_tensor(phi_h).cuda()
    fm.populate(phi)
    b = torch.as_tensor(b_h).cuda()
    x = torch.empty_like(b)
    L_ = nsl._lib.lib()

    def run(cg, k):
        """k steps: populate (invalidates any factorisation), then the solve."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fm.populate(phi)
            cg(b, out=x)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def leg(cg, steps, warmup, prof):
        run(cg, warmup)
        if prof:
            L_.nslx_prof_reset()
            L_.nslx_prof_enable(args.prof_every)
        t = run(cg, steps)
        text = ""
        if prof:
            L_.nslx_prof_enable(0)
            text = prof_text_of(L_)
        rel = nsl.dd_residual(fm, b, x, getattr(cg, "x_lo", None))
        return t, dict(cg.report), bool(cg.converged), float(rel.max()), text

    # ---- headline: CG preconditioned by the time-slice solve (FP64 by
    # default: its wider chunks give a near-exact preconditioner with a
    # quarter of the factorisation's block steps), double-double correction
    pprec = nsl.F64 if args.precond == "f64" else nsl.F32
    cg_mix = nsl.CGTimeSlice(fm, TOL, 5000, precision=pprec)
    t_mix, rep_mix, conv_mix, rel_mix, prof_text = leg(cg_mix, args.steps, args.warmup, args.prof_every > 0)
    print(f"# headline ({args.precond} preconditioner): {t_mix / args.steps * 1e3:.1f} ms/step, report {rep_mix}, "
          f"dd residual {rel_mix:.3e}", file=sys.stderr)
    chunk_rule = ("chunks of the largest divisor c of Nt with c g <= 12 (c = 32 here: 2 chunks)" if pprec == nsl.F64
                  else "chunks of the largest divisor c of Nt with c g <= 3 (c = 8 here: 8 chunks)")
    print("# kernel timing (name launches total_ms timed flops bytes):\n# " +
          prof_text.strip().replace("\n", "\n# "), file=sys.stderr)
    out = {"metric": METRIC, "value": round(nb * args.steps / t_mix, 3), "unit": "solves/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(t_mix / args.steps * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64+dd" if pprec == nsl.F64 else "f32+f64+dd",
           "data": "synthetic (Philox-seeded phi ~ N(0, sqrt(U~)), Gaussian RHS)",
           "config": {"workload": "precision form: honeycomb 12x12 (Nx=288), Nt=64, beta=8, U=3, "
                                  "16 fields x 8 RHS, CG on MMdagger to 1e-12 true residual per RHS "
                                  "(double-double), populate + solve per step",
                      "Nx": Nx, "Nt": NT, "beta": BETA, "U": U, "fields": NFIELD,
                      "rhs_per_field": NRHS, "tol": TOL,
                      "method": f"CG, time-slice preconditioner {args.precond.upper()} (factorised per step; "
                                f"{chunk_rule}, g = delta ||K|| = 0.375), FP64 recurrence, "
                                "double-double defect correction"},
           "converged": conv_mix, "max_rel_residual": rel_mix, "reached_tol_all": rel_mix <= TOL,
           "iterations": rep_mix["iterations"], "outer_steps": rep_mix["outer_steps"],
           "preconditioner_applies": rep_mix["iterations_total"]}
    rf, rows = roofline_from_prof(prof_text, nb, Nx, NT, {})
    if rf:
        rf["traffic"] = read_traffic(rf["kernel"])
    out["roofline"] = rf
    out["kernels"] = kernel_table(rows)

    if not args.no_fp64:
        # the plain FP64 path: unpreconditioned CG, double-double solution
        cg_64 = nsl.CG(fm, nsl.MMdagger, TOL, 400000, double_double=True)
        t64, rep64, conv64, rel64, _ = leg(cg_64, args.fp64_steps, 1, False)
        v64 = nb * args.fp64_steps / t64
        out["fp64_solves_per_s"] = round(v64, 3)
        out["fp64_ms_per_step"] = round(t64 / args.fp64_steps * 1e3, 3)
        out["fp64_iterations"] = rep64["iterations"]
        out["fp64_max_rel_residual"] = rel64
        out["fp64_reached_tol_all"] = rel64 <= TOL and conv64
        out["speedup_vs_fp64"] = round(out["value"] / v64, 3)
        print(f"# fp64: {t64 / args.fp64_steps * 1e3:.1f} ms/step, report {rep64}", file=sys.stderr)
        # the same preconditioned solver with the time-slice solve in the
        # other precision
        alt = "f32" if pprec == nsl.F64 else "f64"
        cg_alt = nsl.CGTimeSlice(fm, TOL, 5000, precision=nsl.F32 if alt == "f32" else nsl.F64)
        tt, rept, convt, relt, _ = leg(cg_alt, args.steps, 1, False)
        out[f"{alt}_tss"] = {"solves_per_s": round(nb * args.steps / tt, 3),
                             "ms_per_step": round(tt / args.steps * 1e3, 3), "iterations": rept["iterations"],
                             "max_rel_residual": relt, "reached_tol_all": relt <= TOL and convt,
                             "headline_over_this": round(out["value"] / (nb * args.steps / tt), 3)}
    if not args.no_refcrit:
        out["refcrit"] = refcrit_leg(phi_h, b_h, args.refcrit_fields)
    if not args.no_hmc:
        out["hmc"] = hmc_leg()
        out["bicgstab"] = bicgstab_leg(cpu=not args.no_cpu)
    if not args.no_batch:
        out["batching"] = batching_leg()
    if not args.no_serve:
        # the full stream through the time-slice-preconditioned server (FP64
        # and FP32 preconditioner); the plain FP64 path (unpreconditioned CG
        # in the Krylov pool) only on the small requests it finishes in
        # bounded time, beside the preconditioned server on the same subset
        small = lambda nx, nt: nx <= 96 and nt <= 32
        sv_64 = serving_leg(nsl.F64, workers=4)
        sv_32 = serving_leg(nsl.F32, workers=4)
        sub_tss = serving_leg(nsl.F64, workers=4, subset=small)
        sub_k64 = serving_leg(nsl.F64, workers=0, subset=small)
        out["p50_ms"] = sv_64["p50_ms"]
        out["p99_ms"] = sv_64["p99_ms"]
        out["serving"] = {"tss_f64": sv_64, "tss_f32": sv_32, "reported": "tss_f64",
                          "subset_nx_le_96_nt_le_32": {"tss_f64": sub_tss, "plain_fp64_krylov": sub_k64,
                                                       "solves_per_s_ratio": round(sub_tss["solves_per_s"] /
                                                                                   sub_k64["solves_per_s"], 2)},
                          "workload": f"{SERVE_N} correlator requests drawn like the batching mix (honeycomb Nx "
                                      "18-1152, (6,6) nanotubes Nx 96-768, Nt 16-128, beta 2-12, U 1-4), Poisson "
                                      f"{SERVE_RATE}/s, Nx point sources each, per-RHS 1e-10, 4 worker streams"}
    else:
        out["p50_ms"] = None
        out["p99_ms"] = None
    if not args.no_cpu:
        cores = os.cpu_count() or 1
        iters = (out.get("refcrit") or {}).get("iterations", [None])[0] or out.get("fp64_iterations")
        out["cpu_baseline"] = cpu_baseline(phi_h, b_h, cores, iters)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
