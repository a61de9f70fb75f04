#!/usr/bin/env python3
"""Throughput of the explicit residual-and-update loop on one MI355X.

Workload (default): FP64 Navier-Stokes flow past a cylinder, O-grid of
2048 azimuthal x 1024 radial cells (geometric stretching 1.006, R=0.5 to 30),
each quad split into 2 triangles along a seeded diagonal -> 4,194,304 cells;
M=0.2, Re=100, Rusanov flux, Venkatakrishnan limiter (K = 1; --limiter scalar
for the reference's own limiter, solve.d:360-403), TVD-RK2 (2 stages per step),
CFL 0.05.  (The reference's CFL time step, solve.d:1092-1094, has no viscous
limit; with 3.9e-4 wall spacing and nu = 0.01 it diverges within three steps
at CFL 0.5 (the CPU restatement shows the same) and within a hundred steps
at CFL 0.1 -- so the case runs at CFL 0.05.)  A step is one full pass of the reference's time loop
(app/solver.d:119-250): both stages plus RMS residuals, forces, dt/CFL update.

Metric: Mcell-RK-stage updates/s = interior cells * stages * steps / wall time.
Printed as one JSON line; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per interior cell-stage (FP64, triangles): the compulsory
# traffic of one residual + stage update, each datum once, intermediates such as
# gradients not counted -- state in 32 + q^n 32 + new state 32, neighbour and face
# ids 24, least-squares weights 48, centroid/area/d/perimeter 40, 1.5 faces of
# geometry 60.  The tiled kernels finish a whole stage for their cells, so one
# launch processes N cell-stages.
STAGE_BYTES_PER_CELL = 268.0
# split path: its face kernel's own unit
FACE_BYTES_PER_FACE = 88.0      # owner/other ids 8 + normal 16 + midpoint 16 + length 8 + len*flux,len*sMax 40
FACE_BYTES_PER_CELL = 144.0     # state 32 + gradient 64 + limiter 32 + centroid 16, each cell read once
# HBM bytes per launch of the face/update tile kernel on this workload, from
# rocprofv3 FETCH_SIZE and WRITE_SIZE scaled by a plain 1 GiB copy (x2.0 and
# x1.0), averaged over the limited and unlimited stage; the summary and the
# commands are in profiles/hbm_traffic_cylinder_ns_4M.md.  Re-measured whenever
# the kernel changes; reported only for the configurations it was measured on
# (kernel, workload, stage path, tile cells, tile order, cell order, limiter).
MEASURED_TRAFFIC = {
    ("k_face_tile", "cylinder_ns_ogrid_2048x1024x2", "faceTiles", -1, "blocks", "reference", "forward",
     "venkatakrishnan"): 1.3695e9,
    ("k_face_tile", "cylinder_ns_ogrid_2048x1024x2", "faceTiles", -1, "blocks", "reference", "forward",
     "scalar"): 1.3674e9,
    ("k_face_tile", "cylinder_ns_ogrid_2048x1024x2", "faceTiles", -1, "blocks", "tiles", "tilesBackward",
     "venkatakrishnan"): 1.3687e9,
}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cylinder_config(M=0.2, CFL=0.05):
    q = [M, 0.0, 1.0, 1.0]
    return {
        "mesh": "cylinder_ogrid.msh", "dt": 1.0e-6, "tEnd": 1.0e9, "flux": "rusanovFlux", "plotIter": -1,
        "saveIter": -1, "CFL": CFL, "integrator": "RK2_TVD", "forceBoundary": "Wall", "limited": True,
        "order": 2, "initialConditions": q, "viscosity": True, "cflAdjust": True,
        "physicalConfig": {"Pr": 0.7, "R": 8.3145, "gamma": 1.4, "Re": 100.0, "L": 1.0},
        "boudaryConditions": [{"tag": "Wall", "type": "viscidWall", "q": q},
                              {"tag": "Farfield", "type": "fullState", "q": q}],
    }


def cpu_baseline(raw, cfg, steps, n_cells, threads):
    """The oracle (C restatement of the reference loop) on `threads` host threads: its
    per-cell and per-edge loops are OpenMP-parallel and give the single-thread bits."""
    import oracle
    oracle.build()
    oracle.set_threads(threads)
    m = oracle.Mesh(raw)
    c = dict(cfg)
    m.setup(c)
    t0 = time.perf_counter()
    m.run(steps)
    el = time.perf_counter() - t0
    oracle.set_threads(1)
    return n_cells * 2 * steps / el / 1e6, el


def state_error(q64, q32):
    """Per-cell state error of a mixed-precision run against the FP64 one: for each variable
    the largest |difference| over the cells divided by that variable's largest |FP64 value|
    (a near-zero component, such as v in the far field, would make a per-cell ratio
    meaningless), then the worst variable; and the same in the 2-norm."""
    d = np.abs(q64 - q32)
    scale = np.abs(q64).max(axis=0)
    mx = float((d.max(axis=0) / scale).max())
    l2 = float((np.sqrt((d * d).sum(axis=0)) / np.sqrt((q64 * q64).sum(axis=0))).max())
    return mx, l2


def mixed_leg(mesh, cfg, args, N, value64):
    """The same workload in mixed precision (FP32 flux, reconstruction and limiter; FP64
    geometry, accumulation and norms), timed like the FP64 line.  Its error is measured on
    the workload's full run: `--c2-steps` steps (5000) of the FP64 and the mixed solver from
    the same initial state (the FP64 path equals the oracle bit for bit), comparing the
    states, drag and lift."""
    from ebbcfd_amd import Solver
    c = dict(cfg)
    c["extensions"] = dict(cfg["extensions"], precision="mixed")
    s = Solver(mesh, c)
    s.run(args.warmup)
    t0 = time.perf_counter()
    hist, _, _ = s.run(args.steps)
    el = time.perf_counter() - t0
    assert len(hist) == args.steps and np.all(np.isfinite(hist[:, 2:9]))
    ms = el / args.steps * 1e3
    v = N * 2 * args.steps / el / 1e6
    tk = s.time_steps(args.time_steps)
    del s
    log("mixed: %.3f ms/step, %.1f Mcell-stage/s (x%.3f of FP64); tile %.3f ms, gradient %.3f ms"
        % (ms, v, v / value64, tk["fused_ms"], tk["gradient_ms"]))
    out = {"value": round(v, 3), "unit": "Mcell-stage/s", "ms_per_step": round(ms, 4), "dtype": "mixed",
           "speedup_vs_fp64": round(v / value64, 4),
           "kernels_ms": {"tile": round(tk["fused_ms"], 4), "gradient": round(tk["gradient_ms"], 4)},
           "precision": "FP32 flux/reconstruction/gradient/limiter; FP64 geometry, residual sum, update, norms"}
    if args.c2_steps > 0:
        n = args.c2_steps
        runs = {}
        for prec, cc in (("fp64", cfg), ("mixed", c)):
            x = Solver(mesh, cc)
            t0 = time.perf_counter()
            h, t_end, _ = x.run(n)
            runs[prec] = (h, x.q, time.perf_counter() - t0, t_end)
            del x
        (h64, q64, s64, t64), (h32, q32, s32, t32) = runs["fp64"], runs["mixed"]
        assert len(h64) == n and len(h32) == n, "a full-length run stopped early"
        mx, l2 = state_error(q64, q32)
        drag = abs(h64[-1, 7] - h32[-1, 7]) / abs(h64[-1, 7])
        lift = abs(h64[-1, 8] - h32[-1, 8]) / max(abs(h64[-1, 8]), abs(h64[-1, 7]))
        drag_hist = float(np.max(np.abs(h64[:, 7] - h32[:, 7])) / np.max(np.abs(h64[:, 7])))
        log("C2/C3 full runs: %d steps to t = %.4g, FP64 %.2f s, mixed %.2f s; mixed vs FP64: state %.3g (max), "
            "%.3g (2-norm), drag %.3g, lift %.3g (of drag), drag over the history %.3g"
            % (n, t64, s64, s32, mx, l2, drag, lift, drag_hist))
        out["error_vs_fp64"] = {"steps": n, "t_end": t64, "max_rel_state": mx, "l2_rel_state": l2,
                                "drag_rel": drag, "lift_rel_of_drag": lift, "drag_rel_max_over_history": drag_hist,
                                "drag_fp64": h64[-1, 7], "lift_fp64": h64[-1, 8],
                                "strouhal": "not reached: %d steps cover t = %.3g, a shedding period is ~30" % (n, t64)}
        out["full_run_s"] = {"fp64": round(s64, 3), "mixed": round(s32, 3)}
    return out


def bump_request(rng, cells, M_range, seed, precision="mixed", **ext):
    """One channel-bump Euler solve: bump height U(0.02, 0.1) of a sin^2 bump on the lower
    wall, `cells` jittered triangles, M and AoA from the given ranges, Roe flux, scalar
    (Barth-Jespersen-type) limiter, RK4, inviscid walls, full-state inflow/outflow."""
    from ebbcfd_amd import meshgen
    nx, ny = meshgen.channel_bump_cells(cells)
    raw = meshgen.channel_bump(nx, ny, float(rng.uniform(0.02, 0.1)), seed=seed)
    q = [float(rng.uniform(*M_range)), float(rng.uniform(-5.0, 5.0)), 1.0, 1.0]
    bcs = [{"tag": t, "type": "fullState" if t in ("Inflow", "Outflow") else "inviscidWall", "q": q}
           for t in ("Inflow", "Outflow", "Lower", "Upper")]
    cfg = {"mesh": "bump.msh", "dt": 1.0e-5, "tEnd": 1.0e9, "flux": "roeFlux", "plotIter": -1, "saveIter": -1,
           "CFL": 0.5, "integrator": "RK4", "forceBoundary": "Lower", "limited": True, "order": 2,
           "initialConditions": q, "viscosity": False, "cflAdjust": True,
           "physicalConfig": {"Pr": 0.7, "R": 8.3145, "gamma": 1.4, "Re": 1000.0, "L": 1.0},
           "boudaryConditions": bcs, "extensions": dict(precision=precision, **ext)}
    return raw, cfg


def batched_leg(args):
    """Batched Euler: `--batched-meshes` channel-bump meshes, cell counts log-uniform in
    [5e3, 5e5], M U(0.3, 2.0), AoA U(-5, 5) deg, mixed precision, Roe + scalar limiter, RK4;
    every mesh advanced `--batched-steps` steps by one launch per kernel per RK stage
    (HIP-graph replays; tile size and graph length from the autotuner), against the same
    solves run one at a time through the solo path."""
    from ebbcfd_amd import Batch, Solver, build_mesh
    rng = np.random.Generator(np.random.PCG64(args.batched_seed))
    n, K = args.batched_meshes, args.batched_steps
    cells = np.exp(rng.uniform(np.log(5e3), np.log(5e5), n))
    t0 = time.perf_counter()
    reqs = [bump_request(rng, int(c), (0.3, 2.0), args.batched_seed * 1000 + k) for k, c in enumerate(cells)]
    meshes = [build_mesh(raw) for raw, _ in reqs]
    total = sum(m.n_interior for m in meshes)
    log("C4: %d meshes, %d cells in all, built in %.1f s" % (n, total, time.perf_counter() - t0))
    b = Batch(graph_steps=8)
    t0 = time.perf_counter()
    ids = [b.add(m, cfg) for m, (_, cfg) in zip(meshes, reqs)]
    log("C4: admitted (ufvmSetup + upload) in %.1f s, %.2f GB on the device" % (time.perf_counter() - t0,
                                                                              b.info()["device_bytes"] / 1e9))
    b.run(8)
    tune = b.tune(steps=8) if args.batched_tune else None
    if tune:
        log("C4 autotune: tile cells %d, graph steps %d: %.3f ms/step (default %.3f)"
            % (tune["tile_cells"], tune["graph_steps"], tune["ms_per_step"], tune["default_ms_per_step"]))
    b.run(16)
    t0 = time.perf_counter()
    running = b.run(K)
    el = time.perf_counter() - t0
    stopped = n - running
    batched_sps = n / el
    # one at a time through the solo path, on a subsample of the same solves
    sample = list(range(0, n, max(1, n // args.batched_solo_sample)))
    solo_el = 0.0
    for k in sample:
        s = Solver(meshes[k], reqs[k][1])
        s.run(8)
        t0 = time.perf_counter()
        s.run(K)
        solo_el += time.perf_counter() - t0
        del s
    solo_sps = len(sample) / solo_el
    log("C4: batched %.1f solves/s (%d steps each, %.2f ms/step for all %d), one at a time %.1f solves/s "
        "(%d-solve sample): x%.2f" % (batched_sps, K, el / K * 1e3, n, solo_sps, len(sample), batched_sps / solo_sps))
    del b
    return {"workload": "channel_bump_batch_%d_loguniform_5e3_5e5_rk4_mixed" % n, "meshes": n, "cells": int(total),
            "steps_per_solve": K, "batched_solves_per_s": round(batched_sps, 2), "ms_per_step_all": round(el / K * 1e3, 4),
            "Mcell_stage_per_s": round(total * 4 * K / el / 1e6, 1), "solo_solves_per_s": round(solo_sps, 2),
            "solo_sample": len(sample), "speedup": round(batched_sps / solo_sps, 3), "stopped_early": int(stopped),
            "tune": tune, "seed": args.batched_seed}


def serve_bench(args):
    """Serving: Poisson arrivals of steady channel-bump Euler solves (sizes log-uniform in
    [1e4, 2e6] cells, M U(0.3, 1.5), local time steps, Venkatakrishnan limiter K = 5, FP64),
    each run to a 1e-8 drop of the rho RMS residual or the step cap; continuous batching with
    admission by device footprint; p50/p99 time to converged and solves/s per arrival rate,
    beside the same stream served one at a time."""
    from ebbcfd_amd import Server, build_mesh
    rng = np.random.Generator(np.random.PCG64(args.serve_seed))
    n = args.serve_requests
    cells = np.exp(rng.uniform(np.log(1e4), np.log(2e6), n))
    t0 = time.perf_counter()
    ext = {"rmsDrop": 1e-8, "maxSteps": args.serve_max_steps, "limiter": args.serve_limiter}
    if args.serve_limiter == "venkatakrishnan":
        ext["venkatK"] = args.serve_venkat_k
    reqs = [bump_request(rng, int(c), (0.3, 1.5), args.serve_seed * 1000 + k, precision=args.serve_precision, **ext)
            for k, c in enumerate(cells)]
    for _, cfg in reqs:
        cfg["localTimestep"] = True
    meshes = [build_mesh(raw) for raw, _ in reqs]
    log("serve: %d requests, %d cells in all, built in %.1f s" % (n, sum(m.n_interior for m in meshes),
                                                               time.perf_counter() - t0))
    rows = []
    rates = [float(r) for r in args.serve_rates.split(",")]
    for rate in rates:
        for mode, max_solves in (("batched", 256), ("one_at_a_time", 1)):
            # one at a time is compared where both are saturated: the highest rate
            if mode == "one_at_a_time" and (rate != max(rates) or not args.serve_solo):
                continue
            sv = Server(max_solves=max_solves, steps_per_round=args.serve_round, graph_steps=8)
            arr = np.cumsum(rng.exponential(1.0 / rate, n))
            t0 = time.perf_counter()
            for k in range(n):
                d = arr[k] - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                sv.submit(meshes[k], reqs[k][1])
            ok = False
            t_wait = time.perf_counter()
            while time.perf_counter() - t_wait < 3000.0:
                ok = sv.wait(n, 30.0)
                if ok:
                    break
                log("serve %s rate %g/s: %d of %d done" % (mode, rate, int(sv.stats()["done"]), n))
            st = sv.stats()
            polls = [sv.poll(k) for k in range(n)]
            sv.close()
            conv = sum(1 for p in polls if p["converged"])
            steps = np.array([p["steps"] for p in polls])
            wait = np.array([p["admit_s"] - p["arrival_s"] for p in polls])
            row = {"rate_per_s": rate, "mode": mode, "done": int(st["done"]), "failed": int(st["failed"]),
                   "p50_s": round(st["p50_s"], 3), "p99_s": round(st["p99_s"], 3),
                   "solves_per_s": round(st["solves_per_s"], 3), "converged": conv,
                   "steps_min_median_max": [int(steps.min()), int(np.median(steps)), int(steps.max())],
                   "admission_wait_p50_s": round(float(np.median(wait)), 3), "complete": bool(ok)}
            log("serve %s rate %g/s: %s" % (mode, rate, row))
            rows.append(row)
    out = {"metric": "batched solves/s + p99 latency (serving)", "unit": "solves/s",
           "config": {"workload": "channel_bump_poisson_loguniform_1e4_2e6", "requests": n, "max_steps": args.serve_max_steps,
                      "rms_drop": 1e-8, "precision": args.serve_precision, "limiter": args.serve_limiter,
                      "venkatK": args.serve_venkat_k if args.serve_limiter == "venkatakrishnan" else None,
                      "steps_per_round": args.serve_round, "seed": args.serve_seed},
           "rows": rows}
    print(json.dumps(out), flush=True)


def dropin_leg(s, N, stages, ms_per_stage_loop):
    """The drop-in at the reference's operator boundary (source/integrators.d calling
    ufvmSolver once per stage): per stage, the caller's state goes to the device
    (ebb_state_upload), ebb_residual runs and R comes back to the host -- the PCIe-inclusive
    rate of that use.  Host arrays reused across stages: the caller's pageable ones (numpy),
    then page-locked ones from ebb_host_alloc (pinned_array)."""
    from ebbcfd_amd import pinned_array
    q = s.q
    R = np.zeros((N, 4))
    s.ufvm_solver(out=R)

    def per_stage(qbuf, Rbuf):
        t0 = time.perf_counter()
        for _ in range(stages):
            s.q = qbuf
            s.ufvm_solver(out=Rbuf)
        return (time.perf_counter() - t0) / stages

    el = per_stage(q, R)
    qp, Rp = pinned_array((N, 4)), pinned_array((N, 4))
    qp[...] = q
    s.ufvm_solver(out=Rp)
    elp = per_stage(qp, Rp)
    assert np.array_equal(Rp, R)
    del qp, Rp
    log("drop-in at ufvmSolver (upload q + residual + download R): %.2f ms per stage from pageable host arrays, "
        "%.2f ms from page-locked ones, vs %.3f ms per stage inside the device loop"
        % (el * 1e3, elp * 1e3, ms_per_stage_loop))
    return {"ms_per_stage": round(el * 1e3, 3), "Mcell_stage_per_s": round(N / el / 1e6, 1),
            "pinned_ms_per_stage": round(elp * 1e3, 3), "pinned_Mcell_stage_per_s": round(N / elp / 1e6, 1),
            "bytes_over_pcie_per_stage": int(N * 64), "stages": stages,
            "loop_ms_per_stage": round(ms_per_stage_loop, 4)}


def c1_leg(steps):
    """Latency of the manufactured-solution Euler workload: 8,192 jittered triangles,
    Roe, scalar limiter, RK4, CFL 0.5 -- cache-resident, reported for latency."""
    from ebbcfd_amd import Batch, Solver, build_mesh, meshgen
    raw = meshgen.mms_square(64, 0.1, 1)
    q = [0.1, 0.0, 1.1, 1.0]
    cfg = {"mesh": "mms.msh", "dt": 1.0e-5, "tEnd": 1.0e9, "flux": "roeFlux", "plotIter": -1, "saveIter": -1,
           "CFL": 0.5, "integrator": "RK4", "forceBoundary": "Wall", "limited": True, "order": 2,
           "initialConditions": q, "viscosity": False,
           "physicalConfig": {"Pr": 0.71, "R": 1.0, "gamma": 1.4, "Re": 1000.0, "L": 1.0},
           "boudaryConditions": [{"tag": "Freestream", "type": "dirichlet", "q": q}], "extensions": {"mms": True}}
    mesh = build_mesh(raw)
    s = Solver(mesh, cfg)
    s.run(50)
    t0 = time.perf_counter()
    h, _, _ = s.run(steps)
    el = time.perf_counter() - t0
    assert len(h) == steps
    tk = s.time_steps(50)
    # the same solve replayed as 8-step HIP graphs (a one-member batch runs the same kernels)
    b = Batch(graph_steps=8)
    b.add(mesh, cfg)
    b.run(56)
    t0 = time.perf_counter()
    b.run(steps)
    el_graph = time.perf_counter() - t0
    del b
    log("C1: %d RK4 steps of 8192 cells: %.2f us/step (stage kernels %.2f us/step; replayed as 8-step graphs "
        "%.2f us/step)" % (steps, el / steps * 1e6, tk["step_ms"] * 1e3, el_graph / steps * 1e6))
    return {"workload": "mms_euler_64x64x2_jittered_rk4", "cells": 8192, "steps": steps,
            "us_per_step": round(el / steps * 1e6, 2), "stage_kernels_us_per_step": round(tk["step_ms"] * 1e3, 2),
            "graph_replay_us_per_step": round(el_graph / steps * 1e6, 2),
            "Mcell_stage_per_s": round(8192 * 4 * steps / el / 1e6, 2)}


def copy_calibration():
    """A plain device-to-device copy of 2^27 doubles (1 GiB read + 1 GiB written), five times."""
    import torch
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    log("copied %d bytes 5 times" % (x.numel() * 8))


def shedding_frequency(hist, t_from):
    """Strouhal number (D = U = 1) and peak-to-peak lift of a lift history after t_from:
    the mean period between upward zero crossings of (lift - mean), each crossing time
    interpolated linearly between its two samples."""
    t, lift = hist[:, 0], hist[:, 8]
    m = t > t_from
    tt, ll = t[m], lift[m] - lift[m].mean()
    k = np.where((ll[:-1] < 0) & (ll[1:] >= 0))[0]
    if len(k) < 3:
        return None, None, float(ll.max() - ll.min()) if len(ll) else None
    tc = tt[k] - ll[k] * (tt[k + 1] - tt[k]) / (ll[k + 1] - ll[k])
    per = np.diff(tc)
    return float(1.0 / per.mean()), per, float(ll.max() - ll.min())


def strouhal_bench(args):
    """Vortex shedding, FP64 against mixed precision: the workload's physics (Re = 100,
    M = 0.2, Rusanov, Venkatakrishnan K = 1, TVD-RK2) on a 256 x 128 O-grid with radial
    stretch 1.04 (wall spacing 7.8e-3), where the reference's CFL time step is stable at
    CFL 0.5 and `--strouhal-steps` steps reach t ~ 100, some ten shedding periods.  CFL is
    held at 0.5 (no cflAdjust), so both runs take the same time steps up to rounding."""
    from ebbcfd_amd import Solver, build_mesh, meshgen
    if args.strouhal_steps <= 0:
        sys.exit("--strouhal-steps must be positive")
    raw = meshgen.cylinder_ogrid(256, 128, seed=1, stretch=1.04)
    mesh = build_mesh(raw)
    out = {}
    hists = {}
    for prec in ("fp64", "mixed"):
        cfg = cylinder_config(CFL=0.5)
        cfg["cflAdjust"] = False
        cfg["extensions"] = {"limiter": "venkatakrishnan", "venkatK": 1.0, "precision": prec}
        s = Solver(mesh, cfg)
        parts, done, t0 = [], 0, time.perf_counter()
        while done < args.strouhal_steps:
            h, t, dt = s.run(min(50000, args.strouhal_steps - done))
            parts.append(h)
            done += len(h)
            log("%s: %d steps, t = %.2f, lift %.5f, drag %.5f (%.1f s)" % (prec, done, t, h[-1, 8], h[-1, 7],
                                                                          time.perf_counter() - t0))
        hists[prec] = np.concatenate(parts)
        St, per, amp = shedding_frequency(hists[prec], args.strouhal_from)
        out[prec] = {"St": St, "lift_peak_to_peak": amp, "t_end": float(t), "seconds": round(time.perf_counter() - t0, 1),
                     "periods": None if per is None else [round(float(x), 6) for x in per]}
        del s
    a, b = hists["fp64"], hists["mixed"]
    n = min(len(a), len(b))
    dl = np.abs(a[:n, 8] - b[:n, 8]).max() / np.abs(a[:n, 8]).max()
    dd = np.abs(a[:n, 7] - b[:n, 7]).max() / np.abs(a[:n, 7]).max()
    dt_rel = np.abs(a[:n, 0] - b[:n, 0]).max() / a[n - 1, 0]
    err = None
    if out["fp64"]["St"] and out["mixed"]["St"]:
        err = abs(out["mixed"]["St"] - out["fp64"]["St"]) / out["fp64"]["St"]
    log("St: FP64 %s, mixed %s, relative error %s; lift peak-to-peak %s vs %s"
        % (out["fp64"]["St"], out["mixed"]["St"], err, out["fp64"]["lift_peak_to_peak"], out["mixed"]["lift_peak_to_peak"]))
    print(json.dumps({"metric": "Strouhal number, mixed vs FP64", "steps": args.strouhal_steps,
                      "config": {"workload": "cylinder_ns_ogrid_256x128x2_stretch1.04", "Re": 100, "M": 0.2, "CFL": 0.5,
                                 "cflAdjust": False, "limiter": "venkatakrishnan (K=1)", "St_from_t": args.strouhal_from},
                      "fp64": out["fp64"], "mixed": out["mixed"], "St_rel_error": err,
                      "lift_max_diff_rel": float(dl), "drag_max_diff_rel": float(dd), "time_max_diff_rel": float(dt_rel)}),
          flush=True)


def ramp_probe(args):
    """Short timed windows after GPU idle: the workload's solver, `--warmup` steps, then for
    each idle gap a `--steps`-step run after the gap and another right after it.  The
    headline's timed region starts `--warmup` steps after seconds of host-only mesh build,
    so a short run shows how much of it falls in the GPU's clock ramp."""
    from ebbcfd_amd import Solver, build_mesh, meshgen
    raw = meshgen.cylinder_ogrid(args.n_theta, args.n_r, seed=1)
    cfg = cylinder_config()
    cfg["extensions"] = {"stagePath": args.stage_path, "limiter": args.limiter, "venkatK": args.venkat_k}
    s = Solver(build_mesh(raw), cfg)
    K = args.steps

    def timed(x=None):
        x = x or s
        t0 = time.perf_counter()
        x.run(K)
        return round((time.perf_counter() - t0) / K * 1e3, 4)

    s.run(args.warmup)
    rows = [{"idle_s": "mesh build", "first_ms_per_step": timed(), "next_ms_per_step": timed()}]
    s.run(100)
    rows.append({"idle_s": 0.0, "first_ms_per_step": timed(), "next_ms_per_step": timed()})
    for idle in (0.01, 0.1, 0.5, 3.0):
        time.sleep(idle)
        rows.append({"idle_s": idle, "first_ms_per_step": timed(), "next_ms_per_step": timed()})
    for r in rows:
        log("after %s idle: %.4f ms/step, then %.4f" % (r["idle_s"], r["first_ms_per_step"], r["next_ms_per_step"]))
    # the fixed cost of one run call: back-to-back runs of 1-64 steps (minimum of three each)
    s.run(100)
    sweep = {}
    for n in (1, 2, 4, 8, 16, 32, 64):
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            s.run(n)
            ts.append(time.perf_counter() - t0)
        sweep[n] = round(min(ts) * 1e3, 4)
    x = np.array(list(sweep), float)
    slope, icpt = np.polyfit(x, np.array(list(sweep.values())), 1)
    log("runs of 1-64 steps: %s ms; fit %.4f ms per call + %.4f ms per step" % (sweep, icpt, slope))
    # a second solver, set up right after the first one's steps (the GPU idles meanwhile)
    s.run(20)
    s2 = Solver(s.mesh, cfg)
    s2.run(args.warmup)
    fresh = {"first_ms_per_step": timed(s2)}
    fresh["next_ms_per_step"] = timed(s2)
    log("fresh solver after %d warm-up steps: %.4f ms/step, then %.4f"
        % (args.warmup, fresh["first_ms_per_step"], fresh["next_ms_per_step"]))
    print(json.dumps({"probe": "clock ramp", "steps_per_window": K, "warmup": args.warmup, "rows": rows,
                      "run_ms_by_steps": sweep, "fit_ms_per_call": round(float(icpt), 4),
                      "fit_ms_per_step": round(float(slope), 4), "second_solver": fresh}), flush=True)


def host_threads():
    """CPUs the mesh build may use: affinity set, capped by OMP_NUM_THREADS and 32."""
    t = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and 0 < int(omp) < t:
        t = int(omp)
    return min(t, 32)


def ingest_bench(args):
    """Mesh ingest: gmsh 2.2 read + geometry precompute (importMesh + buildMesh,
    source/io.d:689-990, source/mesh.d:217-643) of the benchmark's O-grid through
    ebb_mesh_load, timed from the file on local disk (page cache warm after the write)."""
    import tempfile
    from ebbcfd_amd import import_mesh, meshgen
    log("generating O-grid %d x %d" % (args.n_theta, args.n_r))
    raw = meshgen.cylinder_ogrid(args.n_theta, args.n_r, seed=1)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cylinder.msh")
        meshgen.write_gmsh(raw, path)
        size = os.path.getsize(path)
        log("wrote %s (%.1f MB)" % (path, size / 1e6))
        del raw
        runs = []
        for k in range(args.ingest_repeats):
            t0 = time.perf_counter()
            m = import_mesh(path)
            runs.append(time.perf_counter() - t0)
            N, F = m.n_interior, m.n_edges
            del m
            log("ingest %d: %.3f s" % (k, runs[-1]))
    el = float(np.median(runs))
    cpu = None
    if args.cpu_steps > 0:
        import oracle
        from oracle.gmsh import parse_gmsh
        nt, nr = args.n_theta // 2, args.n_r // 2
        sample = meshgen.cylinder_ogrid(nt, nr, seed=1)
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "sample.msh")
            meshgen.write_gmsh(sample, path)
            t0 = time.perf_counter()
            om = oracle.Mesh(parse_gmsh(path))
            el_cpu = time.perf_counter() - t0
        cpu = {"value": round(om.n_interior / el_cpu / 1e6, 4), "unit": "Mcell/s", "cores": 1, "kind": "port",
               "sample": "oracle gmsh reader + geometry build (numpy, 1 thread) on the %dx%dx2 O-grid "
                         "(%d cells), %.1f s" % (nt, nr, om.n_interior, el_cpu)}
        log("CPU baseline %.3f Mcell/s (%.1f s)" % (cpu["value"], el_cpu))
    out = {"metric": "Mcells ingested/s (gmsh read + geometry precompute)", "value": round(N / el / 1e6, 3),
           "unit": "Mcell/s", "higher_is_better": True, "ms_per_ingest": round(el * 1e3, 1),
           "runs_s": [round(r, 3) for r in runs], "data": "synthetic",
           "config": {"workload": "cylinder_ns_ogrid_%dx%dx2" % (args.n_theta, args.n_r), "cells": N, "edges": F,
                      "file_bytes": size, "host_threads": host_threads()},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n-theta", type=int, default=2048)
    ap.add_argument("--n-r", type=int, default=1024)
    ap.add_argument("--stretch", type=float, default=None,
                    help="radial growth ratio (default: the generator's, 1.006 at 512+ radial cells)")
    ap.add_argument("--cpu-steps", type=int, default=4,
                    help="oracle steps timed for the CPU baseline, about 13 s on one core (0: skip)")
    ap.add_argument("--time-steps", type=int, default=20, help="steps timed per kernel with HIP events")
    ap.add_argument("--stage-path", default="faceTiles", choices=("split", "faceTiles", "fused", "gradTiles"),
                    help="kernels of one stage: gradient/face/update, gradient + face/update tiles, or one fused kernel")
    ap.add_argument("--tile-cells", type=int, default=-1, help="cells per tile of the tiled paths (-1: automatic)")
    ap.add_argument("--tile-order", default="blocks", choices=("blocks", "grow", "morton"),
                    help="how tiles are cut: Morton order of 16-cell blocks, greedily grown regions, Morton order")
    ap.add_argument("--kernel-order", default="tilesBackward",
                    choices=("forward", "tilesBackward", "gradientBackward", "bothBackward"),
                    help="which kernels walk their tiles / cell blocks last to first")
    ap.add_argument("--cell-order", default="tiles", choices=("reference", "tiles"),
                    help="device numbering of the cells: the mesh's, or tile by tile")
    ap.add_argument("--mixed", type=int, default=1, help="also time the mixed-precision form of the workload (0: skip)")
    ap.add_argument("--c2-steps", type=int, default=5000,
                    help="full-length FP64 and mixed runs of the workload compared for the mixed error (0: skip)")
    ap.add_argument("--limiter", default="venkatakrishnan", choices=("venkatakrishnan", "scalar"))
    ap.add_argument("--venkat-k", type=float, default=1.0)
    ap.add_argument("--c1-steps", type=int, default=1000,
                    help="steps of the 8,192-cell manufactured-solution latency leg (0: skip)")
    ap.add_argument("--dropin-stages", type=int, default=10,
                    help="stages timed at the ufvmSolver boundary (state upload + ebb_residual + R download; 0: skip)")
    ap.add_argument("--batched-meshes", type=int, default=256, help="meshes of the batched leg (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=64, help="steps per solve timed in the batched leg")
    ap.add_argument("--batched-seed", type=int, default=1)
    ap.add_argument("--batched-tune", type=int, default=1, help="run the batch autotuner first")
    ap.add_argument("--batched-solo-sample", type=int, default=32, help="solves timed one at a time")
    ap.add_argument("--serve", action="store_true", help="run the serving sweep instead")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rates", default="1,20,200", help="arrival rates (1/s); the highest also runs one at a time")
    ap.add_argument("--serve-max-steps", type=int, default=20000)
    ap.add_argument("--serve-round", type=int, default=200, help="steps per serving round")
    ap.add_argument("--serve-precision", default="fp64", choices=("mixed", "fp64"),
                    help="fp64: a 1e-8 drop of the rho RMS residual is below the mixed form's rounding floor")
    ap.add_argument("--serve-limiter", default="venkatakrishnan", choices=("venkatakrishnan", "scalar"),
                    help="the scalar limiter stalls these steady solves near a 1e-3 drop")
    ap.add_argument("--serve-venkat-k", type=float, default=5.0)
    ap.add_argument("--serve-solo", type=int, default=1, help="also serve the stream one at a time (0: skip)")
    ap.add_argument("--serve-seed", type=int, default=1)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (reference for the HBM byte counters) and exit")
    ap.add_argument("--ingest", action="store_true",
                    help="time the mesh ingest (gmsh read + geometry precompute) of the workload's mesh instead")
    ap.add_argument("--ingest-repeats", type=int, default=3)
    ap.add_argument("--strouhal", action="store_true",
                    help="only run the shedding comparison (FP64 and mixed precision on a 256x128 cylinder) and exit")
    ap.add_argument("--strouhal-steps", type=int, default=400000)
    ap.add_argument("--strouhal-from", type=float, default=40.0, help="shedding frequency measured after this t")
    ap.add_argument("--ramp-probe", action="store_true",
                    help="only time --steps-step windows after GPU idle gaps of 0-3 s and exit")
    args = ap.parse_args()
    if args.ramp_probe:
        ramp_probe(args)
        return
    if args.strouhal:
        strouhal_bench(args)
        return
    if args.ingest:
        ingest_bench(args)
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.serve:
        serve_bench(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the solver keeps the whole mesh on a single MI355X "
              "(multi-GPU sharding is a design note, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from ebbcfd_amd import Solver, build_mesh, meshgen

    log("generating O-grid %d x %d" % (args.n_theta, args.n_r))
    kw = {} if args.stretch is None else {"stretch": args.stretch}
    raw = meshgen.cylinder_ogrid(args.n_theta, args.n_r, seed=1, **kw)
    cfg = cylinder_config()
    cfg["extensions"] = {"stagePath": args.stage_path, "tileCells": args.tile_cells, "tileOrder": args.tile_order,
                         "cellOrder": args.cell_order, "limiter": args.limiter, "venkatK": args.venkat_k,
                         "kernelOrder": args.kernel_order}
    mesh = build_mesh(raw)
    N = mesh.n_interior
    log("mesh built: %d cells, %d edges" % (N, mesh.n_edges))
    s = Solver(mesh, cfg)
    log("solver on device (%.2f GB)" % (s.device_bytes / 1e9))
    stages = 2
    s.run(args.warmup)
    log("warmup done")
    # timed region: device-resident time loop, K steps, synchronised on both sides
    t0 = time.perf_counter()
    hist, t_end, dt_end = s.run(args.steps)
    el = time.perf_counter() - t0
    assert len(hist) == args.steps, "solver stopped early"
    assert np.all(np.isfinite(hist[:, 2:6])), "non-finite residuals"
    ms_per_step = el / args.steps * 1e3
    value = N * stages * args.steps / el / 1e6
    log("timed %d steps: %.3f ms/step, %.1f Mcell-stage/s" % (args.steps, ms_per_step, value))

    # per-kernel device time with HIP events on the solver stream
    tk = s.time_steps(args.time_steps)
    F = mesh.n_edges
    stage_ms = tk["step_ms"] / stages
    stage_gbs = N * STAGE_BYTES_PER_CELL / (stage_ms * 1e-3) / 1e9
    log("kernels: tile %.3f ms | gradient %.3f ms, faces %.3f ms, update %.3f ms | stages of a step %.3f ms"
        % (tk["fused_ms"], tk["gradient_ms"], tk["faces_ms"], tk["update_ms"], tk["step_ms"]))
    if tk["fused_ms"] > 0 and args.stage_path == "faceTiles":
        kernel, launch_ms = "k_face_tile", tk["fused_ms"]
        launch_bytes = N * STAGE_BYTES_PER_CELL
    elif tk["fused_ms"] > 0:
        kernel, launch_ms = "k_stage_tile", tk["fused_ms"]
        launch_bytes = N * STAGE_BYTES_PER_CELL
    else:
        kernel, launch_ms = "k_faces", tk["faces_ms"]
        launch_bytes = F * FACE_BYTES_PER_FACE + N * FACE_BYTES_PER_CELL
    achieved = launch_bytes / (launch_ms * 1e-3) / 1e9
    workload = "cylinder_ns_ogrid_%dx%dx2" % (args.n_theta, args.n_r)
    if args.stretch is not None:
        workload += "_stretch%g" % args.stretch
    traffic = MEASURED_TRAFFIC.get((kernel, workload, args.stage_path, args.tile_cells, args.tile_order,
                                    args.cell_order, args.kernel_order, args.limiter))

    dropin = None
    if args.dropin_stages > 0:
        # the drop-in keeps the reference's cell order on the device, so q and R cross PCIe as
        # they are (a renumbering solver would permute them on the host)
        sd = s
        if args.cell_order != "reference":
            ext = dict(cfg["extensions"], cellOrder="reference", kernelOrder="forward")
            sd = Solver(mesh, dict(cfg, extensions=ext))
        dropin = dropin_leg(sd, N, args.dropin_stages, stage_ms)
        del sd
    mixed = None
    if args.mixed:
        mixed = mixed_leg(mesh, cfg, args, N, value)
    c1 = None
    if args.c1_steps > 0:
        c1 = c1_leg(args.c1_steps)
    batched = None
    if args.batched_meshes > 0:
        del s
        batched = batched_leg(args)

    cpu = None
    if args.cpu_steps > 0:
        T = host_threads()
        log("CPU baseline: oracle, %d steps on the same mesh, %d threads, then 1" % (args.cpu_steps * 4, T))
        vT, elT = cpu_baseline(raw, cfg, args.cpu_steps * 4, N, T)
        v1, el1 = cpu_baseline(raw, cfg, args.cpu_steps, N, 1)
        cpu = {"value": round(vT, 4), "unit": "Mcell-stage/s", "cores": T, "kind": "port",
               "sample": "oracle (C restatement, gcc -O2 -ffp-contract=off, OpenMP over cells and edges, %d threads): "
                         "%d TVD-RK2 steps of the same 4.2M-cell cylinder case, %.1f s" % (T, args.cpu_steps * 4, elT),
               "single_core": {"value": round(v1, 4), "cores": 1, "steps": args.cpu_steps, "seconds": round(el1, 1)}}
        log("CPU baseline %.3f Mcell-stage/s on %d threads (%.1f s), %.3f on 1 (%.1f s)" % (vT, T, elT, v1, el1))

    out = {
        "metric": "Mcell-RK-stage updates/s @4M cells (%HBM roof)",
        "value": round(value, 3),
        "unit": "Mcell-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "cells": N,
                   "stage_path": args.stage_path, "tile_cells": args.tile_cells, "tile_order": args.tile_order,
                   "cell_order": args.cell_order, "kernel_order": args.kernel_order,
                   "edges": F, "flux": "rusanovFlux",
                   "limiter": ("venkatakrishnan (K=%g)" % args.venkat_k) if args.limiter == "venkatakrishnan"
                   else "scalar (solve.d:360-403)",
                   "integrator": "RK2_TVD", "stages_per_step": stages, "M": 0.2, "Re": 100, "CFL": cfg["CFL"]},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else int(traffic),
                     "avg_launch_ms": round(launch_ms, 4), "algorithmic_bytes_per_launch": int(launch_bytes)},
        "stage_roofline": {"bytes_per_cell_stage": STAGE_BYTES_PER_CELL, "stage_ms": round(stage_ms, 4),
                           "achieved": round(stage_gbs, 1), "frac": round(stage_gbs / HBM_PEAK_GBS, 4),
                           "unit": "GB/s"},
        "cpu_baseline": cpu,
        "mixed": mixed,
        "dropin_ufvmSolver": dropin,
        "c1": c1,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
