"""Serving loop: requests that arrive while others run are admitted by device footprint,
stepped together (continuous batching) and retired when they stop; each result must
equal the same request run alone (Solver.run, which stops at the same step through
extensions.rmsDrop / maxSteps) bit for bit."""
import time

import numpy as np
import pytest

from ebbcfd_amd import Server, Solver, build_mesh, estimate_device_bytes, meshgen

from helpers import base_config

pytestmark = pytest.mark.gpu


def _request(seed, cells, rms_drop=1e-4, max_steps=3000):
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny = meshgen.channel_bump_cells(cells)
    raw = meshgen.channel_bump(nx, ny, float(rng.uniform(0.02, 0.1)), seed=seed)
    q = [float(rng.uniform(0.3, 1.5)), float(rng.uniform(-5, 5)), 1.0, 1.0]
    bcs = [{"tag": "Inflow", "type": "fullState", "q": q}, {"tag": "Outflow", "type": "fullState", "q": q},
           {"tag": "Lower", "type": "inviscidWall", "q": q}, {"tag": "Upper", "type": "inviscidWall", "q": q}]
    cfg = base_config(initialConditions=q, boudaryConditions=bcs, forceBoundary="Lower", flux="roeFlux",
                      integrator="RK4", localTimestep=True, plotIter=-1,
                      extensions={"rmsDrop": rms_drop, "maxSteps": max_steps})
    return raw, cfg


def test_served_results_equal_solo_runs():
    reqs = [_request(100 + k, c) for k, c in enumerate([2000, 9000, 600, 30000, 4000, 12000])]
    meshes = [build_mesh(raw) for raw, _ in reqs]
    sv = Server(max_solves=8, steps_per_round=50, graph_steps=8, keep_results=True)
    tickets = [sv.submit(meshes[k], reqs[k][1]) for k in range(3)]
    time.sleep(0.05)                                  # the rest arrive while the first run
    tickets += [sv.submit(meshes[k], reqs[k][1]) for k in range(3, 6)]
    assert sv.wait(6, 300.0)
    st = sv.stats()
    print("served", st)
    assert st["done"] == 6 and st["failed"] == 0
    for k, t in enumerate(tickets):
        p = sv.poll(t)
        s = Solver(meshes[k], reqs[k][1])
        h, _, _ = s.run()
        print("request %d: %d cells, %d steps (solo %d), converged %d, latency %.3f s"
              % (k, p["cells"], p["steps"], len(h), p["converged"], p["done_s"] - p["arrival_s"]))
        assert p["state"] == 2 and p["steps"] == len(h)
        np.testing.assert_array_equal(sv.result(t), s.q)
    sv.close()


def test_served_mix_of_kinds_equals_solo_runs():
    """Requests of different kinds served together -- FP64 and mixed precision, face tiles
    and gradient tiles, the scalar and the Venkatakrishnan limiter -- form separate launch
    groups in one batch; each result still equals its request run alone."""
    kinds = [{}, {"precision": "mixed"}, {"stagePath": "gradTiles"},
             {"stagePath": "gradTiles", "precision": "mixed", "limiter": "venkatakrishnan", "venkatK": 5.0},
             {"limiter": "venkatakrishnan", "venkatK": 1.0}]
    reqs = []
    for k, ext in enumerate(kinds):
        raw, cfg = _request(400 + k, [3000, 7000, 1500, 20000, 5000][k], rms_drop=1e-3, max_steps=600)
        cfg["extensions"] = dict(cfg["extensions"], **ext)
        reqs.append((raw, cfg))
    meshes = [build_mesh(raw) for raw, _ in reqs]
    sv = Server(max_solves=8, steps_per_round=40, graph_steps=8, keep_results=True)
    tickets = [sv.submit(m, c) for m, (_, c) in zip(meshes, reqs)]
    assert sv.wait(len(reqs), 300.0)
    for k, t in enumerate(tickets):
        p = sv.poll(t)
        s = Solver(meshes[k], reqs[k][1])
        h, _, _ = s.run()
        assert p["state"] == 2 and p["steps"] == len(h), (k, p, len(h))
        np.testing.assert_array_equal(sv.result(t), s.q)
    sv.close()


def test_admission_waits_for_memory():
    reqs = [_request(200 + k, 8000, rms_drop=1e-3, max_steps=400) for k in range(4)]
    meshes = [build_mesh(raw) for raw, _ in reqs]
    need = max(estimate_device_bytes(m, c) for m, (_, c) in zip(meshes, reqs))
    sv = Server(device_bytes_limit=int(2.2 * need), max_solves=8, steps_per_round=40, keep_results=True)
    tickets = [sv.submit(m, c) for m, (_, c) in zip(meshes, reqs)]
    assert sv.wait(4, 300.0)
    polls = [sv.poll(t) for t in tickets]
    assert all(p["state"] == 2 for p in polls)
    # only two fit at once: the third and fourth were admitted after an earlier one finished
    first_done = min(p["done_s"] for p in polls[:2])
    assert polls[2]["admit_s"] >= first_done and polls[3]["admit_s"] >= first_done
    for m, (raw, cfg), t in zip(meshes, reqs, tickets):
        s = Solver(m, cfg)
        s.run()
        np.testing.assert_array_equal(sv.result(t), s.q)
    # a request larger than the whole limit fails instead of waiting forever
    big = _request(300, 60000)
    bm = build_mesh(big[0])
    t = sv.submit(bm, big[1])
    assert sv.wait(5, 60.0)
    assert sv.poll(t)["state"] == 3 and sv.poll(t)["code"] == 10
    sv.close()
