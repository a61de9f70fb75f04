"""Batched solves (many meshes per launch) against solo runs.

Each solve in a batch runs the solo kernels' code on its own (mesh, block/tile) work
items, so its per-step history and its states must equal a solo Solver.run of the same
mesh and config bit for bit -- whatever else is in the batch, with or without graph
capture, and after the autotuner (which must also restore every state).  Admission is
by device footprint.
"""
import copy

import numpy as np
import pytest

from ebbcfd_amd import Batch, Solver, build_mesh, estimate_device_bytes, meshgen

from helpers import (base_config, channel_case, cylinder_case, flat_plate_fixture, hanging_node_mesh, mixed_mesh,
                     mms_case, tiny_mesh)

pytestmark = pytest.mark.gpu


def _channel(nx, ny, h, seed, M, aoa, **over):
    raw = meshgen.channel_bump(nx, ny, h, seed=seed)
    q = [M, aoa, 1.0, 1.0]
    bcs = [{"tag": "Inflow", "type": "fullState", "q": q}, {"tag": "Outflow", "type": "fullState", "q": q},
           {"tag": "Lower", "type": "inviscidWall", "q": q}, {"tag": "Upper", "type": "inviscidWall", "q": q}]
    cfg = base_config(initialConditions=q, boudaryConditions=bcs, forceBoundary="Lower", flux="roeFlux",
                      integrator="RK4", **over)
    return raw, cfg


def _cases():
    rng = np.random.Generator(np.random.PCG64(11))
    out = []
    for k in range(6):
        nx = int(rng.integers(12, 90))
        out.append(_channel(nx, max(4, nx // 3), float(rng.uniform(0.02, 0.1)), int(rng.integers(1, 1000)),
                            float(rng.uniform(0.3, 1.2)), float(rng.uniform(-5, 5))))
    out.append(mms_case(n=24))                                        # Dirichlet + source, RK4
    out.append(cylinder_case(n_theta=64, n_r=32, CFL=0.05))           # viscous, Rusanov, TVD-RK2, force faces
    raw, cfg = channel_case(integrator="RK2", localTimestep=True)     # local time steps, RK2
    out.append((raw, cfg))
    return out


@pytest.mark.parametrize("graph_steps", [0, 8])
def test_batch_equals_solo_runs(graph_steps):
    cases = _cases()
    steps = 120
    b = Batch(graph_steps=graph_steps, keep_history=True)
    ids = [b.add(build_mesh(raw), cfg) for raw, cfg in cases]
    assert b.run(steps) == len(cases)
    for (raw, cfg), sid in zip(cases, ids):
        s = Solver(build_mesh(raw), cfg)
        hs, t, dt = s.run(steps)
        hb = b.history(sid)
        st = b.status(sid)
        assert len(hb) == steps and st["steps"] == steps
        np.testing.assert_array_equal(hb, hs)
        np.testing.assert_array_equal(b.q(sid), s.q)
        assert (st["t"], st["dt"]) == (t, dt)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_grad_tiles_equal_solo_runs(precision):
    """Members on the gradient-tile path (border blocks as gradient items, tiles that compute
    their own gradients) beside members on face tiles, in one graph-captured batch."""
    cases = _cases()
    steps = 60
    cfgs = []
    for k, (raw, cfg) in enumerate(cases):
        c = copy.deepcopy(cfg)
        c["extensions"] = dict(c.get("extensions", {}), stagePath="gradTiles" if k % 3 else "faceTiles",
                               precision=precision)
        cfgs.append(c)
    b = Batch(graph_steps=8, keep_history=True)
    ids = [b.add(build_mesh(raw), c) for (raw, _), c in zip(cases, cfgs)]
    assert b.run(steps) == len(cases)
    for (raw, _), c, sid in zip(cases, cfgs, ids):
        s = Solver(build_mesh(raw), c)
        hs, t, dt = s.run(steps)
        np.testing.assert_array_equal(b.history(sid), hs)
        np.testing.assert_array_equal(b.q(sid), s.q)
        st = b.status(sid)
        assert (st["t"], st["dt"]) == (t, dt)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_mixed_element_shapes_equal_solo_runs(precision):
    """One batch holding meshes of every cell shape the device lays out: triangles (3 slots
    per cell), triangles beside quads (4), pentagons at hanging nodes (6) and the reference's
    viscous flat-plate quads (fullState, viscidWall, constP, symmetry; Roe, Euler steps).
    Each shape is its own launch group; every member still equals its solo run."""
    cases = [
        channel_case(integrator="RK4"),
        (mixed_mesh(), channel_case(integrator="RK2_TVD", viscosity=True)[1]),
        (hanging_node_mesh(), channel_case(integrator="RK2", viscosity=True)[1]),
        flat_plate_fixture(0),
        (mixed_mesh(nx=30, ny=10, seed=9), channel_case(integrator="RK4", flux="rusanovFlux")[1]),
    ]
    cfgs = []
    for raw, cfg in cases:
        c = copy.deepcopy(cfg)
        c["extensions"] = dict(c.get("extensions", {}), precision=precision)
        cfgs.append(c)
    steps = 40
    b = Batch(graph_steps=8, keep_history=True)
    ids = [b.add(build_mesh(raw), c) for (raw, _), c in zip(cases, cfgs)]
    assert b.run(steps) == len(cases)
    for (raw, _), c, sid in zip(cases, cfgs, ids):
        s = Solver(build_mesh(raw), c)
        hs, t, dt = s.run(steps)
        np.testing.assert_array_equal(b.history(sid), hs)
        np.testing.assert_array_equal(b.q(sid), s.q)
        st = b.status(sid)
        assert (st["t"], st["dt"]) == (t, dt)


def test_batch_tiny_members_equal_solo_runs():
    """Members of one triangle, two triangles and one quad (every cell on the boundary: a
    tile shorter than a 16-cell block, an RMS sum of the element tail only) beside an
    ordinary mesh, on both stage paths."""
    q = [0.3, 10.0, 1.0, 1.0]
    cases = []
    for kind in ("tri1", "tri2", "quad1"):
        for path in ("faceTiles", "gradTiles"):
            cfg = base_config(initialConditions=q, boudaryConditions=[{"tag": "Farfield", "type": "fullState", "q": q}],
                              forceBoundary="Farfield", integrator="RK2_TVD")
            cfg["extensions"] = {"stagePath": path}
            cases.append((tiny_mesh(kind), cfg))
    cases.append(channel_case(integrator="RK4"))
    steps = 20
    b = Batch(graph_steps=8, keep_history=True)
    ids = [b.add(build_mesh(raw), cfg) for raw, cfg in cases]
    assert b.run(steps) == len(cases)
    for (raw, cfg), sid in zip(cases, ids):
        s = Solver(build_mesh(raw), cfg)
        hs, t, dt = s.run(steps)
        np.testing.assert_array_equal(b.history(sid), hs)
        np.testing.assert_array_equal(b.q(sid), s.q)
        assert (b.status(sid)["t"], b.status(sid)["dt"]) == (t, dt)


def test_batch_members_join_and_leave():
    """Continuous batching: a solve that joins late, or whose neighbours leave, still equals
    its solo run; stopped solves drop out."""
    cases = _cases()[:5]
    b = Batch(graph_steps=8, keep_history=True)
    first = [b.add(build_mesh(raw), cfg) for raw, cfg in cases[:3]]
    b.run(40)
    b.remove(first[1])
    late = [b.add(build_mesh(raw), cfg) for raw, cfg in cases[3:]]
    b.run(60)
    plan = {first[0]: (cases[0], 100), first[2]: (cases[2], 100), late[0]: (cases[3], 60), late[1]: (cases[4], 60)}
    for sid, ((raw, cfg), n) in plan.items():
        s = Solver(build_mesh(raw), cfg)
        hs, _, _ = s.run(n)
        np.testing.assert_array_equal(b.history(sid), hs)
        np.testing.assert_array_equal(b.q(sid), s.q)


@pytest.mark.parametrize("path,precision", [("faceTiles", "fp64"), ("gradTiles", "fp64"), ("gradTiles", "mixed")])
def test_batch_member_retiled_as_batch_grows_and_shrinks(path, precision):
    """Automatic tiles follow the cells running together: a small solve admitted alone gets
    32-cell tiles (as a solo run); when a ~60k-cell solve joins, the cells sharing the
    launches grow more than twofold and the small solve is cut into larger tiles (on the
    gradient-tile path with new ring and border tables); when it leaves, back.  Both must
    still equal their solo runs bit for bit."""
    def spread(cells, cus):   # cells per automatic face tile (two tiles per CU, 16-cell blocks, 32-768)
        return max(32, min(768, cells // (2 * cus) // 16 * 16))

    raw_s, cfg_s = _channel(40, 12, 0.05, 5, 0.5, 2.0, extensions={"stagePath": path, "precision": precision})
    raw_b, cfg_b = _channel(300, 100, 0.06, 6, 0.7, -1.0, extensions={"stagePath": path, "precision": precision})
    ms, mb = build_mesh(raw_s), build_mesh(raw_b)
    # the join and the leave cross the twofold re-tiling threshold on any GPU of 32 to 304
    # CUs (MI355X: 256, where the small solve goes from 32-cell tiles to 112 and back)
    for cus in (32, 64, 128, 256, 304):
        alone, joined = spread(ms.n_interior, cus), spread(ms.n_interior + mb.n_interior, cus)
        assert joined > 2 * alone, (cus, alone, joined)
    b = Batch(graph_steps=8, keep_history=True)
    small = b.add(ms, cfg_s)
    b.run(30)
    big = b.add(mb, cfg_b)
    b.run(30)
    hb_big, q_big = b.history(big), b.q(big)
    b.remove(big)
    b.run(30)
    s = Solver(build_mesh(raw_s), cfg_s)
    hs, _, _ = s.run(90)
    np.testing.assert_array_equal(b.history(small), hs)
    np.testing.assert_array_equal(b.q(small), s.q)
    s = Solver(build_mesh(raw_b), cfg_b)
    hs, _, _ = s.run(30)
    np.testing.assert_array_equal(hb_big, hs)
    np.testing.assert_array_equal(q_big, s.q)


def test_batch_convergence_stop_and_max_steps():
    raw, cfg = _channel(30, 10, 0.05, 3, 0.5, 1.0, limited=False, localTimestep=True)
    cfg["extensions"] = {"rmsDrop": 1e-3}
    raw2, cfg2 = _channel(20, 8, 0.05, 4, 0.6, 0.0)
    cfg2["extensions"] = {"maxSteps": 37}
    b = Batch(graph_steps=8, keep_history=True)
    a = b.add(build_mesh(raw), cfg)
    c = b.add(build_mesh(raw2), cfg2)
    left = 2
    for _ in range(100):
        left = b.run(400)
        if left == 0:
            break
    sa, sc = b.status(a), b.status(c)
    assert sc["steps"] == 37 and not sc["running"]
    assert sa["converged"] == 1 and sa["rho_rms"] <= 1e-3 * sa["rho_rms_first"]
    s = Solver(build_mesh(raw), cfg)
    hs, _, _ = s.run()
    assert len(hs) == sa["steps"]
    np.testing.assert_array_equal(b.q(a), s.q)


def test_batch_member_error_is_reported_and_isolated():
    """A member that diverges stops with the error its solo run reports (kind and reference
    edge or cell); the members beside it run on and equal their solo runs."""
    cases = _cases()[:3]
    raw_bad, cfg_bad = _channel(30, 10, 0.05, 7, 0.8, 2.0)
    cfg_bad["CFL"] = 40.0                     # far above the explicit limit: NaN within a few steps
    cfg_bad["cflAdjust"] = False
    solo = Solver(build_mesh(raw_bad), cfg_bad)
    with pytest.raises(Exception) as e:
        solo.run(200)
    b = Batch(graph_steps=8, keep_history=True)
    ids = [b.add(build_mesh(raw), cfg) for raw, cfg in cases]
    bad = b.add(build_mesh(raw_bad), cfg_bad)
    assert b.run(200) == len(cases)
    st = b.status(bad)
    assert not st["running"] and int(st["error"]) == e.value.code and int(st["error_index"]) == e.value.index
    for (raw, cfg), sid in zip(cases, ids):
        s = Solver(build_mesh(raw), cfg)
        hs, _, _ = s.run(200)
        np.testing.assert_array_equal(b.history(sid), hs)
        np.testing.assert_array_equal(b.q(sid), s.q)


def test_batch_admission_by_footprint():
    raw, cfg = _channel(60, 20, 0.05, 5, 0.5, 0.0)
    m = build_mesh(raw)
    need = estimate_device_bytes(m, cfg)
    s = Solver(m, cfg)
    assert 0.85 * s.device_bytes <= need <= 1.25 * s.device_bytes, (need, s.device_bytes)
    b = Batch(device_bytes_limit=int(2.5 * need))
    b.add(m, cfg)
    b.add(m, cfg)
    with pytest.raises(Exception) as e:
        b.add(m, cfg)
    assert getattr(e.value, "code", None) == 10
    assert b.info()["solves"] == 2


@pytest.mark.parametrize("path", ["faceTiles", "gradTiles"])
def test_batch_tuner_keeps_states(path):
    """The autotuner re-tiles every member for each trial (and restores every state); the
    members then still equal their solo runs."""
    cases = []
    for raw, cfg in _cases()[:4]:
        c = copy.deepcopy(cfg)
        c["extensions"] = dict(c.get("extensions", {}), stagePath=path)
        cases.append((raw, c))
    b = Batch(graph_steps=8, keep_history=True)
    ids = [b.add(build_mesh(raw), cfg) for raw, cfg in cases]
    b.run(16)
    res = b.tune(steps=8)
    assert res["tile_cells"] in (-1, 512, 384, 256) and res["graph_steps"] in (0, 2, 8, 32)
    b.run(24)
    for (raw, cfg), sid in zip(cases, ids):
        s = Solver(build_mesh(raw), cfg)
        hs, _, _ = s.run(40)
        np.testing.assert_array_equal(b.history(sid), hs)
        np.testing.assert_array_equal(b.q(sid), s.q)
