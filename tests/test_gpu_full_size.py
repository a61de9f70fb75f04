"""The throughput workload at its full size, bit for bit: FP64 Navier-Stokes flow past a
cylinder on a 2048 x 1024 O-grid split into 4,194,304 triangles (Rusanov flux, TVD-RK2,
viscous wall and far field) with the reference's scalar limiter, and with the
Venkatakrishnan limiter (K = 1) exactly as bench.py runs it (device cells numbered tile by
tile, each XCD's slice of tiles walked last to first), and in the reference numbering.

Two steps (four stages: ghosts, gradients, limiters, faces, residuals, updates, RMS
residuals, forces, time step and CFL update) on every stage path are compared with the
oracle over the whole per-step history and every interior state.  The oracle's two steps
take a few seconds on one host core at this size, so the full-size check is direct, not
through a size-independent property.
"""
import copy

import numpy as np
import pytest

import bench
import oracle
from ebbcfd_amd import Solver, build_mesh, meshgen

from helpers import oracle_cfg

pytestmark = pytest.mark.gpu

STEPS = 2


@pytest.fixture(scope="module")
def workload():
    raw = meshgen.cylinder_ogrid(2048, 1024, seed=1)
    cfg = bench.cylinder_config()
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    assert om.n_interior == 4_194_304
    hist, t, dt = om.run(STEPS)
    yield raw, cfg, hist, t, dt, om.q[: om.n_interior].copy()


@pytest.mark.parametrize("path,cells,korder", [("faceTiles", "reference", "forward"), ("split", "reference", "forward"),
                                               ("fused", "reference", "forward"), ("gradTiles", "reference", "forward"),
                                               ("faceTiles", "tiles", "tilesBackward")])
def test_full_size_steps_match_oracle(workload, path, cells, korder):
    raw, cfg, hist_o, t_o, dt_o, q_o = workload
    c = copy.deepcopy(cfg)
    c["extensions"] = {"stagePath": path, "cellOrder": cells, "kernelOrder": korder}
    s = Solver(build_mesh(raw), c)
    hist_g, t_g, dt_g = s.run(STEPS)
    np.testing.assert_array_equal(hist_g, hist_o)
    assert t_g == t_o and dt_g == dt_o
    q_g = s.q
    assert np.all(np.abs(q_g - q_o) <= 1e-11 * np.abs(q_o).max(axis=0))
    np.testing.assert_array_equal(q_g, q_o)


@pytest.fixture(scope="module")
def workload_venkat():
    raw = meshgen.cylinder_ogrid(2048, 1024, seed=1)
    cfg = bench.cylinder_config()
    cfg["extensions"] = {"limiter": "venkatakrishnan", "venkatK": 1.0}
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    hist, t, dt = om.run(STEPS)
    yield raw, cfg, hist, t, dt, om.q[: om.n_interior].copy()


@pytest.mark.parametrize("path,cells,korder", [("faceTiles", "tiles", "tilesBackward"),
                                               ("faceTiles", "reference", "forward"),
                                               ("gradTiles", "reference", "forward")])
def test_full_size_venkatakrishnan_matches_oracle(workload_venkat, path, cells, korder):
    raw, cfg, hist_o, t_o, dt_o, q_o = workload_venkat
    c = copy.deepcopy(cfg)
    c["extensions"] = dict(cfg["extensions"], stagePath=path, cellOrder=cells, kernelOrder=korder)
    s = Solver(build_mesh(raw), c)
    hist_g, t_g, dt_g = s.run(STEPS)
    np.testing.assert_array_equal(hist_g, hist_o)
    assert t_g == t_o and dt_g == dt_o
    np.testing.assert_array_equal(s.q, q_o)
