"""The fused tile stage kernel against the split kernels and the oracle.

The fused kernel evaluates ghosts, gradients/limiters, face fluxes and the stage
update of a tile of cells in one workgroup; faces on tile borders are computed by
both tiles with the same device code.  Every result must be bit-identical to the
three-kernel path (tileCells = 0), for tile sizes down to one cell.
"""
import copy

import numpy as np
import pytest

import oracle
from ebbcfd_amd import Solver, build_mesh

from helpers import channel_case, cylinder_case, mms_case, oracle_cfg, reference_fixture

pytestmark = pytest.mark.gpu

CASES = {
    "euler_roe": lambda: channel_case(integrator="Euler"),
    "rk2_roe_order1": lambda: channel_case(integrator="RK2", order=1),
    "rk2tvd_rusanov_sym_visc": lambda: channel_case(integrator="RK2_TVD", flux="rusanovFlux", walls="symmetry",
                                                    viscosity=True),
    "rk4_unlimited": lambda: channel_case(integrator="RK4", limited=False),
    "rk4_local_dt": lambda: channel_case(integrator="RK4", localTimestep=True),
    "cyl_ns_rk2tvd": lambda: cylinder_case(),
    "cyl_euler_roe_rk4": lambda: cylinder_case(viscous=False, flux="roeFlux", integrator="RK4"),
    "mms_viscous_rk4": lambda: mms_case(viscosity=True),
    "mms_fixture_tri_lp": lambda: reference_fixture("ms_refine_tri_pert_0"),
    "mms_fixture_quad_lp": lambda: reference_fixture("ms_refine_1"),
}


def with_tiles(cfg, n):
    c = copy.deepcopy(cfg)
    ext = dict(c.get("extensions", {}))
    ext["tileCells"] = n
    c["extensions"] = ext
    return c


@pytest.mark.parametrize("tile", [1, 13, 64, 128])
@pytest.mark.parametrize("name", sorted(CASES))
def test_fused_equals_split(name, tile):
    raw, cfg = CASES[name]()
    split = Solver(build_mesh(raw), with_tiles(cfg, 0))
    fused = Solver(build_mesh(raw), with_tiles(cfg, tile))
    h_s, t_s, dt_s = split.run(15)
    h_f, t_f, dt_f = fused.run(15)
    np.testing.assert_array_equal(h_f, h_s)
    np.testing.assert_array_equal(fused.q_all(), split.q_all())
    assert t_f == t_s and dt_f == dt_s
    g_s, l_s = split.cell_gradients()
    g_f, l_f = fused.cell_gradients()
    np.testing.assert_array_equal(g_f, g_s)
    np.testing.assert_array_equal(l_f, l_s)
    eq_s, fl_s, sm_s = split.edge_states()
    eq_f, fl_f, sm_f = fused.edge_states()
    np.testing.assert_array_equal(eq_f, eq_s)
    np.testing.assert_array_equal(fl_f, fl_s)


def test_fused_matches_oracle_many_tiles():
    raw, cfg = cylinder_case(n_theta=96, n_r=40, integrator="RK4", flux="roeFlux")
    cfg = with_tiles(cfg, 32)
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    s = Solver(build_mesh(raw), cfg)
    h_o, t_o, _ = om.run(25)
    h_g, t_g, _ = s.run(25)
    np.testing.assert_array_equal(h_g, h_o)
    np.testing.assert_array_equal(s.q, om.q[: om.n_interior])


def test_fused_large_mesh_equals_split():
    raw, cfg = cylinder_case(n_theta=512, n_r=256, stretch=1.015, CFL=0.05)
    split = Solver(build_mesh(raw), with_tiles(cfg, 0))
    fused = Solver(build_mesh(raw), with_tiles(cfg, -1))   # automatic tile size
    h_s, _, _ = split.run(12)
    h_f, _, _ = fused.run(12)
    np.testing.assert_array_equal(h_f, h_s)
    np.testing.assert_array_equal(fused.q, split.q)
    tf = fused.time_steps(2)
    ts = split.time_steps(2)
    assert tf["fused_ms"] > 0 and tf["faces_ms"] == 0
    assert ts["fused_ms"] == 0 and ts["faces_ms"] > 0


@pytest.mark.parametrize("tile", [0, 5, 128])
def test_fused_nan_cell_and_edge_reported_like_split(tile):
    raw, cfg = channel_case(integrator="RK2")
    ref = Solver(build_mesh(raw), with_tiles(cfg, 0))
    s = Solver(build_mesh(raw), with_tiles(cfg, tile))
    for x in (ref, s):
        q = x.q
        q[17, 0] = -1.0
        x.q = q
    with pytest.raises(Exception) as er:
        ref.run(3)
    with pytest.raises(Exception) as eg:
        s.run(3)
    assert eg.value.code == er.value.code
    assert eg.value.index == er.value.index and eg.value.step == er.value.step
