"""HIP path vs the CPU oracle, called through the C ABI on the GPU.

Bar: the FP64 kernels keep the reference's operation order with no FMA
contraction, so per-stage results are compared bit for bit (assert_array_equal
treats +0 and -0 as equal); multi-step trajectories must stay within 1e-11
relative (they are expected to stay bit-identical as well).  The ConstP and
TempPresInflow ghost states use the device pow(), so those cases carry a
1e-13 relative tolerance.
"""
import numpy as np
import pytest

import oracle
from ebbcfd_amd import Solver, build_mesh

from helpers import (channel_case, cylinder_case, flat_plate_fixture, hanging_node_mesh, mixed_mesh, mms_case, oracle_cfg, perturb,
                     reference_fixture)

pytestmark = pytest.mark.gpu


def _pair(raw, cfg, q=None):
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    s = Solver(build_mesh(raw), cfg)
    if q is not None:
        qq = om.q
        qq[: om.n_interior] = q
        om.q = qq
        s.q = q
    else:
        np.testing.assert_array_equal(s.q, om.q[: om.n_interior])
    return om, s


RESIDUAL_CASES = {
    "roe_inviscid_wall": lambda: channel_case(flux="roeFlux"),
    "rusanov_inviscid_wall": lambda: channel_case(flux="rusanovFlux"),
    "roe_symmetry": lambda: channel_case(walls="symmetry", flux="roeFlux"),
    "roe_symmetry_viscous": lambda: channel_case(walls="symmetry", flux="roeFlux", viscosity=True),
    "roe_unlimited": lambda: channel_case(flux="roeFlux", limited=False),
    "roe_order1": lambda: channel_case(flux="roeFlux", order=1),
    "cyl_ns_rusanov": lambda: cylinder_case(),
    "cyl_ns_roe": lambda: cylinder_case(flux="roeFlux"),
    "cyl_euler_wall": lambda: cylinder_case(viscous=False, flux="roeFlux"),
    "mms_dirichlet": lambda: mms_case(),
    "mms_dirichlet_viscous": lambda: mms_case(viscosity=True),
    "mms_fixture_lp": lambda: reference_fixture("ms_refine_tri_pert_1"),
    "mms_fixture_quads_lp": lambda: reference_fixture("ms_refine_1"),
    "flat_plate_fixture": lambda: flat_plate_fixture(0),
    "pentagons_viscous": lambda: (hanging_node_mesh(), channel_case(flux="roeFlux", viscosity=True)[1]),
    "mixed_viscous": lambda: (mixed_mesh(), channel_case(flux="roeFlux", viscosity=True)[1]),
    "mixed_lp": lambda: (mixed_mesh(), channel_case(flux="rusanovFlux", lpThresh=1.0)[1]),
}


@pytest.mark.parametrize("name", sorted(RESIDUAL_CASES))
@pytest.mark.parametrize("limit", [True, False])
def test_residual_bitwise(name, limit):
    raw, cfg = RESIDUAL_CASES[name]()
    om0 = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    q = perturb(om0.q[: om0.n_interior], 0.03, seed=7)
    om, s = _pair(raw, cfg, q)
    R_o, nd_o, rm_o = om.residual(limit=limit, dt_update=True)
    R_g, nd_g, rm_g = s.ufvm_solver(limit=limit, dt_update=True)
    np.testing.assert_array_equal(R_g, R_o)
    assert nd_g == nd_o
    assert rm_g == rm_o
    # ghost states, gradients/limiters and edge data of the same evaluation
    np.testing.assert_array_equal(s.q_all(), om.q)
    g_g, l_g = s.cell_gradients()
    g_o, l_o, _ = om.cell_data()
    if cfg.get("order", 2) > 1:
        np.testing.assert_array_equal(g_g, g_o[: om.n_interior])
        np.testing.assert_array_equal(l_g, l_o[: om.n_interior])
    eq_g, fl_g, sm_g = s.edge_states()
    eq_o, fl_o, sm_o = om.edge_data()
    np.testing.assert_array_equal(eq_g, eq_o)
    # the device stores len*flux; dividing back by len is exact to 1 ulp
    np.testing.assert_allclose(fl_g, fl_o, rtol=4e-16, atol=1e-300)
    np.testing.assert_allclose(sm_g, sm_o, rtol=4e-16, atol=1e-300)


STEP_CASES = {
    "euler_roe": lambda: channel_case(integrator="Euler"),
    "rk2_roe": lambda: channel_case(integrator="RK2"),
    "rk2tvd_rusanov": lambda: channel_case(integrator="RK2_TVD", flux="rusanovFlux"),
    "rk4_roe": lambda: channel_case(integrator="RK4"),
    "rk4_nomultistage": lambda: channel_case(integrator="RK4", multistageLimiting=False),
    "local_dt_rk2tvd": lambda: channel_case(integrator="RK2_TVD", localTimestep=True),
    "local_dt_rk4": lambda: channel_case(integrator="RK4", localTimestep=True),
    "local_dt_euler": lambda: channel_case(integrator="Euler", localTimestep=True),
    "cyl_ns_rk2tvd": lambda: cylinder_case(),
    "mms_rk4": lambda: mms_case(),
    "mms_fixture_euler_lp": lambda: reference_fixture("ms_refine_tri_pert_0"),
    "flat_plate_fixture_0": lambda: flat_plate_fixture(0),
    "flat_plate_fixture_1": lambda: flat_plate_fixture(1),
    "pentagons_rk2tvd_viscous": lambda: (hanging_node_mesh(), channel_case(integrator="RK2_TVD", viscosity=True)[1]),
    "mixed_rk2tvd_viscous": lambda: (mixed_mesh(), channel_case(integrator="RK2_TVD", viscosity=True)[1]),
    "mixed_rk4_lp": lambda: (mixed_mesh(seed=8), channel_case(integrator="RK4", lpThresh=1.0)[1]),
    # the Venkatakrishnan limiter (an extensions option; the oracle restates it, pinned by
    # tests/test_oracle_mesh.py's property checks)
    "venkat_rk4_roe": lambda: channel_case(integrator="RK4", extensions={"limiter": "venkatakrishnan", "venkatK": 0.5}),
    "venkat_cyl_ns_rk2tvd": lambda: cylinder_case(extensions={"limiter": "venkatakrishnan", "venkatK": 1.0}),
    "venkat_quads_mixed": lambda: (mixed_mesh(seed=3), channel_case(integrator="RK2", extensions={"limiter": "venkatakrishnan"})[1]),
    # as the serving requests run: Roe, RK4, local time steps, K = 5
    "venkat_k5_local_dt_rk4": lambda: channel_case(nx=40, ny=12, M=1.2, integrator="RK4", localTimestep=True,
                                                   extensions={"limiter": "venkatakrishnan", "venkatK": 5.0}),
}


@pytest.mark.parametrize("name", sorted(STEP_CASES))
def test_run_matches_oracle(name):
    raw, cfg = STEP_CASES[name]()
    nsteps = 40
    om, s = _pair(raw, cfg)
    hist_o, t_o, dt_o = om.run(nsteps)
    hist_g, t_g, dt_g = s.run(nsteps)
    assert hist_g.shape == hist_o.shape
    qo = om.q[: om.n_interior]
    qg = s.q
    scale = np.abs(qo).max(axis=0)
    assert np.all(np.abs(qg - qo) <= 1e-11 * scale), np.abs(qg - qo).max()
    np.testing.assert_array_equal(qg, qo)
    np.testing.assert_array_equal(hist_g, hist_o)
    assert t_g == t_o and dt_g == dt_o


def test_step_api_matches_oracle():
    raw, cfg = channel_case(integrator="RK4")
    om, s = _pair(raw, cfg)
    dt = cfg["dt"]
    for _ in range(5):
        dt_o, rm_o = om.step(dt)
        dt_g, rm_g = s.step(dt)
        assert dt_g == dt_o and rm_g == rm_o
        dt = dt_o
    np.testing.assert_array_equal(s.q, om.q[: om.n_interior])


def test_run_in_batches_equals_one_run():
    raw, cfg = channel_case(integrator="RK2_TVD")
    s1 = Solver(build_mesh(raw), cfg)
    s2 = Solver(build_mesh(raw), cfg)
    h1, _, _ = s1.run(30)
    parts = [s2.run(7)[0], s2.run(11)[0], s2.run(12)[0]]
    np.testing.assert_array_equal(np.concatenate(parts), h1)
    np.testing.assert_array_equal(s1.q, s2.q)


def test_run_across_history_batches_matches_oracle():
    """One run longer than a history batch (256 steps): the batch boundary inside the run
    sums the last step's boundary force before the host copies the rows."""
    raw, cfg = channel_case(integrator="RK2_TVD", viscosity=True, walls="viscidWall")
    om, s = _pair(raw, cfg)
    hist_o, t_o, dt_o = om.run(300)
    hist_g, t_g, dt_g = s.run(300)
    assert len(hist_g) == 300
    np.testing.assert_array_equal(hist_g, hist_o)
    np.testing.assert_array_equal(s.q, om.q[: om.n_interior])
    assert t_g == t_o and dt_g == dt_o


def test_stops_at_tend():
    raw, cfg = channel_case(integrator="Euler", dynamicDt=False, dt=1.0e-3, tEnd=0.0105)
    om, s = _pair(raw, cfg)
    hist_o, t_o, _ = om.run(10 ** 6)
    hist_g, t_g, _ = s.run(-1)
    assert len(hist_g) == len(hist_o) == 11
    assert t_g == t_o
    np.testing.assert_array_equal(s.q, om.q[: om.n_interior])
    # every row, drag and lift included: the loop stops mid-batch, and the last step's
    # force is summed after the loop has stopped
    np.testing.assert_array_equal(hist_g, hist_o)


def test_nan_edge_reported_like_reference():
    raw, cfg = channel_case(integrator="Euler")
    om, s = _pair(raw, cfg)
    q = s.q
    q[5, 0] = -1.0          # negative density: NaN Roe average on that cell's faces
    s.q = q
    qq = om.q
    qq[5, 0] = -1.0
    om.q = qq
    with pytest.raises(oracle.OracleError) as eo:
        om.residual()
    with pytest.raises(Exception) as eg:
        s.ufvm_solver()
    assert eg.value.code == 7
    assert eg.value.index == eo.value.index


def test_error_step_is_the_step_that_failed():
    """A fixed time step too large for the channel makes the run diverge after a few steps:
    the device stops in the step the oracle stops in, on the same edge, and reports that
    step (the number of steps completed before it)."""
    raw, cfg = channel_case(integrator="Euler", dynamicDt=False, dt=0.03)
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    done = 0
    with pytest.raises(oracle.OracleError) as eo:
        while done < 100:
            om.run(1)
            done += 1
    assert done >= 2
    s = Solver(build_mesh(raw), cfg)
    with pytest.raises(Exception) as eg:
        s.run(100)
    assert eg.value.code == 7 and eg.value.index == eo.value.index
    assert eg.value.step == done


def test_solvers_release_device_memory():
    """Ten solvers created, run (with edge states, which allocate on first use) and
    destroyed leave the device's free memory where it was."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    raw, cfg = channel_case(nx=200, ny=100, integrator="RK2_TVD")   # 40,000 cells
    mesh = build_mesh(raw)
    s = Solver(mesh, cfg)
    s.run(2)
    del s
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    f0 = free.value
    for _ in range(10):
        s = Solver(mesh, cfg)
        s.run(2)
        s.edge_states()
        del s
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    assert f0 - free.value < (64 << 20), (f0, free.value)


def test_run_after_an_error_needs_a_reset():
    """The reference aborts on a NaN (app/solver.d:395-423).  Here the loop stops, further
    runs are refused (EBB_ERR_STATE), and ebb_run_reset plus a new state starts it again
    exactly as a fresh solver would."""
    raw, cfg = channel_case(integrator="RK2")
    s = Solver(build_mesh(raw), cfg)
    q0 = s.q
    q = q0.copy()
    q[17, 0] = -1.0
    s.q = q
    with pytest.raises(Exception) as e1:
        s.run(3)
    assert e1.value.code in (7, 8) and e1.value.step == 0   # the first step, for edges and cells alike
    with pytest.raises(Exception) as e2:
        s.run(3)
    assert e2.value.code == 9
    s.run_reset()
    s.q = q0
    h, t, dt = s.run(5)
    fresh = Solver(build_mesh(raw), cfg)
    hf, tf, dtf = fresh.run(5)
    np.testing.assert_array_equal(h, hf)
    assert t == tf and dt == dtf
    np.testing.assert_array_equal(s.q, fresh.q)
