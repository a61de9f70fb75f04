"""The tiled stage paths against the split kernels and the oracle.

"faceTiles": the gradient kernel, then one kernel that computes every face of a
tile of cells into LDS and updates those cells.  "fused": one kernel per stage
that also computes ghosts and gradients/limiters for the tile and its halo.  "gradTiles":
the face tiles' kernel also computes its own cells' ghosts, gradients and limiters into LDS;
a border kernel computes those of the cells other tiles need (each tile's ring).
Faces on tile borders are computed by both tiles with the same device code, so
every result must be bit-identical to the three-kernel path, for tile sizes
down to one cell.
"""
import copy

import numpy as np
import pytest

import oracle
from ebbcfd_amd import Solver, build_mesh, meshgen

from helpers import (base_config, channel_case, cylinder_case, flat_plate_fixture, hanging_node_mesh, mixed_mesh, mms_case,
                     oracle_cfg, perturb, reference_fixture, tiny_mesh)

pytestmark = pytest.mark.gpu

CASES = {
    "euler_roe": lambda: channel_case(integrator="Euler"),
    "rk2tvd_odd_count": lambda: channel_case(nx=25, ny=7, integrator="RK2_TVD"),   # 350 cells: short last block
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
    "flat_plate_fixture": lambda: flat_plate_fixture(0),
    "pentagons_rk2tvd_visc": lambda: (hanging_node_mesh(), channel_case(integrator="RK2_TVD", viscosity=True)[1]),
    "mixed_rk2tvd_visc": lambda: (mixed_mesh(), channel_case(integrator="RK2_TVD", viscosity=True)[1]),
    "mixed_rk4_lp": lambda: (mixed_mesh(seed=8), channel_case(integrator="RK4", lpThresh=1.0)[1]),
    "quads_rk2tvd_visc": lambda: (meshgen.channel_bump_quads(24, 8),
                                  channel_case(integrator="RK2_TVD", viscosity=True)[1]),
}


def with_path(cfg, path, n=-1, order="blocks", cells="reference"):
    c = copy.deepcopy(cfg)
    ext = dict(c.get("extensions", {}))
    ext.update(stagePath=path, tileCells=n, tileOrder=order, cellOrder=cells)
    c["extensions"] = ext
    return c


@pytest.mark.parametrize("path,tile,order,cells", [
    ("split", -1, "blocks", "reference"),
    ("faceTiles", -1, "blocks", "reference"), ("faceTiles", 32, "blocks", "tiles"),
    ("faceTiles", 29, "morton", "reference"), ("faceTiles", 1, "grow", "tiles"), ("faceTiles", 256, "grow", "reference"),
    ("fused", 1, "grow", "tiles"), ("fused", 13, "grow", "reference"), ("fused", 64, "blocks", "tiles"),
    ("fused", 128, "morton", "reference"),
    ("gradTiles", -1, "blocks", "reference"), ("gradTiles", 32, "blocks", "tiles"), ("gradTiles", 29, "morton", "reference"),
    ("gradTiles", 1, "grow", "tiles"), ("gradTiles", 256, "grow", "reference")])
@pytest.mark.parametrize("name", sorted(CASES))
def test_tiled_equals_split(name, path, tile, order, cells):
    raw, cfg = CASES[name]()
    split = Solver(build_mesh(raw), with_path(cfg, "split", order="grow"))
    fused = Solver(build_mesh(raw), with_path(cfg, path, tile, order, cells))
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


@pytest.mark.parametrize("path", ["fused", "faceTiles", "gradTiles"])
def test_tiled_matches_oracle_many_tiles(path):
    raw, cfg = cylinder_case(n_theta=96, n_r=40, integrator="RK4", flux="roeFlux")
    cfg = with_path(cfg, path, 32)
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    s = Solver(build_mesh(raw), cfg)
    h_o, t_o, _ = om.run(25)
    h_g, t_g, _ = s.run(25)
    np.testing.assert_array_equal(h_g, h_o)
    np.testing.assert_array_equal(s.q, om.q[: om.n_interior])


@pytest.mark.parametrize("path", ["fused", "faceTiles", "gradTiles"])
def test_tiled_large_mesh_equals_split(path):
    raw, cfg = cylinder_case(n_theta=512, n_r=256, stretch=1.015, CFL=0.05)
    split = Solver(build_mesh(raw), with_path(cfg, "split"))
    tiled = Solver(build_mesh(raw), with_path(cfg, path))   # automatic tile size
    h_s, _, _ = split.run(12)
    h_f, _, _ = tiled.run(12)
    np.testing.assert_array_equal(h_f, h_s)
    np.testing.assert_array_equal(tiled.q, split.q)
    tf = tiled.time_steps(2)
    ts = split.time_steps(2)
    assert tf["fused_ms"] > 0 and tf["faces_ms"] == 0
    assert ts["fused_ms"] == 0 and ts["faces_ms"] > 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("cells", ["reference", "tiles"])
@pytest.mark.parametrize("korder", ["tilesBackward", "gradientBackward", "bothBackward"])
def test_kernel_order_gives_the_same_bits(korder, cells, precision):
    """Walking the tiles or the cell blocks of each XCD's slice last to first (and the
    device renumbering) changes which workgroup takes which item, not a result: 256 tiles
    (odd slices too) over 15 steps, history, states, gradients and edge states."""
    raw, cfg = cylinder_case(n_theta=200, n_r=84, stretch=1.03, CFL=0.05)
    cfg = copy.deepcopy(cfg)
    cfg["extensions"] = dict(cfg.get("extensions", {}), precision=precision, limiter="venkatakrishnan")
    base = Solver(build_mesh(raw), with_path(cfg, "faceTiles", 128))
    c = with_path(cfg, "faceTiles", 128, cells=cells)
    c["extensions"]["kernelOrder"] = korder
    s = Solver(build_mesh(raw), c)
    h_b, t_b, dt_b = base.run(15)
    h_s, t_s, dt_s = s.run(15)
    np.testing.assert_array_equal(h_s, h_b)
    np.testing.assert_array_equal(s.q_all(), base.q_all())
    assert (t_s, dt_s) == (t_b, dt_b)
    g_b, l_b = base.cell_gradients()
    g_s, l_s = s.cell_gradients()
    np.testing.assert_array_equal(g_s, g_b)
    np.testing.assert_array_equal(l_s, l_b)
    np.testing.assert_array_equal(s.edge_states()[0], base.edge_states()[0])


@pytest.mark.parametrize("path,tile", [("split", -1), ("fused", 5), ("fused", 128), ("faceTiles", 7), ("gradTiles", 7),
                                       ("gradTiles", -1)])
def test_tiled_nan_cell_and_edge_reported_like_split(path, tile):
    raw, cfg = channel_case(integrator="RK2")
    ref = Solver(build_mesh(raw), with_path(cfg, "split"))
    s = Solver(build_mesh(raw), with_path(cfg, path, tile))
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


@pytest.mark.parametrize("path", ["fused", "faceTiles", "gradTiles"])
def test_renumbered_interfaces_match_split(path, tmp_path):
    """Tiled paths renumber the device cells; every interface still speaks the reference numbering."""
    raw, cfg = channel_case(integrator="RK2_TVD", viscosity=True)
    split = Solver(build_mesh(raw), with_path(cfg, "split"))
    tiled = Solver(build_mesh(raw), with_path(cfg, path, 17, "morton", "tiles"))
    q = split.q
    q[:, 0] *= 1.0 + 0.01 * np.sin(np.arange(len(q)))
    split.q = q
    tiled.q = q
    np.testing.assert_array_equal(tiled.q, q)
    for x in (split, tiled):
        x.run(4)
    np.testing.assert_array_equal(tiled.q_all(), split.q_all())
    R_s, nd_s, rm_s = split.ufvm_solver()
    R_t, nd_t, rm_t = tiled.ufvm_solver()
    np.testing.assert_array_equal(R_t, R_s)
    assert nd_t == nd_s and rm_t == rm_s
    for a, b in zip(tiled.cell_gradients(), split.cell_gradients()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(tiled.edge_states(), split.edge_states()):
        np.testing.assert_array_equal(a, b)
    split.save_solution(str(tmp_path / "s.esln"))
    tiled.save_solution(str(tmp_path / "t.esln"))
    assert (tmp_path / "s.esln").read_bytes() == (tmp_path / "t.esln").read_bytes()
    tiled2 = Solver(build_mesh(raw), with_path(cfg, path, 17, "morton", "tiles"))
    tiled2.load_solution(str(tmp_path / "t.esln"))
    np.testing.assert_array_equal(tiled2.q, tiled.q)


def test_quad_mesh_takes_face_tiles_by_default():
    """Quads have ~2.3 faces per cell: the automatic tile size shrinks until a tile's face
    results fit the LDS of three resident workgroups, instead of falling back to the split
    kernels, and the results stay those of the split kernels and the oracle."""
    _, cfg = channel_case(integrator="RK2_TVD", viscosity=True)
    raw = meshgen.channel_bump_quads(200, 60)
    tiled = Solver(build_mesh(raw), cfg)
    split = Solver(build_mesh(raw), with_path(cfg, "split"))
    h_t, _, _ = tiled.run(10)
    h_s, _, _ = split.run(10)
    np.testing.assert_array_equal(h_t, h_s)
    np.testing.assert_array_equal(tiled.q, split.q)
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    h_o, _, _ = om.run(10)
    np.testing.assert_array_equal(h_t, h_o)
    np.testing.assert_array_equal(tiled.q, om.q[: om.n_interior])
    tt = tiled.time_steps(1)
    assert tt["fused_ms"] > 0 and tt["faces_ms"] == 0


@pytest.mark.parametrize("path", ["split", "faceTiles", "fused", "gradTiles"])
@pytest.mark.parametrize("kind", ["tri1", "tri2", "quad1"])
def test_tiny_meshes_match_oracle(kind, path):
    """One or two cells, all on the boundary: a tile shorter than a 16-cell block and an
    RMS sum made of the element tail only."""
    raw = tiny_mesh(kind)
    q = [0.3, 10.0, 1.0, 1.0]
    cfg = with_path(base_config(initialConditions=q, boudaryConditions=[{"tag": "Farfield", "type": "fullState", "q": q}],
                                forceBoundary="Farfield", integrator="RK2_TVD"), path)
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    s = Solver(build_mesh(raw), cfg)
    qp = perturb(om.q[: om.n_interior], 0.05, seed=1)
    qq = om.q
    qq[: om.n_interior] = qp
    om.q = qq
    s.q = qp
    h_o, t_o, _ = om.run(5)
    h_g, t_g, _ = s.run(5)
    np.testing.assert_array_equal(h_g, h_o)
    np.testing.assert_array_equal(s.q, om.q[: om.n_interior])
    assert t_g == t_o


MIXED_CASES = ["cyl_ns_rk2tvd", "cyl_euler_roe_rk4", "rk4_local_dt", "quads_rk2tvd_visc", "pentagons_rk2tvd_visc"]


@pytest.mark.parametrize("limiter", ["scalar", "venkatakrishnan"])
@pytest.mark.parametrize("tile", [-1, 19])
@pytest.mark.parametrize("name", MIXED_CASES)
def test_grad_tiles_equal_face_tiles_mixed_and_venkat(name, tile, limiter):
    """The mixed-precision and Venkatakrishnan forms of the gradient-tile path compute the
    same FP32 (or FP64) gradients and faces as the face-tile path: bit-identical."""
    raw, cfg = CASES[name]()
    for prec in ("mixed", "fp64"):
        c = copy.deepcopy(cfg)
        c["extensions"] = dict(c.get("extensions", {}), precision=prec, limiter=limiter)
        ft = Solver(build_mesh(raw), with_path(c, "faceTiles", tile))
        gt = Solver(build_mesh(raw), with_path(c, "gradTiles", tile))
        h_f, t_f, _ = ft.run(12)
        h_g, t_g, _ = gt.run(12)
        np.testing.assert_array_equal(h_g, h_f)
        np.testing.assert_array_equal(gt.q_all(), ft.q_all())
        assert t_g == t_f
        assert gt.time_steps(1)["gradient_ms"] > 0
