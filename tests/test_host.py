"""Host side of the C ABI, no GPU needed: the library loads and exports every
symbol include/ebbcfd.h declares; the C++ gmsh reader and geometry precompute
agree bit for bit with the oracle; the config reader keeps the reference's
keys, types and defaults (source/config.d:112-313)."""
import glob
import json
import os
import re

import numpy as np
import pytest

import oracle
from oracle.gmsh import parse_gmsh
from ebbcfd_amd import _lib, build_mesh, import_mesh, load_config, meshgen
from ebbcfd_amd.solver import Config

from helpers import FLAT_DIR, MMS_DIR, ROOT, hanging_node_mesh, mixed_mesh


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "ebbcfd.h")).read()
    declared = set(re.findall(r"\b(ebb_\w+)\s*\(", header))
    assert declared == set(_lib.SYMBOLS), declared ^ set(_lib.SYMBOLS)
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), name


def _declarations(text):
    """name -> parameter count of every `ebb_*` function declared in C or D text."""
    out = {}
    for m in re.finditer(r"\b(ebb_\w+)\s*\(", text):
        depth, k = 1, m.end()
        while depth:   # D writes const(char)*: match the parentheses
            depth += {"(": 1, ")": -1}.get(text[k], 0)
            k += 1
        if not re.match(r"\s*;", text[k:]):
            continue   # a call, not a declaration
        params = text[m.end():k - 1]
        out[m.group(1)] = 0 if params.strip() in ("", "void") else params.count(",") + 1
    return out


def test_d_binding_declares_every_entry_point():
    """The extern(C) D module include/ebbcfd.d names every function of the C ABI with the
    header's parameter count."""
    header = _declarations(open(os.path.join(ROOT, "include", "ebbcfd.h")).read())
    d = _declarations(open(os.path.join(ROOT, "include", "ebbcfd.d")).read())
    assert set(header) == set(_lib.SYMBOLS)
    assert d == header, {k: (header.get(k), d.get(k)) for k in set(header) | set(d) if header.get(k) != d.get(k)}


def _compare_geometry(mesh, om):
    g, og = mesh.cell_geometry(), om.cells()
    N = om.n_interior
    for k in ("area", "d", "perim", "centroid"):
        np.testing.assert_array_equal(g[k], og[k], err_msg=k)
    np.testing.assert_array_equal(g["gradMat"], og["gradMat"][:N])
    np.testing.assert_array_equal(g["edges"], og["edges"][:N])
    np.testing.assert_array_equal(g["fm"], og["fm"][:N].astype(np.int8))
    np.testing.assert_array_equal(g["nbrs"], og["nbrs"][:N])
    e, oe = mesh.edge_geometry(), om.edges()
    for k in e:
        np.testing.assert_array_equal(e[k], oe[k], err_msg=k)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(MMS_DIR, "*.msh")) + glob.glob(os.path.join(FLAT_DIR, "*.msh"))),
                         ids=os.path.basename)
def test_gmsh_reader_and_geometry_match_oracle(path):
    m = import_mesh(path)
    om = oracle.Mesh(parse_gmsh(path))
    assert (m.n_cells, m.n_interior, m.n_edges) == (om.n_cells, om.n_interior, om.n_edges)
    _compare_geometry(m, om)


@pytest.mark.parametrize("maker", ["mms", "cyl", "bump"])
def test_generated_mesh_roundtrip_through_gmsh_file(maker, tmp_path):
    raw = {"mms": lambda: meshgen.mms_square(10, 0.1, 4), "cyl": lambda: meshgen.cylinder_ogrid(32, 12, stretch=1.1),
           "bump": lambda: meshgen.channel_bump(18, 6, 0.05, 2)}[maker]()
    path = str(tmp_path / "m.msh")
    meshgen.write_gmsh(raw, path)
    a, b = import_mesh(path), build_mesh(raw)
    om = oracle.Mesh(raw)
    _compare_geometry(a, om)
    _compare_geometry(b, om)


@pytest.mark.parametrize("maker", ["tri", "quad"])
def test_threaded_build_matches_oracle(maker, tmp_path):
    """Above 65536 cells the geometry precompute runs on several threads and fills
    its arrays without a zero pass; the result must not change by a bit."""
    raw = {"tri": lambda: meshgen.mms_square(190, 0.1, 2),
           "quad": lambda: meshgen.cylinder_ogrid(384, 192, stretch=1.02)}[maker]()
    om = oracle.Mesh(raw)
    assert om.n_interior > 65536
    b = build_mesh(raw)
    assert (b.n_cells, b.n_interior, b.n_edges) == (om.n_cells, om.n_interior, om.n_edges)
    _compare_geometry(b, om)
    path = str(tmp_path / "m.msh")
    meshgen.write_gmsh(raw, path)
    _compare_geometry(import_mesh(path), om)


def test_pentagon_mesh_geometry_matches_oracle():
    """Coarse quads next to a refined column: the interface cells have a hanging node and
    five edges."""
    raw = hanging_node_mesh()
    b = build_mesh(raw)
    om = oracle.Mesh(raw)
    assert b.max_edges == 5
    _compare_geometry(b, om)


def test_mixed_mesh_geometry_matches_oracle():
    """Triangles and quadrilaterals side by side in one mesh."""
    raw = mixed_mesh()
    assert sorted(set(np.diff(raw.elem_offsets).tolist())) == [3, 4]
    b = build_mesh(raw)
    om = oracle.Mesh(raw)
    assert b.max_edges == 4
    _compare_geometry(b, om)


def test_negative_area_rejected():
    raw = meshgen.mms_square(4, 0.0, 1)
    raw.elem_nodes = raw.elem_nodes.reshape(-1, 3)[:, ::-1].reshape(-1).copy()   # clockwise
    with pytest.raises(Exception):
        build_mesh(raw)


def test_c_abi_error_codes(tmp_path):
    """The status codes of include/ebbcfd.h on bad input, through the C ABI itself:
    null handles and pointers (1), unreadable files (2), malformed input (3), meshes that
    buildMesh rejects (4)."""
    import ctypes as C
    L = _lib.lib()
    vp = C.c_void_p()
    d = C.c_double()
    assert L.ebb_mesh_load(None, C.byref(vp)) == 1
    assert L.ebb_mesh_load(str(tmp_path / "missing.msh").encode(), C.byref(vp)) == 2
    (tmp_path / "mesh.txt").write_text("not a mesh")
    assert L.ebb_mesh_load(str(tmp_path / "mesh.txt").encode(), C.byref(vp)) == 3   # neither .msh nor .gri
    (tmp_path / "bad.msh").write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n3\n1 0 0\n")
    assert L.ebb_mesh_load(str(tmp_path / "bad.msh").encode(), C.byref(vp)) == 3     # truncated node block
    assert L.ebb_config_parse(b"{not json", C.byref(vp)) == 3
    assert L.ebb_config_parse(None, C.byref(vp)) == 1
    assert L.ebb_config_get(None, b"CFL", C.byref(d)) == 1
    assert L.ebb_mesh_info(None, (C.c_int64 * 8)()) == 1
    assert L.ebb_mesh_tile_stats(None, 128, 0, (C.c_double * 8)()) == 1
    assert L.ebb_solver_create(None, None, 0, C.byref(vp)) == 1
    for name in ("ebb_state_upload", "ebb_state_download", "ebb_run_reset"):
        assert getattr(L, name)(None, *([None] if name != "ebb_run_reset" else [0.0, 0.0])) == 1
    assert L.ebb_history(None, 0, 0, None) == 1
    # batch arguments are checked before any device call: odd graph lengths, and graphs
    # longer than the 256-row history ring a run is cut to
    for steps in (-2, 3, 258):
        assert L.ebb_batch_create(0, C.c_int64(1 << 30), steps, 0, C.byref(vp)) == 1
    assert L.ebb_server_create(0, C.c_int64(1 << 30), 0, 100, 8, 0, C.byref(vp)) == 1
    assert L.ebb_host_alloc(0, C.byref(vp)) == 1 and not vp.value   # nothing to allocate
    assert L.ebb_host_alloc(16, None) == 1
    L.ebb_host_free(None)                                           # a no-op

    def create(nodes, eoff, enodes, bnodes):
        nodes = np.ascontiguousarray(nodes, np.float64)
        eoff = np.ascontiguousarray(eoff, np.uint32)
        enodes = np.ascontiguousarray(enodes, np.uint32)
        bnodes = np.ascontiguousarray(bnodes, np.uint32).reshape(-1, 2)
        tags = (C.c_char_p * 1)(b"Farfield")
        starts = np.zeros(1, np.uint64)
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        h = C.c_void_p()
        rc = L.ebb_mesh_create(nodes.shape[0], P(nodes, C.c_double), eoff.size - 1, P(eoff, C.c_uint32),
                               P(enodes, C.c_uint32), bnodes.shape[0], P(bnodes, C.c_uint32), 1, tags,
                               P(starts, C.c_uint64), C.byref(h))
        if rc == 0:
            L.ebb_mesh_destroy(h)
        return rc

    square = [[0, 0], [1, 0], [1, 1], [0, 1]]
    ring = [[0, 1], [1, 2], [2, 3], [3, 0]]
    assert create(square, [0, 3, 6], [1, 2, 3, 1, 3, 4], ring) == 0
    assert create(square, [0, 3, 6], [1, 2, 3, 1, 3, 5], ring) == 4      # node 5 of 4 (1-based)
    assert create(square, [0, 3, 6], [1, 2, 3, 1, 3, 0], ring) == 4      # node 0 (elements are 1-based)
    assert create(square, [0, 3, 6], [1, 2, 3, 1, 3, 4], [[0, 1], [1, 7]]) == 4   # boundary node out of range
    assert create(square, [0, 2, 6], [1, 2, 3, 1, 3, 4], ring) == 4      # a two-node element
    assert create(square, [0, 3, 6], [3, 2, 1, 1, 3, 4], ring) == 4      # clockwise: negative area


REF_CFG = json.load(open(os.path.join(MMS_DIR, "ms_refine_tri_0.json")))


def test_config_reference_fixture():
    c = load_config(json.dumps(REF_CFG))
    assert c["dt"] == 1e-5 and c["tEnd"] == 10000 and c["CFL"] == 0.2 and c["lpThresh"] == 1
    assert c["limited"] == 1 and c["order"] == 2 and c["viscosity"] == 0 and c["cflAdjust"] == 0
    assert c["integrator"] == 0 and c["flux"] == 1 and c["dynamicDt"] == 1 and c["localTimestep"] == 0
    assert c["Pr"] == 0.71 and c["R"] == 1 and c["Re"] == 1000
    assert c.mesh_file == "ms_refine_tri_0.msh"


def test_config_defaults():
    minimal = {k: REF_CFG[k] for k in ("mesh", "dt", "tEnd", "flux", "plotIter", "saveIter", "forceBoundary",
                                       "initialConditions", "boudaryConditions")}
    c = load_config(json.dumps(minimal))
    # config.d:137-260
    assert c["CFL"] == 0.5 and c["dynamicDt"] == 1 and c["integrator"] == 0 and c["limited"] == 1
    assert c["localTimestep"] == 0 and c["lpThresh"] == -1 and c["order"] == 2 and c["multistageLimiting"] == 1
    assert c["cflAdjust"] == 1 and c["viscosity"] == 0
    assert c["Pr"] == 0.7 and c["R"] == 8.3145 and c["Re"] == 1000 and c["L"] == 1


def test_config_type_rules():
    bad = dict(REF_CFG)
    bad["dt"] = 1          # dt must be a JSON float (config.d:119)
    with pytest.raises(Exception):
        load_config(json.dumps(bad))
    bad = dict(REF_CFG)
    bad["saveIter"] = 10.0
    with pytest.raises(Exception):
        load_config(json.dumps(bad))
    bad = dict(REF_CFG)
    del bad["boudaryConditions"]
    bad["boundaryConditions"] = REF_CFG["boudaryConditions"]     # the reference's key is misspelt
    with pytest.raises(Exception):
        load_config(json.dumps(bad))
    odd = dict(REF_CFG)
    odd["order"] = 3       # invalid order falls back to 2 (config.d:200-204)
    assert load_config(json.dumps(odd))["order"] == 2


def test_config_extension_keys():
    """The extension keys this build adds: parsed into the config, unknown values refused."""
    c = dict(REF_CFG)
    c["extensions"] = {"precision": "mixed", "limiter": "venkatakrishnan", "venkatK": 2.5, "rmsDrop": 1e-8,
                       "stagePath": "gradTiles", "maxSteps": 20000, "kernelOrder": "tilesBackward"}
    cfg = load_config(json.dumps(c))
    assert cfg["precision"] == 1 and cfg["limiter"] == 1 and cfg["venkatK"] == 2.5 and cfg["rmsDrop"] == 1e-8
    assert cfg["stagePath"] == 3 and cfg["maxSteps"] == 20000 and cfg["kernelOrder"] == 1
    for name, number in (("forward", 0), ("tilesBackward", 1), ("gradientBackward", 2), ("bothBackward", 3)):
        c["extensions"] = {"kernelOrder": name}
        assert load_config(json.dumps(c))["kernelOrder"] == number
    plain = load_config(json.dumps(REF_CFG))
    assert plain["precision"] == 0 and plain["limiter"] == 0 and plain["rmsDrop"] == 0 and plain["stagePath"] == 1
    assert plain["kernelOrder"] == 0
    # enumerated keys set by number take only their values' numbers
    plain["kernelOrder"] = 3
    assert plain["kernelOrder"] == 3
    for key, value in (("kernelOrder", 4), ("stagePath", 4), ("precision", 2), ("limiter", -1), ("stagePath", 1.5)):
        with pytest.raises(Exception):
            plain[key] = value
    for key, value in (("precision", "fp16"), ("limiter", "minmod"), ("stagePath", "tiles"), ("kernelOrder", "reverse")):
        bad = dict(REF_CFG)
        bad["extensions"] = {key: value}
        with pytest.raises(Exception):
            load_config(json.dumps(bad))


def test_config_set_get_roundtrip():
    c = Config(REF_CFG)
    c["CFL"] = 0.3
    c["maxSteps"] = 17
    assert c["CFL"] == 0.3 and c["maxSteps"] == 17


def test_solver_needs_a_device_or_fails_loudly():
    """No silent CPU fallback: without a HIP device the solver refuses to start."""
    import ctypes
    if ctypes.CDLL("libamdhip64.so").hipGetDeviceCount is None:  # pragma: no cover
        pytest.skip("no HIP runtime")
    from ebbcfd_amd import Solver
    n = ctypes.c_int(0)
    ctypes.CDLL("libamdhip64.so").hipGetDeviceCount(ctypes.byref(n))
    if n.value > 0:
        pytest.skip("a device is present")
    raw = meshgen.mms_square(4, 0.0, 1)
    cfg = dict(REF_CFG, extensions={"mms": True})
    with pytest.raises(Exception) as e:
        Solver(build_mesh(raw), cfg)
    assert "device" in str(e.value)


@pytest.mark.parametrize("q", [1, 2])
def test_gri_reader_matches_oracle(q, tmp_path):
    from oracle.gmsh import parse_gri
    raw = meshgen.channel_bump(14, 5, 0.05, 3)
    path = str(tmp_path / "m.gri")
    meshgen.write_gri(raw, path, q=q)
    m = import_mesh(path)
    ro = parse_gri(path)
    om = oracle.Mesh(ro)
    _compare_geometry(m, om)
    # same geometry as the mesh it was written from (q=2 only adds edge midpoints)
    ob = oracle.Mesh(raw)
    np.testing.assert_array_equal(om.cells()["area"], ob.cells()["area"])


def test_host_mms_values_equal_the_oracle():
    """The manufactured solution and source the library precomputes on the host (for the
    device's Dirichlet ghosts, edge states and per-cell source) equal the oracle's bit for
    bit on the 8,192 jittered centroids of the 64 x 64 square.  Both call sin and cos
    separately, as the reference does; a sincos fusion by either compiler changes the last
    bit of cos for some of these arguments."""
    import ctypes as C
    from ebbcfd_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    src = L._ZN3ebb13mmsSourceTermEddbdddPd
    src.argtypes = [C.c_double, C.c_double, C.c_bool, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
    sol = L._ZN3ebb11mmsSolutionEddPd
    sol.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double)]
    raw = meshgen.mms_square(64, 0.1, 1)
    m = oracle.Mesh(raw)
    cen = m.cells()["centroid"]
    out = (C.c_double * 4)()
    bad = 0
    for x, y in cen:
        src(x, y, False, 1.0, 0.0, 0.71, out)
        bad += list(out) != oracle.source_term(x, y, viscosity=False).tolist()
        src(x, y, True, 1.0, 0.01, 0.71, out)
        bad += list(out) != oracle.source_term(x, y, viscosity=True, R=1.0, mu=0.01, Pr=0.71).tolist()
        sol(x, y, out)
        bad += list(out) != oracle.solution(x, y).tolist()
    assert bad == 0
