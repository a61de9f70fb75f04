"""Shared test inputs: matching meshes/configs for the HIP path and the oracle."""
import copy
import json
import os

import numpy as np

from ebbcfd_amd import meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MMS_DIR = os.path.join(ROOT, "tests", "golden", "ms_refinement")
FLAT_DIR = os.path.join(ROOT, "tests", "golden", "flatPlate")


def base_config(**over):
    cfg = {
        "mesh": "synthetic.msh", "dt": 1.0e-4, "tEnd": 1.0e9, "flux": "roeFlux", "plotIter": 100, "saveIter": -1,
        "CFL": 0.5, "integrator": "Euler", "forceBoundary": "Wall", "limited": True, "order": 2,
        "initialConditions": [0.5, 0.0, 1.0, 1.0], "viscosity": False, "cflAdjust": True,
        "physicalConfig": {"Pr": 0.7, "R": 8.3145, "gamma": 1.4, "Re": 100.0, "L": 1.0},
        "boudaryConditions": [],
    }
    cfg.update(over)
    return cfg


def channel_case(nx=24, ny=8, height=0.06, seed=3, walls="inviscidWall", M=0.5, aoa=2.0, **over):
    raw = meshgen.channel_bump(nx, ny, height, seed=seed)
    q = [M, aoa, 1.0, 1.0]
    bcs = [{"tag": "Inflow", "type": "fullState", "q": q}, {"tag": "Outflow", "type": "fullState", "q": q},
           {"tag": "Lower", "type": walls, "q": q}, {"tag": "Upper", "type": walls, "q": q}]
    cfg = base_config(initialConditions=q, boudaryConditions=bcs, forceBoundary="Lower", **over)
    return raw, cfg


def cylinder_case(n_theta=64, n_r=32, M=0.2, viscous=True, flux="rusanovFlux", integrator="RK2_TVD", r_far=30.0,
                  stretch=None, **over):
    if stretch is None:
        stretch = 1.006 if n_r >= 512 else 1.08
    raw = meshgen.cylinder_ogrid(n_theta, n_r, r_far=r_far, stretch=stretch)
    q = [M, 0.0, 1.0, 1.0]
    bcs = [{"tag": "Wall", "type": "viscidWall" if viscous else "inviscidWall", "q": q},
           {"tag": "Farfield", "type": "fullState", "q": q}]
    cfg = base_config(initialConditions=q, boudaryConditions=bcs, forceBoundary="Wall", flux=flux,
                      integrator=integrator, viscosity=viscous,
                      physicalConfig={"Pr": 0.7, "R": 8.3145, "gamma": 1.4, "Re": 100.0, "L": 1.0}, **over)
    return raw, cfg


def mms_case(n=16, jitter=0.1, seed=1, integrator="RK4", **over):
    raw = meshgen.mms_square(n, jitter, seed)
    q = [0.1, 0.0, 1.1, 1.0]
    cfg = base_config(initialConditions=q, boudaryConditions=[{"tag": "Freestream", "type": "dirichlet", "q": q}],
                      forceBoundary="Wall", integrator=integrator, dt=1.0e-5, CFL=0.5,
                      physicalConfig={"Pr": 0.71, "R": 1.0, "gamma": 1.4, "Re": 1000.0, "L": 1.0},
                      extensions={"mms": True}, **over)
    return raw, cfg


def oracle_cfg(cfg):
    """The oracle reads the MMS switch at the top level."""
    c = copy.deepcopy(cfg)
    c["mms"] = bool(cfg.get("extensions", {}).get("mms", False))
    return c


def reference_fixture(name):
    from oracle.gmsh import parse_gmsh
    raw = parse_gmsh(os.path.join(MMS_DIR, name + ".msh"))
    cfg = json.load(open(os.path.join(MMS_DIR, name + ".json")))
    cfg["extensions"] = {"mms": True}
    return raw, cfg


def flat_plate_fixture(k=0):
    """The reference's viscous flat-plate case (flatPlate/flatPlate_<k>.msh + .json): quads,
    Roe flux, Euler steps, fullState / viscidWall / constP / symmetry boundaries."""
    from oracle.gmsh import parse_gmsh
    raw = parse_gmsh(os.path.join(FLAT_DIR, "flatPlate_%d.msh" % k))
    cfg = json.load(open(os.path.join(FLAT_DIR, "flatPlate_%d.json" % k)))
    return raw, cfg


def hanging_node_mesh(nc=6, ny=4):
    """Channel [0,2]x[0,1]: nc x ny coarse quads on x < 1, 2nc x 2ny fine quads on x > 1.
    The coarse cells on x = 1 carry the fine column's midpoint as a hanging node, which makes
    them pentagons (5 edges, so the device uses 6 slots per cell).  Tags as channel_case."""
    hc, hy = 1.0 / nc, 1.0 / ny
    nodes = []
    C = {}
    for i in range(nc):                      # coarse columns x < 1
        for j in range(ny + 1):
            C[i, j] = len(nodes)
            nodes.append([i * hc, j * hy])
    Fn = {}
    for i in range(2 * nc + 1):              # fine columns from x = 1
        for j in range(2 * ny + 1):
            Fn[i, j] = len(nodes)
            nodes.append([1.0 + i * hc / 2, j * hy / 2])
    elems = []
    for i in range(nc):
        for j in range(ny):
            if i < nc - 1:
                elems.append([C[i, j], C[i + 1, j], C[i + 1, j + 1], C[i, j + 1]])
            else:
                elems.append([C[i, j], Fn[0, 2 * j], Fn[0, 2 * j + 1], Fn[0, 2 * j + 2], C[i, j + 1]])
    for i in range(2 * nc):
        for j in range(2 * ny):
            elems.append([Fn[i, j], Fn[i + 1, j], Fn[i + 1, j + 1], Fn[i, j + 1]])
    inflow = [[C[0, j + 1], C[0, j]] for j in range(ny)]
    outflow = [[Fn[2 * nc, j], Fn[2 * nc, j + 1]] for j in range(2 * ny)]
    bottom = [[C[i, 0], C[i + 1, 0]] for i in range(nc - 1)] + [[C[nc - 1, 0], Fn[0, 0]]] + \
             [[Fn[i, 0], Fn[i + 1, 0]] for i in range(2 * nc)]
    top = [[C[i + 1, ny], C[i, ny]] for i in range(nc - 1)] + [[Fn[0, 2 * ny], C[nc - 1, ny]]] + \
          [[Fn[i + 1, 2 * ny], Fn[i, 2 * ny]] for i in range(2 * nc)]
    groups = [inflow, outflow, bottom, top]
    offs = np.cumsum([0] + [len(e) for e in elems]).astype(np.uint32)
    starts = np.cumsum([0] + [len(g) for g in groups[:-1]]).astype(np.uint64)
    return meshgen.RawMesh(nodes=np.array(nodes, float), elem_offsets=offs,
                           elem_nodes=np.array([v + 1 for e in elems for v in e], np.uint32),
                           bnodes=np.array([p for g in groups for p in g], np.uint32),
                           btags=["Inflow", "Outflow", "Lower", "Upper"], bgroup_start=starts)


def mixed_mesh(nx=24, ny=8, frac=0.5, seed=5):
    """The channel_bump quads with a seeded random subset (about ``frac``) split into two
    triangles along a random diagonal: triangles and quads side by side, so the device
    uses 4 slots per cell and pads the triangles.  Tags as channel_case."""
    raw = meshgen.channel_bump_quads(nx, ny, seed=seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    quads = raw.elem_nodes.reshape(-1, 4)
    split = rng.uniform(size=quads.shape[0]) < frac
    diag = rng.integers(0, 2, size=quads.shape[0])
    elems = []
    for q, s, dg in zip(quads.tolist(), split, diag):
        if not s:
            elems.append(q)
        elif dg:   # diagonal 0-2
            elems += [[q[0], q[1], q[2]], [q[0], q[2], q[3]]]
        else:      # diagonal 1-3
            elems += [[q[0], q[1], q[3]], [q[1], q[2], q[3]]]
    raw.elem_offsets = np.cumsum([0] + [len(e) for e in elems]).astype(np.uint32)
    raw.elem_nodes = np.array([v for e in elems for v in e], np.uint32)
    return raw


def tiny_mesh(kind):
    """Meshes whose every cell is a boundary cell: one triangle, two triangles, one quad;
    one far-field tag."""
    if kind == "tri1":
        nodes, elems = [[0, 0], [1, 0], [0, 1]], [[1, 2, 3]]
    elif kind == "tri2":
        nodes, elems = [[0, 0], [1, 0], [1, 1], [0, 1]], [[1, 2, 3], [1, 3, 4]]
    else:
        nodes, elems = [[0, 0], [1, 0], [1, 1], [0, 1]], [[1, 2, 3, 4]]
    nn = len(nodes)
    offs = np.cumsum([0] + [len(e) for e in elems]).astype(np.uint32)
    bnodes = np.array([[k, (k + 1) % nn] for k in range(nn)], np.uint32)
    return meshgen.RawMesh(nodes=np.array(nodes, float), elem_offsets=offs,
                           elem_nodes=np.array([v for e in elems for v in e], np.uint32), bnodes=bnodes,
                           btags=["Farfield"], bgroup_start=np.zeros(1, np.uint64))


def perturb(q, scale=0.02, seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    return q * (1.0 + scale * rng.uniform(-1, 1, size=q.shape))
