"""Synthetic mesh generators for the solver's workloads, and a gmsh 2.2 writer.

Every generator returns a :class:`RawMesh` -- the arrays the reference's mesh
reader produces (source/io.d:689-990) and ``buildMesh`` consumes
(source/mesh.d:217-643): node coordinates, 1-based counter-clockwise elements,
0-based boundary node pairs grouped by physical tag, and the tag names.

Workloads:
  * ``mms_square``      -- [0,1]^2, n x n quads split SW-NE into 2 triangles,
                           interior nodes jittered by U(-j h, j h), tag "Freestream";
  * ``cylinder_ogrid``  -- O-grid around a cylinder of radius 0.5 out to R=30,
                           geometric radial stretching, each quad split along a
                           seeded random diagonal; tags "Wall" and "Farfield";
  * ``channel_bump``    -- channel [0,3]x[0,1] with a sin^2 bump of height h on
                           the lower wall between x=1 and x=2; tags "Inflow",
                           "Outflow", "Lower", "Upper" (``channel_bump_quads``: the
                           same domain in quadrilaterals).
Random numbers come from numpy's PCG64 seeded with ``seed``.
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class RawMesh:
    nodes: np.ndarray          # (n, 2) float64
    elem_offsets: np.ndarray   # (ne + 1,) uint32
    elem_nodes: np.ndarray     # uint32, 1-based node ids, counter-clockwise
    bnodes: np.ndarray         # (nb, 2) uint32, 0-based
    btags: list = field(default_factory=list)
    bgroup_start: np.ndarray = None  # (ntags,) uint64

    @property
    def n_elements(self):
        return int(self.elem_offsets.size - 1)


def _tri_mesh(nodes, tris, groups, names):
    """tris: (ne,3) 0-based, re-oriented counter-clockwise here; groups: list of (nb_g,2) 0-based pairs."""
    ne = tris.shape[0]
    p = nodes[tris]
    cross = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0])
    tris = np.where((cross < 0)[:, None], tris[:, ::-1], tris)
    offs = (np.arange(ne + 1, dtype=np.uint64) * 3).astype(np.uint32)
    bn = np.concatenate(groups).astype(np.uint32) if groups else np.zeros((0, 2), np.uint32)
    starts = np.zeros(len(groups), np.uint64)
    acc = 0
    for i, g in enumerate(groups):
        starts[i] = acc
        acc += g.shape[0]
    return RawMesh(nodes=np.ascontiguousarray(nodes, dtype=np.float64), elem_offsets=offs,
                   elem_nodes=(tris.astype(np.uint32) + 1).reshape(-1), bnodes=bn, btags=list(names),
                   bgroup_start=starts)


def mms_square(n=64, jitter=0.1, seed=1):
    """n x n quads on [0,1]^2, each split along its SW-NE diagonal (2 n^2 triangles)."""
    h = 1.0 / n
    xs = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(xs, xs)           # row j = y index, column i = x index
    X, Y = X.copy(), Y.copy()
    if jitter > 0:
        rng = np.random.Generator(np.random.PCG64(seed))
        d = rng.uniform(-jitter * h, jitter * h, size=(n - 1, n - 1, 2))
        X[1:-1, 1:-1] += d[:, :, 0]
        Y[1:-1, 1:-1] += d[:, :, 1]
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    idx = np.arange((n + 1) * (n + 1)).reshape(n + 1, n + 1)
    sw, se, ne_, nw = idx[:-1, :-1], idx[:-1, 1:], idx[1:, 1:], idx[1:, :-1]
    t1 = np.stack([sw, se, ne_], axis=-1).reshape(-1, 3)
    t2 = np.stack([sw, ne_, nw], axis=-1).reshape(-1, 3)
    tris = np.stack([t1, t2], axis=1).reshape(-1, 3)
    bottom = np.stack([idx[0, :-1], idx[0, 1:]], axis=1)
    right = np.stack([idx[:-1, -1], idx[1:, -1]], axis=1)
    top = np.stack([idx[-1, 1:], idx[-1, :-1]], axis=1)[::-1]
    left = np.stack([idx[1:, 0], idx[:-1, 0]], axis=1)[::-1]
    return _tri_mesh(nodes, tris, [np.concatenate([bottom, right, top, left])], ["Freestream"])


def ogrid_radii(n_r=1024, r0=0.5, r_far=30.0, stretch=1.006):
    dr0 = (r_far - r0) * (stretch - 1.0) / (stretch ** n_r - 1.0)
    j = np.arange(n_r + 1, dtype=np.float64)
    r = r0 + dr0 * (stretch ** j - 1.0) / (stretch - 1.0)
    r[-1] = r_far
    return r


def cylinder_ogrid(n_theta=2048, n_r=1024, r0=0.5, r_far=30.0, stretch=1.006, seed=1):
    """O-grid of n_theta x n_r quads around a cylinder, 2 triangles per quad (4,194,304 at the defaults)."""
    r = ogrid_radii(n_r, r0, r_far, stretch)
    th = 2.0 * np.pi * np.arange(n_theta) / n_theta
    R, T = np.meshgrid(r, th, indexing="ij")           # (n_r+1, n_theta)
    nodes = np.stack([(R * np.cos(T)).ravel(), (R * np.sin(T)).ravel()], axis=1)
    idx = np.arange((n_r + 1) * n_theta).reshape(n_r + 1, n_theta)
    ip = np.roll(idx, -1, axis=1)
    a, b, c, d = idx[:-1], ip[:-1], ip[1:], idx[1:]     # (r_j,th_i) (r_j,th_i+1) (r_j+1,th_i+1) (r_j+1,th_i): CCW
    rng = np.random.Generator(np.random.PCG64(seed))
    diag = rng.integers(0, 2, size=a.shape).astype(bool)
    t1 = np.where(diag[..., None], np.stack([a, b, c], -1), np.stack([a, b, d], -1))
    t2 = np.where(diag[..., None], np.stack([a, c, d], -1), np.stack([b, c, d], -1))
    tris = np.stack([t1, t2], axis=2).reshape(-1, 3)
    wall = np.stack([ip[0], idx[0]], axis=1)            # traversed clockwise: outward normal points into the body
    far = np.stack([idx[-1], ip[-1]], axis=1)
    return _tri_mesh(nodes, tris, [wall, far], ["Wall", "Farfield"])


def channel_bump(nx=96, ny=32, height=0.05, seed=1, jitter=0.1):
    """Channel [0,3]x[0,1] with y_low(x) = height*sin^2(pi(x-1)) on 1<=x<=2; 2*nx*ny triangles."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xs = np.linspace(0.0, 3.0, nx + 1)
    s = np.linspace(0.0, 1.0, ny + 1)
    X, S = np.meshgrid(xs, s)
    yl = np.where((X >= 1.0) & (X <= 2.0), height * np.sin(np.pi * (X - 1.0)) ** 2, 0.0)
    Y = yl + S * (1.0 - yl)
    if jitter > 0:
        hx, hy = 3.0 / nx, 1.0 / ny
        X[1:-1, 1:-1] += rng.uniform(-jitter * hx, jitter * hx, size=(ny - 1, nx - 1)) * (1.0 - yl[1:-1, 1:-1])
        Y[1:-1, 1:-1] += rng.uniform(-jitter * hy, jitter * hy, size=(ny - 1, nx - 1)) * (1.0 - yl[1:-1, 1:-1])
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    idx = np.arange((nx + 1) * (ny + 1)).reshape(ny + 1, nx + 1)
    sw, se, ne_, nw = idx[:-1, :-1], idx[:-1, 1:], idx[1:, 1:], idx[1:, :-1]
    diag = rng.integers(0, 2, size=sw.shape).astype(bool)
    t1 = np.where(diag[..., None], np.stack([sw, se, ne_], -1), np.stack([sw, se, nw], -1))
    t2 = np.where(diag[..., None], np.stack([sw, ne_, nw], -1), np.stack([se, ne_, nw], -1))
    tris = np.stack([t1, t2], axis=2).reshape(-1, 3)
    lower = np.stack([idx[0, :-1], idx[0, 1:]], axis=1)
    outflow = np.stack([idx[:-1, -1], idx[1:, -1]], axis=1)
    upper = np.stack([idx[-1, 1:], idx[-1, :-1]], axis=1)
    inflow = np.stack([idx[1:, 0], idx[:-1, 0]], axis=1)
    return _tri_mesh(nodes, tris, [inflow, outflow, lower, upper], ["Inflow", "Outflow", "Lower", "Upper"])


def channel_bump_quads(nx=96, ny=32, height=0.05, seed=1, jitter=0.1):
    """The channel_bump domain as nx*ny quadrilaterals (no diagonal split)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xs = np.linspace(0.0, 3.0, nx + 1)
    s = np.linspace(0.0, 1.0, ny + 1)
    X, S = np.meshgrid(xs, s)
    yl = np.where((X >= 1.0) & (X <= 2.0), height * np.sin(np.pi * (X - 1.0)) ** 2, 0.0)
    Y = yl + S * (1.0 - yl)
    if jitter > 0:
        hx, hy = 3.0 / nx, 1.0 / ny
        X[1:-1, 1:-1] += rng.uniform(-jitter * hx, jitter * hx, size=(ny - 1, nx - 1)) * (1.0 - yl[1:-1, 1:-1])
        Y[1:-1, 1:-1] += rng.uniform(-jitter * hy, jitter * hy, size=(ny - 1, nx - 1)) * (1.0 - yl[1:-1, 1:-1])
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    idx = np.arange((nx + 1) * (ny + 1)).reshape(ny + 1, nx + 1)
    quads = np.stack([idx[:-1, :-1], idx[:-1, 1:], idx[1:, 1:], idx[1:, :-1]], axis=-1).reshape(-1, 4)  # counter-clockwise
    groups = [np.stack([idx[1:, 0], idx[:-1, 0]], axis=1), np.stack([idx[:-1, -1], idx[1:, -1]], axis=1),
              np.stack([idx[0, :-1], idx[0, 1:]], axis=1), np.stack([idx[-1, 1:], idx[-1, :-1]], axis=1)]
    ne = quads.shape[0]
    starts = np.cumsum([0] + [g.shape[0] for g in groups[:-1]]).astype(np.uint64)
    return RawMesh(nodes=np.ascontiguousarray(nodes, dtype=np.float64),
                   elem_offsets=(np.arange(ne + 1, dtype=np.uint64) * 4).astype(np.uint32),
                   elem_nodes=(quads.astype(np.uint32) + 1).reshape(-1),
                   bnodes=np.concatenate(groups).astype(np.uint32), btags=["Inflow", "Outflow", "Lower", "Upper"],
                   bgroup_start=starts)


def channel_bump_cells(target_cells, aspect=3.0):
    """(nx, ny) so that 2*nx*ny is close to target_cells with nx/ny ~ aspect."""
    ny = max(2, int(round(np.sqrt(target_cells / (2.0 * aspect)))))
    nx = max(2, int(round(target_cells / (2.0 * ny))))
    return nx, ny


def write_gmsh(raw, path):
    """Write a RawMesh as gmsh 2.2 ASCII; physical tags are numbered 1..ntags in btags order
    (the numbering the reader assumes, source/io.d:939-947)."""
    nodes = np.asarray(raw.nodes)
    ntag = len(raw.btags)
    starts = list(np.asarray(raw.bgroup_start, dtype=np.int64)) + [raw.bnodes.shape[0]]
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
        f.write("$PhysicalNames\n%d\n" % (ntag + 1))
        for i, t in enumerate(raw.btags):
            f.write('1 %d "%s"\n' % (i + 1, t))
        f.write('2 %d "Flow"\n$EndPhysicalNames\n' % (ntag + 1))
        f.write("$Nodes\n%d\n" % nodes.shape[0])
        lines = ["%d %.17g %.17g 0\n" % (i + 1, x, y) for i, (x, y) in enumerate(nodes)]
        f.write("".join(lines))
        f.write("$EndNodes\n")
        ne = raw.n_elements
        nb = raw.bnodes.shape[0]
        f.write("$Elements\n%d\n" % (nb + ne))
        k = 1
        out = []
        for g in range(ntag):
            for b in range(starts[g], starts[g + 1]):
                out.append("%d 1 2 %d %d %d %d\n" % (k, g + 1, g + 1, raw.bnodes[b, 0] + 1, raw.bnodes[b, 1] + 1))
                k += 1
        offs = raw.elem_offsets
        for e in range(ne):
            en = raw.elem_nodes[offs[e]:offs[e + 1]]
            et = 2 if en.size == 3 else 3
            out.append("%d %d 2 %d %d %s\n" % (k, et, ntag + 1, ntag + 1, " ".join(str(int(v)) for v in en)))
            k += 1
        f.write("".join(out))
        f.write("$EndElements\n")


def write_gri(raw, path, q=1):
    """Write a single-element-type RawMesh as an XFlow .gri file (the layout read by
    source/io.d:992-1121).  q=2 writes quadratic triangles with edge-midpoint nodes."""
    nodes = np.asarray(raw.nodes)
    offs = np.asarray(raw.elem_offsets)
    nv = np.diff(offs)
    assert np.all(nv == nv[0]), ".gri needs one element type"
    el = np.asarray(raw.elem_nodes).reshape(-1, int(nv[0]))
    extra = []
    if q == 2:
        assert nv[0] == 3
        mids = {}
        base = nodes.shape[0]
        hi = []
        for t in el:
            a, b, c = (int(v) for v in t)
            def mid(u, w):
                key = (min(u, w), max(u, w))
                if key not in mids:
                    mids[key] = base + len(extra) + 1
                    extra.append(0.5 * (nodes[u - 1] + nodes[w - 1]))
                return mids[key]
            # q=2 triangle node order: corners at positions 0, 2, 5
            hi.append([a, mid(a, b), b, mid(a, c), mid(b, c), c])
        el = np.array(hi)
    allnodes = np.vstack([nodes] + ([np.array(extra)] if extra else []))
    starts = list(np.asarray(raw.bgroup_start, dtype=np.int64)) + [raw.bnodes.shape[0]]
    kind = "TriLagrange" if nv[0] == 3 else "QuadLagrange"
    with open(path, "w") as f:
        f.write("%d %d 2\n" % (allnodes.shape[0], el.shape[0]))
        f.write("".join("%.17g %.17g\n" % (x, y) for x, y in allnodes))
        f.write("%d\n" % len(raw.btags))
        for g, tag in enumerate(raw.btags):
            f.write("%d 2 %s\n" % (starts[g + 1] - starts[g], tag))
            for b in range(starts[g], starts[g + 1]):
                f.write("%d %d\n" % (raw.bnodes[b, 0] + 1, raw.bnodes[b, 1] + 1))
        f.write("%d %d %s\n" % (el.shape[0], q, kind))
        f.write("".join(" ".join(str(int(v)) for v in row) + "\n" for row in el))
