"""CPU oracle for the EbbCFD explicit residual-and-update path.

TEST INFRASTRUCTURE ONLY: this package is the checker.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it; the product path (``ebbcfd_amd``) never does.

``ebb_oracle.c`` restates the reference in its own edge-loop form (citations in
that file); this module is a ctypes binding plus the restated gmsh parser
(``oracle.gmsh``).  Pinning of each function (known answers, invariants,
closed forms) lives in ``tests/test_oracle_*.py``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

BT = {"fullState": 0, "constP": 1, "inviscidWall": 2, "viscidWall": 3,
      "symmetry": 4, "tpInflow": 5, "dirichlet": 6}
FLUX = {"rusanovFlux": 0, "roeFlux": 1}
INTEGRATOR = {"Euler": 0, "RK2": 1, "RK2_TVD": 2, "RK4": 3}
HIST_FIELDS = ("t", "dt", "rho_rms", "u_rms", "v_rms", "E_rms", "Rmax", "drag", "lift", "CFL")

ERR_NAMES = {0: "ok", 1: "NaN flux on edge", 2: "NaN state in cell", 3: "boundary condition", 4: "argument", 5: "mesh"}


class OrcConfig(C.Structure):
    _fields_ = [
        ("flux", C.c_int), ("integrator", C.c_int), ("order", C.c_int), ("limited", C.c_int),
        ("localTimestep", C.c_int), ("multistageLimiting", C.c_int), ("cflAdjust", C.c_int),
        ("viscosity", C.c_int), ("dynamicDt", C.c_int), ("mms", C.c_int),
        ("dt", C.c_double), ("tEnd", C.c_double), ("CFL", C.c_double), ("lpThresh", C.c_double),
        ("ic", C.c_double * 4),
        ("Pr", C.c_double), ("R", C.c_double), ("Re", C.c_double), ("L", C.c_double), ("mu", C.c_double),
        ("nBC", C.c_int), ("bcType", C.c_int * 32), ("bcData", (C.c_double * 4) * 32), ("bcNData", C.c_int * 32),
        ("tagToBc", C.c_int * 32), ("nForce", C.c_int), ("forceTags", C.c_int * 32),
        ("limiter", C.c_int), ("venkatK", C.c_double),
    ]


def set_threads(n):
    """OpenMP threads of the residual and stage loops (results do not depend on it)."""
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(os.path.join(_HERE, "ebb_oracle.c")):
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        d, u32, i64 = C.c_double, C.c_uint32, C.c_int64
        P = C.POINTER
        L.orc_mesh_build.restype = C.c_void_p
        L.orc_mesh_build.argtypes = [u32, P(d), u32, P(u32), P(u32), u32, P(u32), u32, P(C.c_uint64)]
        L.orc_mesh_free.argtypes = [C.c_void_p]
        L.orc_mesh_counts.argtypes = [C.c_void_p, P(i64)]
        L.orc_mesh_cells.argtypes = [C.c_void_p, P(d), P(d), P(d), P(d), P(d), P(u32), P(d), P(u32), P(u32)]
        L.orc_mesh_edges.argtypes = [C.c_void_p, P(u32), P(u32), P(d), P(d), P(d), P(C.c_int32)]
        L.orc_boundary_group.argtypes = [C.c_void_p, u32, P(u32), P(u32)]
        L.orc_get_q.argtypes = [C.c_void_p, P(d)]
        L.orc_set_q.argtypes = [C.c_void_p, P(d)]
        L.orc_get_edge_data.argtypes = [C.c_void_p, P(d), P(d), P(d)]
        L.orc_get_cell_data.argtypes = [C.c_void_p, P(d), P(d), P(d)]
        L.orc_last_error.argtypes = [C.c_void_p, P(C.c_int), P(i64)]
        L.orc_setup.argtypes = [C.c_void_p, P(OrcConfig)]
        L.orc_residual.argtypes = [C.c_void_p, P(OrcConfig), P(d), P(d), P(d), C.c_int, C.c_int]
        L.orc_step.argtypes = [C.c_void_p, P(OrcConfig), P(d), P(d)]
        L.orc_run.argtypes = [C.c_void_p, P(OrcConfig), i64, P(d), P(d), P(i64)]
        for name in ("orc_roe_flux", "orc_rusanov_flux"):
            getattr(L, name).argtypes = [P(d), P(d), P(d), P(d), P(d)]
        L.orc_convective_flux.argtypes = [d, d, d, d, d, P(d), P(d)]
        L.orc_diffusive_flux.argtypes = [d, d, P(d), P(d), P(d), P(d)]
        L.orc_solution.argtypes = [d, d, P(d)]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_solution_gradient.argtypes = [d, d, P(d)]
        L.orc_source_term.argtypes = [d, d, C.c_int, d, d, d, P(d)]
        L.orc_limiter1d.restype = d
        L.orc_limiter1d.argtypes = [C.c_int, d, d]
        L.orc_pairwise_sum.restype = d
        L.orc_pairwise_sum.argtypes = [P(d), C.c_uint64]
        L.orc_get_pressure.restype = d
        L.orc_get_pressure.argtypes = [P(d)]
        L.orc_lp.argtypes = [C.c_int, P(d), P(d), P(d), P(d)]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- pointwise functions
def roe_flux(qL, qR, n):
    qL, qR, n = _f64(qL), _f64(qR), _f64(n)
    out = np.zeros(4)
    s = np.zeros(1)
    lib().orc_roe_flux(_p(qL, C.c_double), _p(qR, C.c_double), _p(n, C.c_double), _p(out, C.c_double), _p(s, C.c_double))
    return out, s[0]


def rusanov_flux(qL, qR, n):
    qL, qR, n = _f64(qL), _f64(qR), _f64(n)
    out = np.zeros(4)
    s = np.zeros(1)
    lib().orc_rusanov_flux(_p(qL, C.c_double), _p(qR, C.c_double), _p(n, C.c_double), _p(out, C.c_double), _p(s, C.c_double))
    return out, s[0]


def convective_flux(p, u, v, rho, e, n):
    n = _f64(n)
    out = np.zeros(4)
    lib().orc_convective_flux(p, u, v, rho, e, _p(n, C.c_double), _p(out, C.c_double))
    return out


def diffusive_flux(Pr, mu, q, dq, n):
    q, dq, n = _f64(q), _f64(dq).reshape(8), _f64(n)
    out = np.zeros(4)
    lib().orc_diffusive_flux(Pr, mu, _p(q, C.c_double), _p(dq, C.c_double), _p(n, C.c_double), _p(out, C.c_double))
    return out


def solution(x, y):
    out = np.zeros(4)
    lib().orc_solution(x, y, _p(out, C.c_double))
    return out


def solution_gradient(x, y):
    out = np.zeros(8)
    lib().orc_solution_gradient(x, y, _p(out, C.c_double))
    return out.reshape(4, 2)


def source_term(x, y, viscosity=False, R=1.0, mu=0.0, Pr=0.72):
    out = np.zeros(4)
    lib().orc_source_term(x, y, int(viscosity), R, mu, Pr, _p(out, C.c_double))
    return out


def get_pressure(q):
    q = _f64(q)
    return lib().orc_get_pressure(_p(q, C.c_double))


def limiter1d(kind, Sp, Sm):
    return lib().orc_limiter1d(kind, Sp, Sm)


def pairwise_sum(x):
    x = _f64(x)
    return lib().orc_pairwise_sum(_p(x, C.c_double), x.size)


def lp(A, b, c):
    A, b, c = _f64(A), _f64(b), _f64(c)
    xk = np.zeros(2)
    lib().orc_lp(A.shape[0], _p(A, C.c_double), _p(b, C.c_double), _p(c, C.c_double), _p(xk, C.c_double))
    return xk


# ---------------------------------------------------------------- config
def make_config(cfg, mesh_tags):
    """Translate a reference JSON config (dict, reference keys) into OrcConfig.

    Defaults follow source/config.d:112-313; the MMS flag ("mms") selects the
    manufactured-solution Dirichlet functions and per-stage source term
    (app/ms.d:131-141, 285-298)."""
    c = OrcConfig()
    c.flux = FLUX[cfg["flux"]]
    c.integrator = INTEGRATOR[cfg.get("integrator", "Euler")]
    order = int(cfg.get("order", 2))
    c.order = order if order in (1, 2) else 2
    c.limited = int(bool(cfg.get("limited", True)))
    c.localTimestep = int(bool(cfg.get("localTimestep", False)))
    c.multistageLimiting = int(bool(cfg.get("multistageLimiting", True)))
    c.cflAdjust = int(bool(cfg.get("cflAdjust", True)))
    c.viscosity = int(bool(cfg.get("viscosity", False)))
    c.dynamicDt = int(bool(cfg.get("dynamicDt", True)))
    c.mms = int(bool(cfg.get("mms", False)))
    c.dt = float(cfg["dt"])
    c.tEnd = float(cfg["tEnd"])
    c.CFL = float(cfg.get("CFL", 0.5))
    c.lpThresh = float(cfg.get("lpThresh", -1))
    ext = cfg.get("extensions", {})
    c.limiter = {"scalar": 0, "venkatakrishnan": 1}[ext.get("limiter", "scalar")]
    c.venkatK = float(ext.get("venkatK", 1.0))
    for i in range(4):
        c.ic[i] = float(cfg["initialConditions"][i])
    pc = cfg.get("physicalConfig")
    if pc is None:
        pc = {"Pr": 0.7, "R": 8.3145, "gamma": 1.4, "Re": 1000, "L": 1}
    c.Pr, c.R, c.Re, c.L = float(pc["Pr"]), float(pc["R"]), float(pc["Re"]), float(pc["L"])
    bcs = cfg["boudaryConditions"]
    c.nBC = len(bcs)
    for i, bc in enumerate(bcs):
        c.bcType[i] = BT[bc["type"]]
        data = [float(v) for v in bc["q"]]
        c.bcNData[i] = len(data)
        for k, v in enumerate(data[:4]):
            c.bcData[i][k] = v
    tags = [bc["tag"] for bc in bcs]
    for t, name in enumerate(mesh_tags):
        c.tagToBc[t] = tags.index(name) if name in tags else -1
    fb = cfg.get("forceBoundary", [])
    if isinstance(fb, str):
        fb = [fb]
    c.nForce = len(fb)
    for i, name in enumerate(fb):
        c.forceTags[i] = list(mesh_tags).index(name) if name in mesh_tags else -1
    return c


# ---------------------------------------------------------------- mesh + solver
class OracleError(RuntimeError):
    def __init__(self, code, kind, index):
        super().__init__("oracle error %d (%s) at index %d" % (code, ERR_NAMES.get(kind, "?"), index))
        self.code, self.kind, self.index = code, kind, index


class Mesh:
    """buildMesh (source/mesh.d:217-643) over raw arrays:
    nodes (n,2) f64, elem_offsets (ne+1) u32, elem_nodes u32 1-based,
    bnodes (nb,2) u32 0-based, btags list[str], bgroup_start (ntags) u64."""

    def __init__(self, raw):
        L = lib()
        self.raw = raw
        nodes = _f64(raw.nodes).reshape(-1, 2)
        eoff = np.ascontiguousarray(raw.elem_offsets, dtype=np.uint32)
        enodes = np.ascontiguousarray(raw.elem_nodes, dtype=np.uint32)
        bnodes = np.ascontiguousarray(raw.bnodes, dtype=np.uint32).reshape(-1, 2)
        bgs = np.ascontiguousarray(raw.bgroup_start, dtype=np.uint64)
        self.h = L.orc_mesh_build(nodes.shape[0], _p(nodes, C.c_double), eoff.size - 1, _p(eoff, C.c_uint32),
                                  _p(enodes, C.c_uint32), bnodes.shape[0], _p(bnodes, C.c_uint32),
                                  len(raw.btags), _p(bgs, C.c_uint64))
        if not self.h:
            raise RuntimeError("oracle buildMesh failed (bad element or negative area)")
        cnt = np.zeros(6, dtype=np.int64)
        L.orc_mesh_counts(self.h, _p(cnt, C.c_int64))
        self.n_cells, self.n_interior, self.n_ghost, self.n_edges, self.n_interior_edges, self.n_boundary_edges = map(int, cnt)
        self.cfg = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_free(self.h)
            self.h = None

    def cells(self):
        Cn = self.n_cells
        out = dict(area=np.zeros(Cn), d=np.zeros(Cn), perim=np.zeros(Cn), centroid=np.zeros((Cn, 2)),
                   gradMat=np.zeros((Cn, 2, 6)), edges=np.zeros((Cn, 6), np.uint32), fm=np.zeros((Cn, 6)),
                   nbrs=np.zeros((Cn, 6), np.uint32), nEdges=np.zeros(Cn, np.uint32))
        lib().orc_mesh_cells(self.h, *[_p(out[k], C.c_double if out[k].dtype == np.float64 else C.c_uint32)
                                       for k in ("area", "d", "perim", "centroid", "gradMat", "edges", "fm", "nbrs", "nEdges")])
        return out

    def edges(self):
        E = self.n_edges
        out = dict(nodeIdx=np.zeros((E, 2), np.uint32), cellIdx=np.zeros((E, 2), np.uint32), len=np.zeros(E),
                   normal=np.zeros((E, 2)), mid=np.zeros((E, 2)), bGroup=np.zeros(E, np.int32))
        lib().orc_mesh_edges(self.h, _p(out["nodeIdx"], C.c_uint32), _p(out["cellIdx"], C.c_uint32),
                             _p(out["len"], C.c_double), _p(out["normal"], C.c_double), _p(out["mid"], C.c_double),
                             _p(out["bGroup"], C.c_int32))
        return out

    def boundary_group(self, tag):
        n = np.zeros(1, np.uint32)
        lib().orc_boundary_group(self.h, tag, None, _p(n, C.c_uint32))
        out = np.zeros(int(n[0]), np.uint32)
        lib().orc_boundary_group(self.h, tag, _p(out, C.c_uint32), _p(n, C.c_uint32))
        return out

    # ---- solver
    def setup(self, cfg_dict):
        self.cfg = make_config(cfg_dict, self.raw.btags)
        rc = lib().orc_setup(self.h, C.byref(self.cfg))
        self._check(rc)
        return self

    def _check(self, rc):
        if rc:
            kind, idx = C.c_int(0), C.c_int64(0)
            lib().orc_last_error(self.h, C.byref(kind), C.byref(idx))
            raise OracleError(rc, kind.value, idx.value)

    @property
    def q(self):
        out = np.zeros((self.n_cells, 4))
        lib().orc_get_q(self.h, _p(out, C.c_double))
        return out

    @q.setter
    def q(self, val):
        v = _f64(val).reshape(self.n_cells, 4)
        lib().orc_set_q(self.h, _p(v, C.c_double))

    def residual(self, limit=True, dt_update=True, rmax=0.0):
        R = np.zeros((self.n_cells, 4))
        newDt, Rmax = np.zeros(1), np.array([rmax])
        rc = lib().orc_residual(self.h, C.byref(self.cfg), _p(R, C.c_double), _p(newDt, C.c_double),
                                _p(Rmax, C.c_double), int(limit), int(dt_update))
        self._check(rc)
        return R[: self.n_interior], newDt[0], Rmax[0]

    def step(self, dt):
        dtv, Rmax = np.array([dt]), np.zeros(1)
        rc = lib().orc_step(self.h, C.byref(self.cfg), _p(dtv, C.c_double), _p(Rmax, C.c_double))
        self._check(rc)
        return dtv[0], Rmax[0]

    def run(self, max_steps, t=0.0, dt=None):
        state = np.array([t, self.cfg.dt if dt is None else dt])
        hist = np.zeros((max(int(max_steps), 1), len(HIST_FIELDS)))
        done = C.c_int64(0)
        rc = lib().orc_run(self.h, C.byref(self.cfg), int(max_steps), _p(state, C.c_double), _p(hist, C.c_double),
                           C.byref(done))
        self._check(rc)
        return hist[: done.value], state[0], state[1]

    def edge_data(self):
        E = self.n_edges
        eq, fl, sm = np.zeros((E, 2, 4)), np.zeros((E, 4)), np.zeros(E)
        lib().orc_get_edge_data(self.h, _p(eq, C.c_double), _p(fl, C.c_double), _p(sm, C.c_double))
        return eq, fl, sm

    def cell_data(self):
        Cn = self.n_cells
        g, l, dt = np.zeros((Cn, 4, 2)), np.zeros((Cn, 4, 2)), np.zeros(Cn)
        lib().orc_get_cell_data(self.h, _p(g, C.c_double), _p(l, C.c_double), _p(dt, C.c_double))
        return g, l, dt
