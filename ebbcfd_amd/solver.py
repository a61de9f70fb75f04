"""Python mirror of the reference's solver interface over the C ABI.

Names follow the reference:
  * ``load_config``   -- source/config.d:112 loadConfig (JSON text or dict)
  * ``import_mesh``   -- source/io.d:669 importMesh (+ buildMesh, source/mesh.d:217)
  * ``UMesh2``        -- source/mesh.d:101 (built mesh; geometry accessors)
  * ``Solver``        -- ufvmSetup (source/solve.d:175) on a HIP device, with
                         ``ufvm_solver`` (solve.d:322), ``step`` (integrators.d),
                         ``run`` (app/solver.d:38 runIntegrator), ``save_solution``
                         / ``load_solution`` (io.d:68, 218).
Everything runs through libebbcfd.so on the GPU; nothing here computes.
"""
import ctypes as C
import json

import numpy as np

from ._lib import EbbError, lib

HIST_FIELDS = ("t", "dt", "rho_rms", "u_rms", "v_rms", "E_rms", "Rmax", "drag", "lift", "CFL")


def _p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


def _check(rc, what, solver=None):
    if rc:
        detail = ""
        if solver is not None:
            kind, idx, step = C.c_int(0), C.c_int64(0), C.c_int64(0)
            buf = C.create_string_buffer(512)
            lib().ebb_last_error(solver, C.byref(kind), C.byref(idx), C.byref(step), buf, 512)
            detail = "%s; index %d step %d" % (buf.value.decode(errors="replace"), idx.value, step.value)
            err = EbbError(rc, what, detail)
            err.index, err.step = idx.value, step.value
            raise err
        raise EbbError(rc, what, detail)


class Config:
    """loadConfig result (source/config.d:72-94)."""

    def __init__(self, text):
        if isinstance(text, dict):
            text = json.dumps(text)
        self.text = text
        self.h = C.c_void_p()
        _check(lib().ebb_config_parse(text.encode(), C.byref(self.h)), "loadConfig")

    def __getitem__(self, key):
        v = C.c_double(0)
        _check(lib().ebb_config_get(self.h, key.encode(), C.byref(v)), "config get %s" % key)
        return v.value

    def __setitem__(self, key, value):
        _check(lib().ebb_config_set(self.h, key.encode(), float(value)), "config set %s" % key)

    @property
    def mesh_file(self):
        buf = C.create_string_buffer(4096)
        lib().ebb_config_mesh_file(self.h, buf, 4096)
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "h", None):
            lib().ebb_config_destroy(self.h)
            self.h = None


def load_config(text):
    return Config(text)


class UMesh2:
    """Built mesh (source/mesh.d:101-672)."""

    def __init__(self, handle):
        self.h = handle
        info = np.zeros(8, np.int64)
        _check(lib().ebb_mesh_info(self.h, _p(info, C.c_int64)), "mesh info")
        (self.n_cells, self.n_interior, self.n_ghost, self.n_edges, self.n_interior_edges,
         self.n_boundary_edges, self.max_edges, self.n_tags) = map(int, info)

    @classmethod
    def from_raw(cls, raw):
        nodes = np.ascontiguousarray(raw.nodes, np.float64).reshape(-1, 2)
        eoff = np.ascontiguousarray(raw.elem_offsets, np.uint32)
        enodes = np.ascontiguousarray(raw.elem_nodes, np.uint32)
        bn = np.ascontiguousarray(raw.bnodes, np.uint32).reshape(-1, 2)
        bgs = np.ascontiguousarray(raw.bgroup_start, np.uint64)
        tags = (C.c_char_p * max(1, len(raw.btags)))(*[t.encode() for t in raw.btags])
        h = C.c_void_p()
        _check(lib().ebb_mesh_create(nodes.shape[0], _p(nodes), eoff.size - 1, _p(eoff, C.c_uint32),
                                     _p(enodes, C.c_uint32), bn.shape[0], _p(bn, C.c_uint32), len(raw.btags),
                                     tags, _p(bgs, C.c_uint64), C.byref(h)), "buildMesh")
        m = cls(h)
        m.btags = list(raw.btags)
        return m

    def cell_geometry(self):
        Cn, N = self.n_cells, self.n_interior
        out = dict(area=np.zeros(Cn), d=np.zeros(Cn), perim=np.zeros(Cn), centroid=np.zeros((Cn, 2)),
                   gradMat=np.zeros((N, 2, 6)), edges=np.zeros((N, 6), np.uint32), fm=np.zeros((N, 6), np.int8),
                   nbrs=np.zeros((N, 6), np.uint32), nEdges=np.zeros(N, np.uint8))
        _check(lib().ebb_mesh_cell_geometry(self.h, _p(out["area"]), _p(out["d"]), _p(out["perim"]),
                                            _p(out["centroid"]), _p(out["gradMat"]), _p(out["edges"], C.c_uint32),
                                            _p(out["fm"], C.c_int8), _p(out["nbrs"], C.c_uint32),
                                            _p(out["nEdges"], C.c_uint8)), "cell geometry")
        return out

    def edge_geometry(self):
        E = self.n_edges
        out = dict(nodeIdx=np.zeros((E, 2), np.uint32), cellIdx=np.zeros((E, 2), np.uint32), len=np.zeros(E),
                   normal=np.zeros((E, 2)), mid=np.zeros((E, 2)), bGroup=np.zeros(E, np.int32))
        _check(lib().ebb_mesh_edge_geometry(self.h, _p(out["nodeIdx"], C.c_uint32), _p(out["cellIdx"], C.c_uint32),
                                            _p(out["len"]), _p(out["normal"]), _p(out["mid"]),
                                            _p(out["bGroup"], C.c_int32)), "edge geometry")
        return out

    def tile_stats(self, tile_cells=128, order="grow"):
        """Tiles of the fused stage kernel for this mesh (host only)."""
        out = np.zeros(8)
        _check(lib().ebb_mesh_tile_stats(self.h, int(tile_cells), {"grow": 0, "morton": 1, "blocks": 2}[order], _p(out)),
               "tile stats")
        keys = ("tiles", "cells_per_tile", "grad_ratio", "load_ratio", "local_ratio", "faces_per_cell",
                "max_local", "lds_bytes")
        return dict(zip(keys, out.tolist()))

    def save_mmsh(self, path):
        """triangulate + buildMesh + saveMatlabMesh (source/mesh.d:678-715, source/io.d:534-643)."""
        _check(lib().ebb_mesh_save_mmsh(self.h, str(path).encode()), "saveMatlabMesh")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ebb_mesh_destroy(self.h)
            self.h = None


def import_mesh(path):
    h = C.c_void_p()
    _check(lib().ebb_mesh_load(str(path).encode(), C.byref(h)), "importMesh %s" % path)
    return UMesh2(h)


def build_mesh(raw):
    return UMesh2.from_raw(raw)


class Solver:
    """ufvmSetup + device-resident state for one mesh (source/solve.d:175-279)."""

    def __init__(self, mesh, config, device=0):
        if not isinstance(config, Config):
            config = Config(config)
        self.mesh, self.config = mesh, config
        self.h = C.c_void_p()
        _check(lib().ebb_solver_create(mesh.h, config.h, int(device), C.byref(self.h)), "ufvmSetup")
        self.n_interior, self.n_cells, self.n_edges = mesh.n_interior, mesh.n_cells, mesh.n_edges

    def __del__(self):
        if getattr(self, "h", None):
            lib().ebb_solver_destroy(self.h)
            self.h = None

    # state
    @property
    def q(self):
        out = np.zeros((self.n_interior, 4))
        _check(lib().ebb_state_download(self.h, _p(out)), "state download", self.h)
        return out

    @q.setter
    def q(self, value):
        v = np.ascontiguousarray(value, np.float64).reshape(self.n_interior, 4)
        _check(lib().ebb_state_upload(self.h, _p(v)), "state upload", self.h)

    def q_all(self):
        out = np.zeros((self.n_cells, 4))
        _check(lib().ebb_state_download_all(self.h, _p(out)), "state download", self.h)
        return out

    # operator
    def ufvm_solver(self, limit=True, dt_update=True, rmax=0.0, out=None):
        """One residual evaluation; returns (R, newDt, Rmax).  `out`: an n_interior x 4
        float64 array to receive R (e.g. a pinned_array), else a new one."""
        R = np.zeros((self.n_interior, 4)) if out is None else out
        if R.shape != (self.n_interior, 4) or R.dtype != np.float64 or not R.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float64 array of shape (%d, 4)" % self.n_interior)
        nd, rm = C.c_double(0), C.c_double(rmax)
        _check(lib().ebb_residual(self.h, int(limit), int(dt_update), _p(R), C.byref(nd), C.byref(rm)),
               "ufvmSolver", self.h)
        return R, nd.value, rm.value

    def step(self, dt):
        """Integrator.step: returns (new dt, Rmax)."""
        d, rm = C.c_double(dt), C.c_double(0)
        _check(lib().ebb_step(self.h, C.byref(d), C.byref(rm)), "step", self.h)
        return d.value, rm.value

    def run(self, max_steps=-1):
        """runIntegrator on the device; returns (history rows of this call, t, dt)."""
        start = lib().ebb_history_length(self.h)
        t, dt, n = C.c_double(0), C.c_double(0), C.c_int64(0)
        rc = lib().ebb_run(self.h, int(max_steps), C.byref(t), C.byref(dt), C.byref(n))
        _check(rc, "runIntegrator", self.h)
        return self.history(start, n.value), t.value, dt.value

    def run_state(self):
        t, dt, cfl, n = C.c_double(0), C.c_double(0), C.c_double(0), C.c_int64(0)
        _check(lib().ebb_run_state(self.h, C.byref(t), C.byref(dt), C.byref(cfl), C.byref(n)), "run state", self.h)
        return dict(t=t.value, dt=dt.value, CFL=cfl.value, steps=n.value)

    def run_reset(self, t=0.0, dt=None):
        _check(lib().ebb_run_reset(self.h, float(t), float(self.config["dt"] if dt is None else dt)), "reset", self.h)

    def history(self, first=0, count=None):
        n = lib().ebb_history_length(self.h)
        if count is None:
            count = n - first
        out = np.zeros((count, len(HIST_FIELDS)))
        if count:
            _check(lib().ebb_history(self.h, int(first), int(count), _p(out)), "history", self.h)
        return out

    def edge_states(self):
        E = self.n_edges
        eq, fl, sm = np.zeros((E, 2, 4)), np.zeros((E, 4)), np.zeros(E)
        _check(lib().ebb_edge_states(self.h, _p(eq), _p(fl), _p(sm)), "edge states", self.h)
        return eq, fl, sm

    def cell_gradients(self):
        g, l = np.zeros((self.n_interior, 4, 2)), np.zeros((self.n_interior, 4, 2))
        _check(lib().ebb_cell_gradients(self.h, _p(g), _p(l)), "cell gradients", self.h)
        return g, l

    def save_solution(self, path, t=None, dt=None):
        if t is None:
            _check(lib().ebb_save_solution(self.h, str(path).encode()), "saveSolution", self.h)
        else:
            _check(lib().ebb_save_solution_as(self.h, str(path).encode(), float(t), float(dt)), "saveSolution", self.h)

    def load_solution(self, path):
        _check(lib().ebb_load_solution(self.h, str(path).encode()), "loadSolution", self.h)

    def write_forces(self, path, first=0, count=None, append=False):
        if count is None:
            count = lib().ebb_history_length(self.h) - first
        _check(lib().ebb_write_forces(self.h, str(path).encode(), int(first), int(count), int(append)), "forces")

    def time_steps(self, steps):
        """Device time per kernel launch (ms) over `steps` integrator steps, HIP events on the solver stream."""
        out = np.zeros(6)
        _check(lib().ebb_time_steps(self.h, int(steps), _p(out)), "time steps", self.h)
        return dict(gradient_ms=out[0], faces_ms=out[1], update_ms=out[2], fused_ms=out[3], step_ms=out[4],
                    stages=int(out[5]))

    @property
    def device_bytes(self):
        return lib().ebb_solver_device_bytes(self.h)


class _HostBlock:
    """Page-locked host memory from ebb_host_alloc, freed when the last array viewing it goes."""

    def __init__(self, nbytes):
        self.p = None
        self._free = lib().ebb_host_free   # bound now: module globals may be gone when this is freed
        p = C.c_void_p()
        _check(lib().ebb_host_alloc(C.c_size_t(nbytes), C.byref(p)), "host alloc")
        self.p = p.value

    def __del__(self):
        if self.p:
            self._free(C.c_void_p(self.p))
            self.p = None


def pinned_array(shape, dtype=np.float64):
    """A zero-filled numpy array in page-locked host memory (ebb_host_alloc): state uploads,
    downloads and residuals copied through it run at the PCIe DMA rate."""
    shape = tuple(np.atleast_1d(shape).tolist())
    n = int(np.prod(shape)) * np.dtype(dtype).itemsize
    blk = _HostBlock(max(n, 1))
    buf = (C.c_char * max(n, 1)).from_address(blk.p)
    buf._block = blk   # the array's base keeps the memory alive
    arr = np.frombuffer(buf, dtype=np.uint8, count=n).view(dtype).reshape(shape)
    arr[...] = 0
    return arr


def estimate_device_bytes(mesh, config):
    """Device bytes a Solver for this mesh and config would hold (admission by footprint)."""
    if not isinstance(config, Config):
        config = Config(config)
    v = C.c_int64(0)
    _check(lib().ebb_estimate_device_bytes(mesh.h, config.h, C.byref(v)), "estimate device bytes")
    return v.value


BATCH_STATUS = ("t", "dt", "steps", "running", "converged", "rho_rms_first", "rho_rms", "error", "error_index",
                "device_bytes")


class Batch:
    """Many solves advanced by the same launches (no counterpart in the reference, whose
    integrator scratch is process-static: source/integrators.d:56-75).  Each solve's states
    and history equal its solo Solver.run's."""

    def __init__(self, device=0, device_bytes_limit=200 << 30, graph_steps=8, keep_history=False):
        self.h = C.c_void_p()
        _check(lib().ebb_batch_create(int(device), int(device_bytes_limit), int(graph_steps), int(keep_history),
                                      C.byref(self.h)), "batch create")
        self.meshes = {}

    def add(self, mesh, config):
        """Admit a solve; returns its id.  Raises EbbError (code EBB_ERR_ALLOC = 10) when its
        device footprint does not fit what the limit leaves."""
        if not isinstance(config, Config):
            config = Config(config)
        i = C.c_int64(0)
        _check(lib().ebb_batch_add(self.h, mesh.h, config.h, C.byref(i)), "batch add")
        self.meshes[i.value] = mesh
        return i.value

    def remove(self, sid):
        _check(lib().ebb_batch_remove(self.h, int(sid)), "batch remove")
        self.meshes.pop(sid, None)

    def run(self, steps):
        """Advance every running solve by `steps` steps; returns how many are still running."""
        r = C.c_int64(0)
        _check(lib().ebb_batch_run(self.h, int(steps), C.byref(r)), "batch run")
        return r.value

    def status(self, sid):
        out = np.zeros(10)
        _check(lib().ebb_batch_status(self.h, int(sid), _p(out)), "batch status")
        return dict(zip(BATCH_STATUS, out.tolist()))

    def q(self, sid):
        out = np.zeros((self.meshes[sid].n_interior, 4))
        _check(lib().ebb_batch_state_download(self.h, int(sid), _p(out)), "batch state")
        return out

    def history(self, sid):
        n = lib().ebb_batch_history_length(self.h, int(sid))
        out = np.zeros((n, len(HIST_FIELDS)))
        if n:
            _check(lib().ebb_batch_history(self.h, int(sid), 0, int(n), _p(out)), "batch history")
        return out

    def info(self):
        out = np.zeros(6, np.int64)
        _check(lib().ebb_batch_info(self.h, _p(out, C.c_int64)), "batch info")
        return dict(zip(("solves", "running", "device_bytes", "limit", "tile_cells", "graph_steps"), out.tolist()))

    def tune(self, steps=16):
        """Autotune tile size and graph length on the running solves (their states are kept)."""
        out = np.zeros(4)
        _check(lib().ebb_batch_tune(self.h, int(steps), _p(out)), "batch tune")
        return dict(tile_cells=int(out[0]), graph_steps=int(out[1]), ms_per_step=out[2], default_ms_per_step=out[3])

    def __del__(self):
        if getattr(self, "h", None):
            lib().ebb_batch_destroy(self.h)
            self.h = None


SERVER_STATS = ("submitted", "done", "failed", "running", "queued", "p50_s", "p99_s", "solves_per_s", "device_bytes",
                "batched_run_s", "rounds", "age_s")


class Server:
    """Serving loop over a Batch (library-side loop thread; no counterpart in the reference):
    requests queue as they arrive, are admitted by device footprint, and are stepped together
    with the running solves until each stops (extensions.rmsDrop, maxSteps, tEnd)."""

    def __init__(self, device=0, device_bytes_limit=200 << 30, max_solves=256, steps_per_round=100, graph_steps=8,
                 keep_results=False):
        self.h = C.c_void_p()
        _check(lib().ebb_server_create(int(device), int(device_bytes_limit), int(max_solves), int(steps_per_round),
                                       int(graph_steps), int(keep_results), C.byref(self.h)), "server create")
        self._keep = {}

    def submit(self, mesh, config):
        if not isinstance(config, Config):
            config = Config(config)
        t = C.c_int64(0)
        _check(lib().ebb_server_submit(self.h, mesh.h, config.h, C.byref(t)), "server submit")
        self._keep[t.value] = (mesh, config)     # the request's mesh and config outlive it
        return t.value

    def poll(self, ticket):
        out = np.zeros(16)
        _check(lib().ebb_server_poll(self.h, int(ticket), _p(out)), "server poll")
        d = dict(zip(("state", "code", "arrival_s", "admit_s", "done_s", "cells"), out[:6].tolist()))
        d.update(zip(BATCH_STATUS, out[6:].tolist()))
        return d

    def result(self, ticket):
        mesh = self._keep[ticket][0]
        q = np.zeros((mesh.n_interior, 4))
        _check(lib().ebb_server_result(self.h, int(ticket), _p(q)), "server result")
        return q

    def wait(self, n_done, timeout_s=600.0):
        return lib().ebb_server_wait(self.h, int(n_done), float(timeout_s)) == 0

    def stats(self):
        out = np.zeros(12)
        _check(lib().ebb_server_stats(self.h, _p(out)), "server stats")
        return dict(zip(SERVER_STATS, out.tolist()))

    def close(self):
        if getattr(self, "h", None):
            lib().ebb_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


LIMITERS_1D = ("firstOrderUpwindS", "laxWendroffS", "minmodS", "harmonicS", "geometricS", "superbeeS")


def limiter1d(kind, sp, sm, device=0):
    """source/limiters.d:7-37 on the device, elementwise over the pairs (sp, sm)."""
    k = LIMITERS_1D.index(kind) if isinstance(kind, str) else int(kind)
    a = np.ascontiguousarray(sp, np.float64).ravel()
    b = np.ascontiguousarray(sm, np.float64).ravel()
    out = np.empty_like(a)
    _check(lib().ebb_limiter1d(int(device), k, a.size, _p(a), _p(b), _p(out)), "limiter1d")
    return out
