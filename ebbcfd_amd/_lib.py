"""ctypes binding of libebbcfd.so (the C ABI declared in include/ebbcfd.h).

The shared library is built in-tree (``ebbcfd_amd/libebbcfd.so``) by
``ebbcfd_amd.build.build()``.  There is no fallback: if the library is missing
or cannot be loaded, every call raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libebbcfd.so")

# every symbol include/ebbcfd.h declares
SYMBOLS = (
    "ebb_error_string", "ebb_mesh_load", "ebb_mesh_create", "ebb_mesh_info", "ebb_mesh_cell_geometry",
    "ebb_mesh_edge_geometry", "ebb_mesh_tile_stats", "ebb_mesh_save_mmsh", "ebb_mesh_destroy", "ebb_config_parse", "ebb_config_get", "ebb_config_set",
    "ebb_config_mesh_file", "ebb_config_destroy", "ebb_solver_create", "ebb_state_upload", "ebb_state_download",
    "ebb_state_download_all", "ebb_residual", "ebb_residual_noalloc", "ebb_step", "ebb_run", "ebb_run_state", "ebb_run_reset",
    "ebb_history", "ebb_history_length", "ebb_edge_states", "ebb_cell_gradients", "ebb_save_solution",
    "ebb_save_solution_as",
    "ebb_load_solution", "ebb_write_forces", "ebb_last_error", "ebb_time_steps", "ebb_solver_device_bytes",
    "ebb_solver_destroy", "ebb_host_alloc", "ebb_host_free", "ebb_limiter1d", "ebb_estimate_device_bytes", "ebb_batch_create", "ebb_batch_add", "ebb_batch_remove",
    "ebb_batch_run", "ebb_batch_status", "ebb_batch_state_download", "ebb_batch_history_length", "ebb_batch_history",
    "ebb_batch_info", "ebb_batch_tune", "ebb_batch_device_bytes", "ebb_batch_destroy", "ebb_server_create",
    "ebb_server_submit", "ebb_server_poll", "ebb_server_result", "ebb_server_wait", "ebb_server_stats",
    "ebb_server_destroy",
)

_lib = None


class EbbError(RuntimeError):
    def __init__(self, code, what, detail=""):
        self.code = code
        msg = "%s failed: %s" % (what, lib().ebb_error_string(code).decode())
        if detail:
            msg += " (%s)" % detail
        super().__init__(msg)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libebbcfd.so is not built (%s); run __graft_entry__.build() or "
                           "`make -C ebbcfd_amd/csrc`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P, d, u32, i64, vp = C.POINTER, C.c_double, C.c_uint32, C.c_int64, C.c_void_p
    L.ebb_error_string.restype = C.c_char_p
    L.ebb_error_string.argtypes = [C.c_int]
    L.ebb_mesh_load.argtypes = [C.c_char_p, P(vp)]
    L.ebb_mesh_create.argtypes = [u32, P(d), u32, P(u32), P(u32), u32, P(u32), u32, P(C.c_char_p), P(C.c_uint64), P(vp)]
    L.ebb_mesh_info.argtypes = [vp, P(i64)]
    L.ebb_mesh_cell_geometry.argtypes = [vp, P(d), P(d), P(d), P(d), P(d), P(u32), P(C.c_int8), P(u32), P(C.c_uint8)]
    L.ebb_mesh_edge_geometry.argtypes = [vp, P(u32), P(u32), P(d), P(d), P(d), P(C.c_int32)]
    L.ebb_mesh_tile_stats.argtypes = [vp, C.c_int, C.c_int, P(d)]
    L.ebb_mesh_save_mmsh.argtypes = [vp, C.c_char_p]
    L.ebb_mesh_destroy.argtypes = [vp]
    L.ebb_mesh_destroy.restype = None
    L.ebb_config_parse.argtypes = [C.c_char_p, P(vp)]
    L.ebb_config_get.argtypes = [vp, C.c_char_p, P(d)]
    L.ebb_config_set.argtypes = [vp, C.c_char_p, d]
    L.ebb_config_mesh_file.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.ebb_config_destroy.argtypes = [vp]
    L.ebb_config_destroy.restype = None
    L.ebb_solver_create.argtypes = [vp, vp, C.c_int, P(vp)]
    L.ebb_state_upload.argtypes = [vp, P(d)]
    L.ebb_state_download.argtypes = [vp, P(d)]
    L.ebb_state_download_all.argtypes = [vp, P(d)]
    L.ebb_residual.argtypes = [vp, C.c_int, C.c_int, P(d), P(d), P(d)]
    L.ebb_residual_noalloc.argtypes = [vp, C.c_int, C.c_int, vp, P(d), P(d)]
    L.ebb_step.argtypes = [vp, P(d), P(d)]
    L.ebb_run.argtypes = [vp, i64, P(d), P(d), P(i64)]
    L.ebb_run_state.argtypes = [vp, P(d), P(d), P(d), P(i64)]
    L.ebb_run_reset.argtypes = [vp, d, d]
    L.ebb_history.argtypes = [vp, i64, i64, P(d)]
    L.ebb_history_length.argtypes = [vp]
    L.ebb_history_length.restype = i64
    L.ebb_edge_states.argtypes = [vp, P(d), P(d), P(d)]
    L.ebb_cell_gradients.argtypes = [vp, P(d), P(d)]
    L.ebb_save_solution.argtypes = [vp, C.c_char_p]
    L.ebb_save_solution_as.argtypes = [vp, C.c_char_p, d, d]
    L.ebb_load_solution.argtypes = [vp, C.c_char_p]
    L.ebb_write_forces.argtypes = [vp, C.c_char_p, i64, i64, C.c_int]
    L.ebb_last_error.argtypes = [vp, P(C.c_int), P(i64), P(i64), C.c_char_p, C.c_size_t]
    L.ebb_time_steps.argtypes = [vp, C.c_int, P(d)]
    L.ebb_solver_device_bytes.argtypes = [vp]
    L.ebb_solver_device_bytes.restype = i64
    L.ebb_solver_destroy.argtypes = [vp]
    L.ebb_solver_destroy.restype = None
    L.ebb_limiter1d.argtypes = [C.c_int, C.c_int, i64, P(d), P(d), P(d)]
    L.ebb_host_alloc.argtypes = [C.c_size_t, P(vp)]
    L.ebb_host_free.argtypes = [vp]
    L.ebb_host_free.restype = None
    L.ebb_estimate_device_bytes.argtypes = [vp, vp, P(i64)]
    L.ebb_batch_create.argtypes = [C.c_int, i64, C.c_int, C.c_int, P(vp)]
    L.ebb_batch_add.argtypes = [vp, vp, vp, P(i64)]
    L.ebb_batch_remove.argtypes = [vp, i64]
    L.ebb_batch_run.argtypes = [vp, i64, P(i64)]
    L.ebb_batch_status.argtypes = [vp, i64, P(d)]
    L.ebb_batch_state_download.argtypes = [vp, i64, P(d)]
    L.ebb_batch_history_length.argtypes = [vp, i64]
    L.ebb_batch_history_length.restype = i64
    L.ebb_batch_history.argtypes = [vp, i64, i64, i64, P(d)]
    L.ebb_batch_info.argtypes = [vp, P(i64)]
    L.ebb_batch_tune.argtypes = [vp, C.c_int, P(d)]
    L.ebb_batch_device_bytes.argtypes = [vp]
    L.ebb_batch_device_bytes.restype = i64
    L.ebb_batch_destroy.argtypes = [vp]
    L.ebb_batch_destroy.restype = None
    L.ebb_server_create.argtypes = [C.c_int, i64, C.c_int, C.c_int, C.c_int, C.c_int, P(vp)]
    L.ebb_server_submit.argtypes = [vp, vp, vp, P(i64)]
    L.ebb_server_poll.argtypes = [vp, i64, P(d)]
    L.ebb_server_result.argtypes = [vp, i64, P(d)]
    L.ebb_server_wait.argtypes = [vp, i64, d]
    L.ebb_server_stats.argtypes = [vp, P(d)]
    L.ebb_server_destroy.argtypes = [vp]
    L.ebb_server_destroy.restype = None
    _lib = L
    return L
