"""MI355X-native explicit residual-and-update solver for EbbCFD's 2-D finite-volume path.

The compute path is libebbcfd.so (HIP kernels for gfx950 behind the C ABI in
include/ebbcfd.h); this package is its Python host mirror plus mesh generators.
"""
from .meshgen import RawMesh, channel_bump, cylinder_ogrid, mms_square, write_gmsh  # noqa: F401
from .solver import (HIST_FIELDS, Batch, Config, Server, Solver, UMesh2, build_mesh,  # noqa: F401
                     estimate_device_bytes, import_mesh, limiter1d, load_config, pinned_array)
