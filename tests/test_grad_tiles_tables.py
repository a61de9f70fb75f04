"""The gradient-tile tables (tiles.cpp) checked on the host, without a GPU: the LDS rows
each face names, the ring lists, the border, the copied neighbour rows and the LDS size,
over triangle, quad, mixed and pentagon meshes, tile sizes from one cell to 768 and the
three tile orders (tests/grad_tiles_tables.cpp)."""
import glob
import os
import shutil
import subprocess

import pytest

from ebbcfd_amd import meshgen

from helpers import FLAT_DIR, MMS_DIR, ROOT, hanging_node_mesh, mixed_mesh

CSRC = os.path.join(ROOT, "ebbcfd_amd", "csrc")


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="needs hipcc (host compile only)")
def test_grad_tile_tables(tmp_path):
    exe = str(tmp_path / "grad_tiles_tables")
    cmd = ["hipcc", "-std=c++17", "-O1", "-g", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
           "-fsanitize=undefined", os.path.join(ROOT, "tests", "grad_tiles_tables.cpp"), os.path.join(CSRC, "tiles.cpp"),
           os.path.join(CSRC, "mesh.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, timeout=600, capture_output=True)
    meshes = [os.path.join(MMS_DIR, "ms_refine_tri_pert_2.msh"), os.path.join(MMS_DIR, "ms_refine_2.msh"),
              os.path.join(FLAT_DIR, "flatPlate_1.msh")]
    for name, raw in (("cyl.msh", meshgen.cylinder_ogrid(96, 40)), ("mixed.msh", mixed_mesh()),
                      ("pent.msh", hanging_node_mesh())):
        p = str(tmp_path / name)
        meshgen.write_gmsh(raw, p)
        meshes.append(p)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe] + meshes, capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0 and "VIOLATIONS" not in r.stdout
    assert r.stdout.count(" ok\n") == len(meshes) * 15
