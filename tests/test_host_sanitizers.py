"""The host half of the library (gmsh and .gri readers, buildMesh with its threaded
geometry pass, the config parser) compiled with g++ under AddressSanitizer and
UndefinedBehaviorSanitizer, run over the reference's fixture meshes and configs, a mesh
large enough for the threaded path, and truncated files that must be rejected cleanly."""
import glob
import os
import shutil
import subprocess

import pytest

from ebbcfd_amd import meshgen

from helpers import FLAT_DIR, MMS_DIR, ROOT

CSRC = os.path.join(ROOT, "ebbcfd_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_address_and_undefined_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-pthread", os.path.join(ROOT, "tests", "sanitize_host.cpp"),
           os.path.join(CSRC, "mesh.cpp"), os.path.join(CSRC, "config.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, timeout=600)
    big = str(tmp_path / "cylinder_big.msh")                      # 73,728 cells: threaded geometry
    meshgen.write_gmsh(meshgen.cylinder_ogrid(256, 144, stretch=1.02), big)
    gri = str(tmp_path / "square.gri")
    meshgen.write_gri(meshgen.mms_square(24, 0.1, 3), gri)
    inputs = sorted(glob.glob(os.path.join(MMS_DIR, "*.msh")) + glob.glob(os.path.join(FLAT_DIR, "*.msh")) +
                    glob.glob(os.path.join(MMS_DIR, "*.json")) + glob.glob(os.path.join(FLAT_DIR, "*.json")))
    inputs += [big, gri]
    src = open(os.path.join(MMS_DIR, "ms_refine_tri_pert_1.msh")).read()
    for k, frac in enumerate((0.05, 0.37, 0.63, 0.999)):
        p = str(tmp_path / ("trunc_%d.msh" % k))
        open(p, "w").write(src[: int(len(src) * frac)])
        inputs.append(p)
    g = open(gri).read()
    for k, frac in enumerate((0.1, 0.9)):
        p = str(tmp_path / ("trunc_%d.gri" % k))
        open(p, "w").write(g[: int(len(g) * frac)])
        inputs.append(p)
    # malformed .gri lines: a second-order triangle with 3 of its 6 nodes, a non-numeric node id
    lines = g.splitlines()
    hdr = [k for k, l in enumerate(lines) if "TriLagrange" in l][0]
    bad_q2 = lines[:hdr] + [lines[hdr].replace(" 1 TriLagrange", " 2 TriLagrange")] + lines[hdr + 1:]
    p = str(tmp_path / "trunc_q2.gri")
    open(p, "w").write("\n".join(bad_q2) + "\n")
    inputs.append(p)
    bad_tok = list(lines)
    bad_tok[hdr + 1] = bad_tok[hdr + 1].split()[0] + " x7 " + " ".join(bad_tok[hdr + 1].split()[2:])
    p = str(tmp_path / "trunc_token.gri")
    open(p, "w").write("\n".join(bad_tok) + "\n")
    inputs.append(p)
    p = str(tmp_path / "trunc_deep.json")                         # 200k nested arrays
    open(p, "w").write('{"mesh": ' + "[" * 200000 + "]" * 200000 + "}")
    inputs.append(p)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([exe] + inputs, capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    lines = r.stdout.splitlines()
    assert len(lines) == len(inputs)
    assert sum(l.startswith("ok ") for l in lines) == sum("trunc" not in os.path.basename(p) for p in inputs)
