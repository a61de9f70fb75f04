"""bench.py's command-line contract on the host: it measures one GPU only, and the mesh
ingest mode (host code, no GPU) prints one JSON line with its CPU baseline."""
import json
import os
import subprocess
import sys

from helpers import ROOT


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=300, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    p = _bench("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_ingest_line():
    p = _bench("--ingest", "--n-theta", "64", "--n-r", "32", "--ingest-repeats", "2")
    assert p.returncode == 0, p.stderr
    lines = [l for l in p.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    assert d["unit"] == "Mcell/s" and d["value"] > 0 and len(d["runs_s"]) == 2
    assert d["config"]["cells"] == 64 * 32 * 2 and d["config"]["workload"] == "cylinder_ns_ogrid_64x32x2"
    cpu = d["cpu_baseline"]
    assert cpu["kind"] == "port" and cpu["cores"] == 1 and cpu["value"] > 0
