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


def test_shedding_frequency_of_known_signals():
    """The Strouhal estimate of bench.py --strouhal on lift histories of known frequency:
    a sinusoid on an uneven time grid with an offset and a decaying start-up transient,
    and a non-sinusoidal periodic signal (one upward zero crossing per period)."""
    import numpy as np
    import bench
    rng = np.random.Generator(np.random.PCG64(3))
    dt = rng.uniform(0.8e-3, 1.2e-3, 200_000)
    t = np.cumsum(dt)
    H = np.zeros((t.size, 10))
    H[:, 0] = t
    f = 0.1319
    H[:, 8] = 0.3 + 0.013 * np.sin(2 * np.pi * f * t + 0.4) + 0.5 * np.exp(-t / 3.0)
    St, per, amp = bench.shedding_frequency(H, t_from=40.0)
    assert abs(St - f) <= 1e-6 * f and len(per) >= 15
    assert abs(amp - 0.026) <= 1e-4
    H[:, 8] = np.sign(np.sin(2 * np.pi * 0.25 * t)) + 0.3 * np.sin(2 * np.pi * 0.75 * t)
    St, _, _ = bench.shedding_frequency(H, t_from=10.0)
    assert abs(St - 0.25) <= 1e-4
    H[:, 8] = 1.0   # no oscillation: no frequency
    St, per, _ = bench.shedding_frequency(H, t_from=10.0)
    assert St is None and per is None
