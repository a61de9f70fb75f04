"""File interfaces on the GPU: .esln v2 save/load/restart (source/io.d:55-366),
.frc force records (app/solver.d:178-193), the ebb-solve CLI (app/solver.d:430-470),
and the ConstPressure boundary (solve.d:570-596, 866-880)."""
import json
import os
import signal
import struct
import subprocess
import time
import zlib

import numpy as np
import pytest

import oracle
from ebbcfd_amd import Solver, build_mesh, import_mesh, meshgen, pinned_array

from helpers import ROOT, base_config, channel_case, oracle_cfg, perturb

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "ebbcfd_amd", "ebb-solve")


def read_esln(path):
    """Byte layout of source/io.d:68-161 (big-endian)."""
    b = open(path, "rb").read()
    magic, ver, n = struct.unpack(">III", b[:12])
    t, dt = struct.unpack(">dd", b[12:28])
    M, order = struct.unpack(">II", b[28:36])
    off = 36
    cells = []
    for _ in range(n):
        gid, ne = struct.unpack(">IB", b[off:off + 5])
        off += 5
        q = np.array(struct.unpack(">4d", b[off:off + 32]))
        off += 32
        e = np.array(struct.unpack(">%dd" % (4 * ne), b[off:off + 32 * ne])).reshape(ne, 4)
        off += 32 * ne
        cells.append((gid, ne, q, e))
    crc = b[off:off + 4]
    assert off + 4 == len(b)
    return dict(magic=magic, ver=ver, n=n, t=t, dt=dt, M=M, order=order, cells=cells, crc=crc, body=b[:off])


def test_esln_layout_and_contents(tmp_path):
    raw, cfg = channel_case(integrator="RK4")
    s = Solver(build_mesh(raw), cfg)
    s.run(7)
    path = str(tmp_path / "a.esln")
    s.save_solution(path)
    f = read_esln(path)
    st = s.run_state()
    assert f["magic"] == 0xEA98E1F5 and f["ver"] == 2 and f["M"] == 4 and f["order"] == 2
    assert f["n"] == s.n_interior and f["t"] == st["t"] and f["dt"] == st["dt"]
    q = s.q
    eq, _, _ = s.edge_states()
    g = s.mesh.cell_geometry()
    e = s.mesh.edge_geometry()
    for i, (gid, ne, qc, ev) in enumerate(f["cells"]):
        assert gid == i and ne == g["nEdges"][i]
        np.testing.assert_array_equal(qc, q[i])
        for j in range(ne):
            k = g["edges"][i, j]
            side = 0 if e["cellIdx"][k, 0] == i else 1
            np.testing.assert_array_equal(ev[j], eq[k, side])
    # CRC-32 of the buffer with its 4 trailing bytes zero, little-endian byte order
    assert f["crc"] == struct.pack("<I", zlib.crc32(f["body"] + b"\0\0\0\0") & 0xFFFFFFFF)


def test_esln_edge_states_match_oracle(tmp_path):
    raw, cfg = channel_case(integrator="RK2_TVD")
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    s = Solver(build_mesh(raw), cfg)
    om.run(5)
    s.run(5)
    eq_o, _, _ = om.edge_data()
    eq_g, _, _ = s.edge_states()
    np.testing.assert_array_equal(eq_g, eq_o)


def test_restart_continues_bitwise(tmp_path):
    raw, cfg = channel_case(integrator="RK2_TVD", cflAdjust=False)
    a = Solver(build_mesh(raw), cfg)
    h1, _, _ = a.run(10)
    path = str(tmp_path / "r.esln")
    a.save_solution(path)
    h2, _, _ = a.run(15)
    b = Solver(build_mesh(raw), cfg)
    b.load_solution(path)
    st = b.run_state()
    assert st["dt"] == h1[-1, 1]
    h3, _, _ = b.run(15)
    np.testing.assert_array_equal(b.q, a.q)
    np.testing.assert_array_equal(h3[:, [0, 1, 6, 7, 8]], h2[:, [0, 1, 6, 7, 8]])
    # the restarted loop's 'last' state is the loaded one, so even the first RMS row agrees
    np.testing.assert_array_equal(h3, h2)


def test_forces_file(tmp_path):
    raw, cfg = channel_case(integrator="RK4")
    s = Solver(build_mesh(raw), cfg)
    h, _, _ = s.run(12)
    path = str(tmp_path / "f.frc")
    s.write_forces(path)
    data = np.frombuffer(open(path, "rb").read(), dtype=">f8").reshape(-1, 3)
    np.testing.assert_array_equal(data, h[:, [0, 7, 8]])
    assert np.any(data[:, 1] != 0)


def _write_case(tmp_path, steps_cfg):
    raw, cfg = channel_case(integrator="RK2_TVD", **steps_cfg)
    meshgen.write_gmsh(raw, str(tmp_path / "bump.msh"))
    cfg["mesh"] = "bump.msh"
    with open(tmp_path / "cfg.json", "w") as f:
        json.dump(cfg, f)
    return raw, cfg


def test_cli_matches_library(tmp_path):
    raw, cfg = _write_case(tmp_path, dict(dynamicDt=False, dt=2.0e-4, tEnd=0.00205, saveIter=4, plotIter=3))
    r = subprocess.run([CLI, "-c", "cfg.json"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # the display mesh of app/solver.d:364-372, named after the mesh file
    assert os.path.getsize(tmp_path / "bump_0.mmsh") > 0
    s = Solver(import_mesh(str(tmp_path / "bump.msh")), cfg)
    h, t, dt = s.run(-1)
    assert len(h) == 11
    frc = np.frombuffer(open(tmp_path / "boundaryForces.frc", "rb").read(), dtype=">f8").reshape(-1, 3)
    np.testing.assert_array_equal(frc, h[:, [0, 7, 8]])
    fin = read_esln(str(tmp_path / "final_0.esln"))
    np.testing.assert_array_equal(np.array([c[2] for c in fin["cells"]]), s.q)
    assert fin["t"] == t
    # saves after iterations 0, 4, 8 carry the t of the step that produced them (app/solver.d:232-247)
    for n, it in enumerate((0, 4, 8)):
        sv = read_esln(str(tmp_path / ("save_%d_0.esln" % n)))
        assert sv["t"] == h[it, 0] and sv["dt"] == h[it, 1]
    assert r.stdout.count("rho_RMS") == len(range(0, 11, 3)) + 1
    # restart from save_2 (written after iteration 8, with that step's start time): the
    # reference's loop then runs until t >= tEnd again, i.e. 3 more steps, 12 in all
    r2 = subprocess.run([CLI, "-c", "cfg.json", "-s", "save_2_"], cwd=tmp_path, capture_output=True, text=True,
                        timeout=120)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    fin2 = read_esln(str(tmp_path / "final_0.esln"))
    long_cfg = dict(cfg, tEnd=1.0e9)
    s12 = Solver(import_mesh(str(tmp_path / "bump.msh")), long_cfg)
    s12.run(12)
    np.testing.assert_array_equal(np.array([c[2] for c in fin2["cells"]]), s12.q)


def test_cli_sigusr1_writes_final(tmp_path):
    _write_case(tmp_path, dict(tEnd=1.0e9, saveIter=-1, plotIter=-1))
    p = subprocess.Popen([CLI, "-c", "cfg.json"], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    time.sleep(3.0)
    p.send_signal(signal.SIGUSR1)
    out, err = p.communicate(timeout=60)
    assert p.returncode == 0, out + err
    assert "Signal received" in out
    assert os.path.exists(tmp_path / "final_0.esln")
    assert read_esln(str(tmp_path / "final_0.esln"))["n"] > 0


def test_constant_pressure_outflow_matches_oracle():
    """ConstP ghosts use pow(); the device's pow differs from glibc's by at most an ulp or so."""
    raw = meshgen.channel_bump(24, 8, 0.05, seed=3)
    q = [0.5, 0.0, 1.0, 1.0]
    p_out = 1.0 / (1.4 * 0.5 * 0.5) * 1.02
    bcs = [{"tag": "Inflow", "type": "fullState", "q": q}, {"tag": "Outflow", "type": "constP", "q": [p_out]},
           {"tag": "Lower", "type": "inviscidWall", "q": q}, {"tag": "Upper", "type": "inviscidWall", "q": q}]
    cfg = base_config(initialConditions=q, boudaryConditions=bcs, forceBoundary="Lower", integrator="RK2")
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    s = Solver(build_mesh(raw), cfg)
    qp = perturb(om.q[: om.n_interior], 0.02, seed=2)
    qq = om.q
    qq[: om.n_interior] = qp
    om.q = qq
    s.q = qp
    R_o, nd_o, _ = om.residual()
    R_g, nd_g, _ = s.ufvm_solver()
    np.testing.assert_allclose(R_g, R_o, rtol=1e-12, atol=1e-12 * np.abs(R_o).max())
    ho, _, _ = om.run(20)
    hg, _, _ = s.run(20)
    qo = om.q[: om.n_interior]
    assert np.all(np.abs(s.q - qo) <= 1e-11 * np.abs(qo).max(axis=0))


def _total_conditions_inflow_case(aoa, tt_factor):
    """Channel with a total-temperature/total-pressure inflow (tpInflow, solve.d:625-683) built
    from the M=0.5 freestream; tt_factor scales the total temperature."""
    raw = meshgen.channel_bump(24, 8, 0.05, seed=3)
    M = 0.5
    q = [M, 0.0, 1.0, 1.0]
    p = 1.0 / (1.4 * M * M)
    T = p / 8.3145
    Tt = T * (1 + 0.2 * M * M) * tt_factor
    Pt = p * (1 + 0.2 * M * M) ** 3.5
    bcs = [{"tag": "Inflow", "type": "tpInflow", "q": [Tt, Pt, aoa]}, {"tag": "Outflow", "type": "fullState", "q": q},
           {"tag": "Lower", "type": "inviscidWall", "q": q}, {"tag": "Upper", "type": "inviscidWall", "q": q}]
    return raw, base_config(initialConditions=q, boudaryConditions=bcs, forceBoundary="Lower", integrator="RK2")


def test_total_conditions_inflow_matches_oracle():
    """The ghost keeps its old momentum and energy and gets only component 0 (the reference
    writes q[g][0] four times); pow() in the ghost state allows an ulp or so."""
    raw, cfg = _total_conditions_inflow_case(180.0, 1.2)
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    s = Solver(build_mesh(raw), cfg)
    R_o, nd_o, _ = om.residual()
    R_g, nd_g, _ = s.ufvm_solver()
    np.testing.assert_allclose(R_g, R_o, rtol=1e-12, atol=1e-12 * np.abs(R_o).max())
    np.testing.assert_allclose(s.q_all(), om.q, rtol=1e-13, atol=0)
    ho, _, _ = om.run(20)
    hg, _, _ = s.run(20)
    qo = om.q[: om.n_interior]
    assert np.all(np.abs(s.q - qo) <= 1e-11 * np.abs(qo).max(axis=0))


def test_total_conditions_inflow_without_root_stops_like_reference():
    """No non-negative Mach root: the reference stops (enforce, solve.d:661-670); the device
    reports EBB_ERR_BC with the same reference edge, from the residual and from the loop."""
    raw, cfg = _total_conditions_inflow_case(0.0, 1.0)
    om = oracle.Mesh(raw).setup(oracle_cfg(cfg))
    with pytest.raises(oracle.OracleError) as eo:
        om.residual()
    s = Solver(build_mesh(raw), cfg)
    with pytest.raises(Exception) as eg:
        s.ufvm_solver()
    assert eg.value.code == 5 and eg.value.index == eo.value.index
    s2 = Solver(build_mesh(raw), cfg)
    with pytest.raises(Exception) as er:
        s2.run(5)
    assert er.value.code == 5 and er.value.index == eo.value.index


def test_pinned_host_buffers_give_the_same_results():
    """The drop-in at ufvmSolver with page-locked host arrays (ebb_host_alloc): the state
    uploaded from one and R downloaded into one equal the pageable path bit for bit, on
    repeated calls into the same buffers."""
    raw, cfg = channel_case(nx=40, ny=12, viscosity=True)
    s = Solver(build_mesh(raw), cfg)
    q0 = perturb(s.q, 0.02, seed=3)
    qp, Rp = pinned_array((s.n_interior, 4)), pinned_array((s.n_interior, 4))
    assert not Rp.any()
    for k in range(3):
        qk = perturb(q0, 0.01, seed=10 + k)
        s.q = qk
        R, nd, rm = s.ufvm_solver()
        qp[...] = qk
        s.q = qp
        R2, nd2, rm2 = s.ufvm_solver(out=Rp)
        assert R2 is Rp
        np.testing.assert_array_equal(Rp, R)
        assert (nd2, rm2) == (nd, rm)
        np.testing.assert_array_equal(s.q, qk)
    del qp, Rp   # frees the page-locked memory
