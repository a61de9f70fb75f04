"""The manufactured-solution Euler workload at its full stated size, bit for bit
against the oracle.

64 x 64 quads on [0,1]^2 split into 8,192 triangles, interior vertices jittered
by up to 10 % of h (seed 1), exact Dirichlet boundaries and source
(app/ms.d:131-141, 285-298), Roe flux (source/flux.d:224-303), least-squares
gradients with the scalar limiter (source/solve.d:324-403), RK4
(source/integrators.d:54-134), CFL 0.5, 1,000 steps.  The whole per-step
history and the final states must equal the oracle's exactly; the path's stated
FP64 tolerance (1e-11 relative) is asserted as well.
"""
import numpy as np
import pytest

import oracle
from ebbcfd_amd import Solver, build_mesh

from helpers import mms_case, oracle_cfg

pytestmark = pytest.mark.gpu


def test_mms_8192_tris_rk4_1000_steps_bitwise():
    raw, cfg = mms_case(n=64, jitter=0.1, seed=1, integrator="RK4")
    cfg["plotIter"] = -1
    s = Solver(build_mesh(raw), cfg)
    assert s.n_interior == 8192
    hist, t, dt = s.run(1000)
    om = oracle.Mesh(raw)
    om.setup(oracle_cfg(cfg))
    ohist, ot, odt = om.run(1000)
    assert len(hist) == 1000 and len(ohist) == 1000
    q, oq = s.q, om.q[: om.n_interior]
    rel = float(np.abs(q - oq).max() / np.abs(oq).max())
    print("C1 1000 RK4 steps: t %.6g dt %.6g, max rel state difference %g" % (t, dt, rel))
    assert rel <= 1e-11
    np.testing.assert_array_equal(q, oq)
    np.testing.assert_array_equal(hist, ohist)
    assert (t, dt) == (ot, odt)
