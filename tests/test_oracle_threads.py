"""The oracle's per-cell and per-edge loops run on OpenMP threads; with any thread count
its results (states, per-step history, error reports) equal the single-thread run bit
for bit, so the multi-core CPU baseline times the same computation."""
import numpy as np
import pytest

import oracle

from helpers import channel_case, cylinder_case, mms_case, oracle_cfg


@pytest.mark.parametrize("make", [lambda: mms_case(n=24), lambda: cylinder_case(n_theta=64, n_r=32, CFL=0.05),
                                  lambda: channel_case(integrator="RK4", localTimestep=True),
                                  lambda: channel_case(lpThresh=0.5, integrator="RK2")])
def test_threads_do_not_change_results(make):
    raw, cfg = make()
    out = []
    for n in (1, 4):
        oracle.set_threads(n)
        m = oracle.Mesh(raw).setup(oracle_cfg(cfg))
        h, t, dt = m.run(30)
        out.append((h, m.q, t, dt))
    oracle.set_threads(1)
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert out[0][2:] == out[1][2:]


def test_threads_report_the_first_nan_edge():
    raw, cfg = channel_case()
    idx = []
    for n in (1, 4):
        oracle.set_threads(n)
        m = oracle.Mesh(raw).setup(oracle_cfg(cfg))
        q = m.q
        q[5, 0] = np.nan
        q[40, 0] = np.nan
        m.q = q
        with pytest.raises(oracle.OracleError) as e:
            m.residual()
        idx.append((e.value.kind, e.value.index))
    oracle.set_threads(1)
    assert idx[0] == idx[1]
