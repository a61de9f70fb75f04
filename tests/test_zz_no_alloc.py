"""ebb_residual_noalloc (the allocation-free variant of ebb_residual) against ebb_residual:
same R, new dt and Rmax bit for bit, on two calls in a row, the second on a smaller mesh
(the shared fixed scratch still holds the first call's partials)."""
import ctypes as C

import numpy as np
import pytest

from ebbcfd_amd import Solver, build_mesh
from ebbcfd_amd._lib import lib

from helpers import channel_case, cylinder_case, perturb

pytestmark = pytest.mark.gpu


def _noalloc(s, R_dev, rmax0=0.0, limit=True, dt_update=True):
    nd, rm = C.c_double(0), C.c_double(rmax0)
    rc = lib().ebb_residual_noalloc(s.h, int(limit), int(dt_update), C.c_void_p(R_dev.data_ptr()), C.byref(nd),
                                    C.byref(rm))
    assert rc == 0, rc
    return nd.value, rm.value


def test_noalloc_residual_equals_ebb_residual():
    import torch
    cases = [cylinder_case(n_theta=256, n_r=128, CFL=0.05), channel_case(nx=24, ny=8)]
    for k, (raw, cfg) in enumerate(cases):           # the second call is on the smaller mesh
        s = Solver(build_mesh(raw), cfg)
        s.q = perturb(s.q, 0.02, seed=k)
        R_dev = torch.empty((s.n_interior, 4), dtype=torch.float64, device="cuda")
        nd, rm = _noalloc(s, R_dev, rmax0=0.0)
        torch.cuda.synchronize()
        Rn = R_dev.cpu().numpy()
        R, nd2, rm2 = s.ufvm_solver(limit=True, dt_update=True, rmax=0.0)
        print("case %d: %d cells, newDt %.17g, Rmax %.17g" % (k, s.n_interior, nd, rm))
        np.testing.assert_array_equal(Rn, R)
        assert (nd, rm) == (nd2, rm2)
        # Rmax accumulates from the caller's value, and an unlimited call matches too
        nd3, rm3 = _noalloc(s, R_dev, rmax0=10.0 * rm, limit=False, dt_update=False)
        torch.cuda.synchronize()
        R4, nd4, rm4 = s.ufvm_solver(limit=False, dt_update=False, rmax=10.0 * rm)
        np.testing.assert_array_equal(R_dev.cpu().numpy(), R4)
        assert (nd3, rm3) == (nd4, rm4) and rm3 == 10.0 * rm
