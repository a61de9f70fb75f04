"""The 1-D slope limiters of source/limiters.d:7-37 on the device, bit for bit against the
oracle's restatement (pinned by textbook values in test_oracle_flux.py), over random
pairs of both signs, zeros, and magnitudes from 1e-300 to 1e300."""
import numpy as np
import pytest

import oracle
from ebbcfd_amd import limiter1d
from ebbcfd_amd.solver import LIMITERS_1D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", range(6))
def test_limiter1d_equals_oracle(kind):
    rng = np.random.default_rng(kind)
    n = 20000
    sp = rng.uniform(-2, 2, n) * 10.0 ** rng.integers(-300, 300, n)
    sm = rng.uniform(-2, 2, n) * 10.0 ** rng.integers(-300, 300, n)
    sp[:50] = 0.0
    sm[25:75] = -0.0
    sm[100:200] = -sp[100:200]
    sm[200:300] = sp[200:300]
    out = limiter1d(LIMITERS_1D[kind], sp, sm)
    ref = np.array([oracle.limiter1d(kind, a, b) for a, b in zip(sp, sm)])
    np.testing.assert_array_equal(out, ref)      # NaN where the reference gives NaN (inf * sgn(0))
