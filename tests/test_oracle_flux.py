"""Pinning of the oracle's pointwise functions (CPU only).

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign, swapped index, transposed matrix) fails it:
  * roeFlux: the reference's 16 known answers (source/flux.d:161-215);
  * both fluxes: consistency F(q,q,n) = physical flux, conservation
    F(qL,qR,-n) = -F(qR,qL,n), rotational invariance;
  * diffusiveFlux: closed-form Navier-Stokes stresses and heat flux from
    primitive gradients, zero for a uniform state, linear in the gradient;
  * manufactured solution: gradient and Euler source term against central
    finite differences of the exact fields and fluxes;
  * 1-D limiters: textbook values; LP: optimum agrees with scipy linprog.
"""
import math

import numpy as np
import pytest

import oracle

G = 1.4

KAT = [
    ((1, 1), (2, 1), (1.0, 0.0), (0.087503926254201, 1.039732984117400, 0.008215342318582, 0.367160441647966)),
    ((1, 1), (1, 2), (0.0, 1.0), (0.117173193168260, 0.010523092695679, 1.005792391773638, 0.482469106067425)),
    ((1, 1), (2, 2), (math.sqrt(2.0) / 2.0, math.sqrt(2.0) / 2.0),
     (0.116317039552757, 0.743804395816748, 0.743854568709977, 0.490010270963920)),
    ((1, 1), (1.1, 1.1), (math.sqrt(2.0) / 2.0, math.sqrt(2.0) / 2.0),
     (0.115207589649517, 0.744728044864156, 0.744751077184082, 0.485506505793277)),
]


@pytest.mark.parametrize("a,b,n,ref", KAT)
def test_roe_known_answers(a, b, n, ref):
    F, _ = oracle.roe_flux(oracle.solution(*a), oracle.solution(*b), n)
    # 15 significant digits are printed in the reference
    np.testing.assert_allclose(F, ref, rtol=0, atol=5e-16 * max(1.0, np.abs(ref).max()) + 1e-15)


def physical_flux(q, n):
    rho, u, v = q[0], q[1] / q[0], q[2] / q[0]
    p = (G - 1) * (q[3] - 0.5 * rho * (u * u + v * v))
    f = np.array([rho * u, rho * u * u + p, rho * u * v, u * (q[3] + p)])
    h = np.array([rho * v, rho * u * v, rho * v * v + p, v * (q[3] + p)])
    return f * n[0] + h * n[1]


def random_state(rng, M=0.5):
    rho = rng.uniform(0.5, 2.0)
    p = rng.uniform(0.5, 2.0)
    c = math.sqrt(G * p / rho)
    th = rng.uniform(0, 2 * math.pi)
    u, v = M * c * math.cos(th), M * c * math.sin(th)
    return np.array([rho, rho * u, rho * v, p / (G - 1) + 0.5 * rho * (u * u + v * v)])


@pytest.mark.parametrize("flux", [oracle.roe_flux, oracle.rusanov_flux])
def test_flux_consistency_and_conservation(flux):
    rng = np.random.default_rng(5)
    for _ in range(50):
        q = random_state(rng, M=rng.uniform(0.05, 2.0))
        th = rng.uniform(0, 2 * math.pi)
        n = np.array([math.cos(th), math.sin(th)])
        F, _ = flux(q, q, n)
        np.testing.assert_allclose(F, physical_flux(q, n), rtol=1e-13, atol=1e-13)
        qR = random_state(rng, M=rng.uniform(0.05, 2.0))
        F1, s1 = flux(q, qR, n)
        F2, s2 = flux(qR, q, -n)
        np.testing.assert_allclose(F1, -F2, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("flux", [oracle.roe_flux, oracle.rusanov_flux])
def test_flux_rotation_invariance(flux):
    rng = np.random.default_rng(9)
    for _ in range(30):
        qL, qR = random_state(rng), random_state(rng)
        th = rng.uniform(0, 2 * math.pi)
        n = np.array([math.cos(th), math.sin(th)])
        c, s = math.cos(th), math.sin(th)

        def rot(q):  # express momentum in the (n, t) frame
            return np.array([q[0], c * q[1] + s * q[2], -s * q[1] + c * q[2], q[3]])

        F, sm = flux(qL, qR, n)
        Fr, smr = flux(rot(qL), rot(qR), np.array([1.0, 0.0]))
        back = np.array([Fr[0], c * Fr[1] - s * Fr[2], s * Fr[1] + c * Fr[2], Fr[3]])
        np.testing.assert_allclose(F, back, rtol=1e-12, atol=1e-12)
        assert abs(sm - smr) <= 1e-12 * sm


def test_rusanov_is_central_plus_jump():
    """C1 = C2 = 0 exactly (flux.d:133-138): F = (fL+fR)/2 - sMax/2 (qR-qL)."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        qL, qR = random_state(rng), random_state(rng)
        n = np.array([0.6, 0.8])
        F, sm = oracle.rusanov_flux(qL, qR, n)
        jump = np.array([qR[0] - qL[0], qR[0] * (qR[1] / qR[0]) - qL[0] * (qL[1] / qL[0]),
                         qR[0] * (qR[2] / qR[0]) - qL[0] * (qL[2] / qL[0]), qR[3] - qL[3]])
        ref = 0.5 * (physical_flux(qL, n) + physical_flux(qR, n)) - 0.5 * sm * jump
        np.testing.assert_allclose(F, ref, rtol=1e-12, atol=1e-13)


def test_viscous_flux_closed_form():
    """Navier-Stokes stresses and heat flux from primitive gradients, with the reference's
    one off-textbook entry (A10 row 2) spelled out."""
    rng = np.random.default_rng(11)
    for _ in range(30):
        rho, u, v, e = rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(1, 3)
        rx, ry, ux, uy, vx, vy, ex, ey = rng.uniform(-1, 1, 8)
        mu, Pr = rng.uniform(0.001, 0.1), 0.7
        E = e + 0.5 * (u * u + v * v)
        q = np.array([rho, rho * u, rho * v, rho * E])
        Ex, Ey = ex + u * ux + v * vx, ey + u * uy + v * vy
        dq = np.array([[rx, ry], [u * rx + rho * ux, u * ry + rho * uy], [v * rx + rho * vx, v * ry + rho * vy],
                       [E * rx + rho * Ex, E * ry + rho * Ey]])
        txx = mu * (4.0 / 3.0 * ux - 2.0 / 3.0 * vy)
        tyy = mu * (4.0 / 3.0 * vy - 2.0 / 3.0 * ux)
        txy = mu * (uy + vx)
        k = mu * G / Pr   # heat flux on the internal energy per unit mass
        # flux.d:48-51: A10 row 2 applies -2/3 to d(rho v)/dx where tau_yy needs d(rho u)/dx;
        # the y-momentum row keeps that term as written, the energy row is textbook
        tyy_ref = mu * 4.0 / 3.0 * vy + (mu / rho) * (2.0 / 3.0 * u * rx - 2.0 / 3.0 * (v * rx + rho * vx))
        Gx = np.array([0.0, txx, txy, u * txx + v * txy + k * ex])
        Gy = np.array([0.0, txy, tyy_ref, u * txy + v * tyy + k * ey])
        for n in (np.array([1.0, 0.0]), np.array([0.0, 1.0]), np.array([0.6, -0.8])):
            Fv = oracle.diffusive_flux(Pr, mu, q, dq, n)
            np.testing.assert_allclose(Fv, Gx * n[0] + Gy * n[1], rtol=1e-11, atol=1e-13)


def test_viscous_flux_zero_for_uniform_state_and_linear():
    q = np.array([1.2, 0.3, -0.1, 2.5])
    n = np.array([0.8, 0.6])
    np.testing.assert_array_equal(oracle.diffusive_flux(0.7, 0.05, q, np.zeros((4, 2)), n), np.zeros(4))
    rng = np.random.default_rng(2)
    a, b = rng.uniform(-1, 1, (4, 2)), rng.uniform(-1, 1, (4, 2))
    lhs = oracle.diffusive_flux(0.7, 0.05, q, a + 2.0 * b, n)
    rhs = oracle.diffusive_flux(0.7, 0.05, q, a, n) + 2.0 * oracle.diffusive_flux(0.7, 0.05, q, b, n)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12, atol=1e-14)


def test_mms_gradient_matches_finite_differences():
    h = 1e-6
    for (x, y) in [(0.1, 0.2), (0.5, 0.5), (0.9, 0.3), (1.3, -0.4)]:
        g = oracle.solution_gradient(x, y)
        fx = (oracle.solution(x + h, y) - oracle.solution(x - h, y)) / (2 * h)
        fy = (oracle.solution(x, y + h) - oracle.solution(x, y - h)) / (2 * h)
        np.testing.assert_allclose(g[:, 0], fx, rtol=1e-7, atol=1e-8)
        np.testing.assert_allclose(g[:, 1], fy, rtol=1e-7, atol=1e-8)


def test_mms_euler_source_is_flux_divergence():
    h = 1e-5
    for (x, y) in [(0.1, 0.2), (0.5, 0.5), (0.9, 0.3)]:
        S = oracle.source_term(x, y, viscosity=False)
        ex, ey = np.array([1.0, 0.0]), np.array([0.0, 1.0])
        div = (physical_flux(oracle.solution(x + h, y), ex) - physical_flux(oracle.solution(x - h, y), ex)) / (2 * h) \
            + (physical_flux(oracle.solution(x, y + h), ey) - physical_flux(oracle.solution(x, y - h), ey)) / (2 * h)
        np.testing.assert_allclose(S, div, rtol=1e-7, atol=1e-8)


def test_mms_viscous_source_momentum_matches_stress_divergence():
    """Momentum rows of the NS source (manufacturedsolution.d:200-239) equal -div(tau) of the
    exact solution with the textbook stress tau = mu(grad u + grad u^T - 2/3 div u I).  (The
    energy row mixes a dimensional heat term into a non-dimensional flux, which the reference's
    author flags in TheoryGuide2.tex:194-196; it is restated as written and not pinned.)"""
    h = 1e-4
    mu, Pr, Rg = 0.01, 0.71, 1.0

    def vel(x, y):
        q = oracle.solution(x, y)
        return q[1] / q[0], q[2] / q[0]

    def tau(x, y):
        e = 1e-5
        up, vp = vel(x + e, y); um, vm = vel(x - e, y)
        ux, vx = (up - um) / (2 * e), (vp - vm) / (2 * e)
        up, vp = vel(x, y + e); um, vm = vel(x, y - e)
        uy, vy = (up - um) / (2 * e), (vp - vm) / (2 * e)
        d = ux + vy
        return np.array([[2 * ux - 2.0 / 3.0 * d, uy + vx], [uy + vx, 2 * vy - 2.0 / 3.0 * d]]) * mu

    for (x, y) in [(0.2, 0.3), (0.6, 0.7)]:
        S = oracle.source_term(x, y, viscosity=True, R=Rg, mu=mu, Pr=Pr) - oracle.source_term(x, y, viscosity=False)
        div = (tau(x + h, y)[0] - tau(x - h, y)[0]) / (2 * h) + (tau(x, y + h)[1] - tau(x, y - h)[1]) / (2 * h)
        np.testing.assert_allclose(S[1:3], -div, rtol=2e-4, atol=1e-8)


def test_limiters_1d_textbook_values():
    # source/limiters.d:9-37
    assert oracle.limiter1d(0, 0.3, 0.7) == 0.0
    assert oracle.limiter1d(1, 0.3, 0.7) == 0.3
    assert oracle.limiter1d(2, 1.0, 2.0) == 1.0          # minmod
    assert oracle.limiter1d(2, -1.0, 2.0) == 1.0         # sign of the sum
    assert oracle.limiter1d(2, -1.0, -2.0) == -1.0
    assert abs(oracle.limiter1d(3, 1.0, 3.0) - 1.5) < 1e-11   # harmonic 2ab/(a+b)
    assert oracle.limiter1d(3, -1.0, 3.0) == 0.0
    assert abs(oracle.limiter1d(4, 1.0, 4.0) - 2.0) < 1e-15   # geometric sqrt(ab)
    assert oracle.limiter1d(4, -1.0, 4.0) == 0.0
    assert oracle.limiter1d(5, 1.0, 3.0) == 2.0           # superbee max(min(2a,b),min(a,2b))
    assert oracle.limiter1d(5, 1.0, 1.0) == 1.0


def test_lp_reaches_linprog_optimum():
    """LP!m (solve.d:37-173): maximise |g|.phi over 0<=phi<=1 with neighbour bounds; compare the
    objective with scipy's linprog on the same inequality system A x >= b."""
    scipy_opt = pytest.importorskip("scipy.optimize")
    rng = np.random.default_rng(4)
    checked = 0
    for _ in range(200):
        m = int(rng.choice([3, 4]))
        gx, gy = rng.uniform(-1, 1, 2)
        dx, dy = rng.uniform(-1, 1, (2, m))
        dq = rng.uniform(0.05, 0.5)
        A = [[1, 0], [0, 1], [-1, 0], [0, -1]]
        b = [0, 0, -1, -1]
        for j in range(m):
            A.append([dx[j] * gx, dy[j] * gy]); b.append(-dq)
            A.append([-dx[j] * gx, -dy[j] * gy]); b.append(-dq)
        A, b = np.array(A, float), np.array(b, float)
        c = np.array([-abs(gx), -abs(gy)])
        x = oracle.lp(A, b, c)
        res = scipy_opt.linprog(c, A_ub=-A, b_ub=-b, bounds=[(None, None)] * 2, method="highs")
        if res.status != 0:
            continue
        assert np.all(A @ x >= b - 1e-9)
        assert c @ x <= res.fun + 1e-9
        checked += 1
    assert checked > 150


def test_pairwise_sum_structure():
    """Phobos pairwise summation: blocks of 16 by a fixed tree, perfect trees over runs
    of blocks, remainder likewise; checked against an independent recursive form."""

    def tree(xs):
        if len(xs) == 1:
            return xs[0]
        h = len(xs) // 2
        return tree(xs[:h]) + tree(xs[h:])

    def recursive(xs):
        nb = len(xs) // 16
        parts = []
        k, pos = nb, 0
        for bit in reversed(range(64)):
            if k & (1 << bit):
                parts.append(tree([tree(xs[(pos + j) * 16:(pos + j + 1) * 16]) for j in range(1 << bit)]))
                pos += 1 << bit
        rem = xs[nb * 16:]
        r, pos = len(rem), 0
        for bit in reversed(range(5)):
            if r & (1 << bit):
                parts.append(tree(rem[pos:pos + (1 << bit)]))
                pos += 1 << bit
        s = parts[-1]
        for p in reversed(parts[:-1]):
            s = s + p
        return 0.0 + s

    rng = np.random.default_rng(8)
    for n in [1, 2, 15, 16, 17, 31, 33, 100, 257, 1000, 4097, 16 * 37 + 13]:
        x = rng.standard_normal(n) * 10.0 ** rng.uniform(-8, 8, n)
        assert oracle.pairwise_sum(x) == recursive(list(x)), n
        ints = np.round(rng.uniform(-1e6, 1e6, n))
        assert oracle.pairwise_sum(ints) == float(np.sum(ints))


def test_rusanov_wave_speed():
    """sMax of rusanovFlux (flux.d:88-131): for qL = qR it is |u.n| + a, the largest
    eigenvalue, whatever the averaging (the Harten fix only acts below 0.01 a); it is
    symmetric in the two states and in the sign of n; and for distinct states it lies
    between the wave speeds of the two states' own Harten-free maxima, since the
    averaged velocity, pressure and density are means of the two sides'."""
    rng = np.random.default_rng(7)
    for _ in range(50):
        q = random_state(rng)
        th = rng.uniform(0, 2 * np.pi)
        n = np.array([np.cos(th), np.sin(th)])
        _, sm = oracle.rusanov_flux(q, q, n)
        rho, u, v = q[0], q[1] / q[0], q[2] / q[0]
        p = (G - 1) * (q[3] - 0.5 * rho * (u * u + v * v))
        ref = abs(u * n[0] + v * n[1]) + math.sqrt(G * p / rho)
        assert abs(sm - ref) <= 4e-15 * ref, (sm, ref)
    for _ in range(50):
        qL, qR = random_state(rng), random_state(rng)
        n = np.array([0.6, -0.8])
        _, s1 = oracle.rusanov_flux(qL, qR, n)
        _, s2 = oracle.rusanov_flux(qR, qL, n)
        _, s3 = oracle.rusanov_flux(qL, qR, -n)
        assert s1 == s2 and abs(s1 - s3) <= 1e-15 * s1

        def speeds(q):
            rho, u, v = q[0], q[1] / q[0], q[2] / q[0]
            p = (G - 1) * (q[3] - 0.5 * rho * (u * u + v * v))
            return abs(u * n[0] + v * n[1]), p, rho

        (uL, pL, rL), (uR, pR, rR) = speeds(qL), speeds(qR)
        hi = max(uL, uR) + math.sqrt(G * max(pL, pR) / min(rL, rR))
        lo = math.sqrt(G * min(pL, pR) / max(rL, rR)) - max(uL, uR)
        assert lo - 1e-12 <= s1 <= hi + 1e-12, (lo, s1, hi)


def _mms_fields(x, y):
    q = oracle.solution(x, y)
    rho, u, v = q[0], q[1] / q[0], q[2] / q[0]
    p = (G - 1) * (q[3] - 0.5 * rho * (u * u + v * v))
    return rho, u, v, p


def _d(f, x, y, ax, h):
    if ax == 0:
        return (f(x + h, y) - f(x - h, y)) / (2 * h)
    return (f(x, y + h) - f(x, y - h)) / (2 * h)


def test_mms_viscous_source_energy_row():
    """Energy row of the NS source (manufacturedsolution.d:219-236), in two parts that
    finite differences of the exact fields fix independently of the restatement:
      * heat part, proportional to k = gamma R mu / ((gamma - 1) Pr): varying Pr at fixed
        mu changes only it.  It equals k times d2T/dx2 + d2T/dy2 of T = p / (rho R) plus
        the reference's extra 2 rho_x p_x / (R rho^2) + 2 rho_y p_y / (R rho^2) (its
        d2Tdx2 expands -(1/r^3)(r p_x - 2 p r_x) r_x, which cancels the +(1/r^2) r_x p_x
        term where the true derivative needs -2 (1/r^2) r_x p_x);
      * work part: d/dx(u tau_xx + v tau_xy) + d/dy(u tau_xy + v tau_yy) with the
        textbook stress.
    A dropped term, a sign flip or the textbook d2T/dx2 in place of the reference's
    fails one of the two."""
    mu, Rg = 0.01, 1.0
    k = lambda Pr: G * Rg * mu / ((G - 1) * Pr)
    # points where the reference's extra term is a sizable part of the heat term
    for (x, y) in [(0.5, 0.95), (0.3, 0.55), (0.15, 0.25)]:
        S_inv = oracle.source_term(x, y, viscosity=False)
        S1 = oracle.source_term(x, y, viscosity=True, R=Rg, mu=mu, Pr=0.71)
        S2 = oracle.source_term(x, y, viscosity=True, R=Rg, mu=mu, Pr=1.3)
        # S = S_inv - (work + k * heat)
        heat_src = -(S1[3] - S2[3]) / (k(0.71) - k(1.3))
        work_src = -(S1[3] - S_inv[3]) - k(0.71) * heat_src

        T = lambda a, b: _mms_fields(a, b)[3] / (_mms_fields(a, b)[0] * Rg)
        h = 1e-4
        Txx = (T(x + h, y) - 2 * T(x, y) + T(x - h, y)) / (h * h)
        Tyy = (T(x, y + h) - 2 * T(x, y) + T(x, y - h)) / (h * h)
        rho = _mms_fields(x, y)[0]
        rx = _d(lambda a, b: _mms_fields(a, b)[0], x, y, 0, 1e-6)
        ry = _d(lambda a, b: _mms_fields(a, b)[0], x, y, 1, 1e-6)
        px = _d(lambda a, b: _mms_fields(a, b)[3], x, y, 0, 1e-6)
        py = _d(lambda a, b: _mms_fields(a, b)[3], x, y, 1, 1e-6)
        extra = 2 * rx * px / (Rg * rho * rho) + 2 * ry * py / (Rg * rho * rho)
        np.testing.assert_allclose(heat_src, Txx + Tyy + extra, rtol=2e-5)
        assert abs(extra) > 0.05 * abs(Txx + Tyy + extra)   # dropping it would fail the check above

        def work_flux(a, b, ax):
            e = 1e-5
            _, u, v, _ = _mms_fields(a, b)
            ux = _d(lambda s, t: _mms_fields(s, t)[1], a, b, 0, e)
            uy = _d(lambda s, t: _mms_fields(s, t)[1], a, b, 1, e)
            vx = _d(lambda s, t: _mms_fields(s, t)[2], a, b, 0, e)
            vy = _d(lambda s, t: _mms_fields(s, t)[2], a, b, 1, e)
            txx = mu * (4.0 / 3.0 * ux - 2.0 / 3.0 * vy)
            tyy = mu * (4.0 / 3.0 * vy - 2.0 / 3.0 * ux)
            txy = mu * (uy + vx)
            return u * txx + v * txy if ax == 0 else u * txy + v * tyy

        hw = 1e-3
        div = _d(lambda a, b: work_flux(a, b, 0), x, y, 0, hw) + _d(lambda a, b: work_flux(a, b, 1), x, y, 1, hw)
        np.testing.assert_allclose(work_src, div, rtol=1e-4, atol=1e-9)
