"""Pins the oracle's integratorTypes 2-5 (src/integrator.cpp:339-723):
integrateMagnusCFET (2), integrateRK45Quaternion (3), integrateRKF45 (4) and
integrateRKF45Quaternion (5).

Independent anchors:
* Fehlberg's RKF4(5) tableau (include/coeff.h:220-256) satisfies the order-5
  (weights C) and order-4 (weights CR) conditions in exact rationals;
* RKF45 steps with the stale first stage equal a numpy restatement of the
  tableau recursion;
* RKF45's "error" is the length of the fourth-order increment, so its steps
  settle where h |S x gamma B| ~ rtol -- a step density no embedded error
  estimate would give;
* on a field of fixed direction whose size is quadratic in time the
  quaternion and Magnus integrators are exact (their weights integrate
  quadratics exactly), so they match the closed-form precession angle;
* single steps on a field that turns equal compositions of rotations from
  scipy.spatial.transform (an independent quaternion library);
* the Magnus integrator converges at order 8 against the oracle's DOP853,
  itself pinned to scipy's DOP853.
"""
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

from test_oracle_integrators import field_opt, np_rhs

GAMMA = -1.832e8

# Fehlberg RKF4(5), include/coeff.h:220-256 (published tableau)
F_A = [Fr(0), Fr(1, 4), Fr(3, 8), Fr(12, 13), Fr(1), Fr(1, 2)]
F_B = [[],
       [Fr(1, 4)],
       [Fr(3, 32), Fr(9, 32)],
       [Fr(1932, 2197), Fr(-7200, 2197), Fr(7296, 2197)],
       [Fr(439, 216), Fr(-8), Fr(3680, 513), Fr(-845, 4104)],
       [Fr(-8, 27), Fr(2), Fr(-3544, 2565), Fr(1859, 4104), Fr(-11, 40)]]
F_C = [Fr(16, 135), Fr(0), Fr(6656, 12825), Fr(28561, 56430), Fr(-9, 50), Fr(2, 55)]
F_CR = [Fr(25, 216), Fr(0), Fr(1408, 2565), Fr(2197, 4104), Fr(-1, 5), Fr(0)]
# RK45 weights of the quaternion integrator (include/coeff.h:178-218)
R_A = [0.0, 2 / 9, 1 / 3, 3 / 4, 1.0, 5 / 6]
R_CH = [47 / 450, 0.0, 12 / 25, 32 / 225, 1 / 30, 6 / 25]
R_C = [1 / 9, 0.0, 9 / 20, 16 / 45, 1 / 12]


def order_conditions(b, c, A, order):
    for i in range(1, 6):
        assert sum(A[i]) == c[i]
    Ac = [sum(A[i][j] * c[j] for j in range(i)) for i in range(6)]
    Ac2 = [sum(A[i][j] * c[j] ** 2 for j in range(i)) for i in range(6)]
    AAc = [sum(A[i][j] * Ac[j] for j in range(i)) for i in range(6)]
    S = lambda f: sum(b[i] * f(i) for i in range(6))
    conds = [(1, lambda i: 1, 1), (2, lambda i: c[i], Fr(1, 2)), (3, lambda i: c[i] ** 2, Fr(1, 3)),
             (3, lambda i: Ac[i], Fr(1, 6)), (4, lambda i: c[i] ** 3, Fr(1, 4)),
             (4, lambda i: c[i] * Ac[i], Fr(1, 8)), (4, lambda i: Ac2[i], Fr(1, 12)),
             (4, lambda i: AAc[i], Fr(1, 24)), (5, lambda i: c[i] ** 4, Fr(1, 5))]
    for p, f, want in conds:
        if p <= order:
            assert S(f) == want, (p, want)


def test_rkf45_tableau_orders():
    order_conditions(F_C, F_A, F_B, 5)
    order_conditions(F_CR, F_A, F_B, 4)
    # CR is a solution weight vector, not an error vector: it sums to 1
    assert sum(F_CR) == 1


def rot_angle_closed_form(o, t0, t1, p0, v0):
    """gamma * int B_x dt for B = B0x + Gx . r(t), r quadratic (gravity on)."""
    g = np.array(o["Gx"])
    a = np.array([0.0, -9.81, 0.0])
    d = t1 - t0
    integral = o["B0"][0] * d + g @ (np.asarray(p0) * d + np.asarray(v0) * d * d / 2 + a * d ** 3 / 6)
    return o["gamma"] * integral


def commuting_opt(oracle, **kw):
    o = oracle.default_options()
    o["B0"] = (3e-6, 0.0, 0.0)
    o["E"] = (0.0, 0.0, 0.0)
    o["Gx"] = (2e-4, 1e-4, -1.5e-4)      # B_x changes by ~20% per ms
    o["Gy"] = (0.0, 0.0, 0.0)
    o["Gz"] = (0.0, 0.0, 0.0)
    o["gravity"] = True
    o["gamma"] = GAMMA
    for k, v in kw.items():
        o[k] = v
    return o


@pytest.mark.parametrize("integ,rtol", [(2, 1e-12), (3, 1e-12), (5, 1e-3)])
@pytest.mark.parametrize("fixed", [False, True])
def test_exact_on_commuting_quadratic_field(oracle, integ, rtol, fixed):
    # field along x only: the exact solution rotates (Sy, Sz) by the integral
    # of gamma*B_x, which Gauss (Magnus) and RK45/RKF45 (quaternion) weights
    # integrate exactly for a quadratic in t
    p0, v0 = np.array([0.01, -0.02, 0.1]), np.array([3.0, -2.0, 4.0])
    T, y0 = 0.05, np.array([0.2, 0.6, math.sqrt(1 - 0.4)])
    if fixed:
        n = 50
        o = commuting_opt(oracle, rtol=rtol, fixedStepSize=True, h=T / n, nmax=n)
        st, y, h, na, nr = oracle.integrate(o, 0.0, T, y0, p0, v0, T / n, integrator=integ)
        # a fixed step size never ends these calls: they stop at nmax with -2
        # after exactly nmax steps of h
        assert st == -2 and na == n
        tend = n * (T / n)
    else:
        o = commuting_opt(oracle, rtol=rtol, hmax=1e-3)
        st, y, h, na, nr = oracle.integrate(o, 0.0, T, y0, p0, v0, 1e-4, integrator=integ)
        assert st == 0 and na >= 50
        tend = T
    phi = rot_angle_closed_form(o, 0.0, tend, p0, v0)
    want = np.array([y0[0], y0[1] * math.cos(phi) + y0[2] * math.sin(phi),
                     y0[2] * math.cos(phi) - y0[1] * math.sin(phi)])
    assert abs(phi) > 20                    # many turns, strongly non-uniform rate
    assert np.abs(y - want).max() < 2e-12, (np.abs(y - want).max(), na, nr)


def test_cross_term_matches_numpy(oracle):
    o = field_opt(oracle, a=2e-6, w=377.0)
    rng = np.random.default_rng(5)
    for _ in range(20):
        p0, v0 = rng.uniform(-0.03, 0.03, 3), rng.normal(0, 5, 3)
        t0, t = rng.uniform(0, 1), rng.uniform(0, 1)
        w = oracle.cross_term(o, t, t0, p0, v0)
        for y in np.eye(3):
            assert np.allclose(np.cross(y, w), np_rhs(o, t, y, t0, p0, v0), rtol=1e-13, atol=1e-16)


def turning_opt(oracle, **kw):
    # gradients comparable to B0 across the cell: the field direction turns
    o = field_opt(oracle)
    o["Gx"] = (1e-4, -2e-4, 3e-4)
    o["Gy"] = (2e-4, 1e-4, -1e-4)
    o["Gz"] = (-3e-4, 2e-4, 1e-4)
    for k, v in kw.items():
        o[k] = v
    return o


def one_fixed_step(oracle, o, integ, t0, h, y, p0, v0):
    o = o.copy()
    o["fixedStepSize"], o["h"], o["nmax"] = True, h, 1
    st, yo, ho, na, nr = oracle.integrate(o, t0, t0 + 10 * h, y, p0, v0, h, integrator=integ)
    assert st == -2 and na == 1
    return yo


@pytest.mark.parametrize("integ", [3, 5])
def test_quaternion_step_is_a_rotation_product(oracle, integ):
    from scipy.spatial.transform import Rotation
    o = turning_opt(oracle)
    A = R_A if integ == 3 else [float(x) for x in F_A]
    Wt = R_CH if integ == 3 else [float(x) for x in F_C]
    rng = np.random.default_rng(6)
    for _ in range(20):
        y = rng.normal(size=3)
        y /= np.linalg.norm(y)
        p0, v0 = rng.uniform(-0.03, 0.03, 3), rng.normal(0, 5, 3)
        t0, h = rng.uniform(0, 1), 10 ** rng.uniform(-4, -2.5)
        yo = one_fixed_step(oracle, o, integ, t0, h, y, p0, v0)
        rot = Rotation.identity()
        for a, wgt in zip(A, Wt):
            k = oracle.cross_term(o, t0 + a * h, t0, p0, v0)
            # dS/dt = S x k turns S by -|k| dt about k
            rot = Rotation.from_rotvec(-k * h * wgt) * rot
        assert np.abs(yo - rot.apply(y)).max() < 5e-15


CF8 = np.array([
    [-0.005984812750244602, 0.020217897716672675, -0.03646660716686794, 0.04245408899634467, 0.14949496424802844],
    [0.02512881778391826, -0.08955337580937157, 0.19032828600393248, 0.28920194841768715, -0.03568486888016069],
    [-0.04972613261435144, 0.6572496090804754, 0.11725149008224456, -0.43611911073577403, 0.18080345083145516],
    [0.0464837870954061, -0.6265183532355708, -0.11144873822286155, 0.4154988625423615, -0.17224148557040517],
    [0.015567390500242426, -0.0677120111619075, -0.15241336773272682, -0.08803798509100481, -0.0013284996209207986],
    [-0.034049168494873015, 0.12263276452977083, 0.2699423185170027, 0.12263276452977094, -0.03404916849487304]])
GL5 = np.array([0.5 - math.sqrt(5 + 2 * math.sqrt(10 / 7)) / 6, 0.5 - math.sqrt(5 - 2 * math.sqrt(10 / 7)) / 6,
                0.5, 0.5 + math.sqrt(5 - 2 * math.sqrt(10 / 7)) / 6, 0.5 + math.sqrt(5 + 2 * math.sqrt(10 / 7)) / 6])


def test_magnus_coefficients_consistent():
    # the 11 exponents k11..k1 sum to h * sum_i w_i B_i with w the 5-point
    # Gauss-Legendre weights on [0, 1]: the scheme is exact for a constant field
    w = np.array([(322 - 13 * math.sqrt(70)) / 900, (322 + 13 * math.sqrt(70)) / 900, 128 / 225,
                  (322 + 13 * math.sqrt(70)) / 900, (322 - 13 * math.sqrt(70)) / 900]) / 2
    tot = CF8[:5, ::-1].sum(axis=0) + CF8[5] + CF8[:5].sum(axis=0)
    assert np.allclose(tot, w, rtol=0, atol=1e-15)
    # the oracle's nodes are the 5-point Gauss-Legendre nodes on [0, 1]
    assert np.allclose(GL5, [0.046910077030668074, 0.2307653449471585, 0.5, 0.7692346550528415,
                             0.9530899229693319], rtol=0, atol=1e-16)


def test_magnus_step_is_a_rotation_product(oracle):
    from scipy.spatial.transform import Rotation
    o = turning_opt(oracle)
    rng = np.random.default_rng(7)
    rows = [CF8[r, ::-1] for r in range(5)] + [CF8[5]] + [CF8[r] for r in range(4, -1, -1)]
    for _ in range(20):
        y = rng.normal(size=3)
        y /= np.linalg.norm(y)
        p0, v0 = rng.uniform(-0.03, 0.03, 3), rng.normal(0, 5, 3)
        t0, h = rng.uniform(0, 1), 10 ** rng.uniform(-4, -2.5)
        yo = one_fixed_step(oracle, o, 2, t0, h, y, p0, v0)
        B = np.array([oracle.cross_term(o, t0 + x * h, t0, p0, v0) for x in GL5])
        rot = Rotation.identity()
        for c in rows:                       # k11 acts first, k1 last
            rot = Rotation.from_rotvec(-(c @ B) * h) * rot
        assert np.abs(yo - rot.apply(y)).max() < 5e-15


def test_magnus_order_8(oracle):
    o = turning_opt(oracle)
    y0, p0, v0 = np.array([1.0, 0, 0]), np.array([0.01, -0.02, 0.1]), np.array([3.0, -2.0, 4.0])
    T = 0.005           # |gamma B| T ~ 25 rad; the DOP853 reference is good to ~3e-13
    od = o.copy()
    od["fixedStepSize"] = True
    ref = oracle.integrate(od, 0.0, T, y0, p0, v0, T / 3072 * 0.999999)[1]
    errs = []
    for n in (12, 24, 48):
        of = o.copy()
        of["fixedStepSize"], of["h"], of["nmax"] = True, T / n, n
        st, y, h, na, nr = oracle.integrate(of, 0.0, T, y0, p0, v0, T / n, integrator=2)
        assert st == -2 and na == n
        errs.append(np.abs(y - ref).max())
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(7.3 < q < 8.9 for q in orders), (errs, orders)


def test_rkf45_steps_with_stale_k1(oracle):
    """Two RKF45 steps on a constant field: the first k1 is f(t, 0) = 0, the
    second f(t + h, stage-6 input of step 1).  rtol is huge so both are
    accepted; hmax = h keeps the first at h, and the second ends the call."""
    o = oracle.default_options()
    o["B0"] = (3e-6, 1e-6, 0.0)
    o["E"] = (0.0, 0.0, 0.0)
    o["gravity"] = False
    o["rtol"] = 1e6
    h = 1e-3
    o["hmax"] = h
    w = o["gamma"] * np.array(o["B0"])
    f = lambda y: np.cross(y, w)
    A, B, C = [float(x) for x in F_A], [[float(x) for x in r] for r in F_B], [float(x) for x in F_C]
    y = np.array([0.0, 0.0, 1.0])
    stale = np.zeros(3)
    for _ in range(2):
        k = [f(stale)]
        for s in range(1, 6):
            yy = y + h * sum(B[s][j] * k[j] for j in range(s))
            k.append(f(yy))
            if s == 5:
                stale = yy
        y = y + h * sum(C[j] * k[j] for j in range(6))
    st, yo, ho, na, nr = oracle.integrate(o, 0.0, 2 * h, (0, 0, 1), (0, 0, 0), (0, 0, 0), h, integrator=4)
    assert st == 0 and na == 2 and nr == 0
    assert np.allclose(yo, y, rtol=0, atol=1e-14), (yo, y)
    # the call returns the step it last proposed (4h after q = 4), not the
    # truncated one: lastH is set before the end-of-interval truncation
    assert ho == 4 * h


def test_rkf45_error_is_the_increment(oracle):
    # |h * sum CR_i k_i| ~ h |S x w|: steps settle at h ~ rtol / |S x w|
    o = oracle.default_options()
    o["B0"] = (3e-6, 0.0, 0.0)
    o["E"] = (0.0, 0.0, 0.0)
    o["gravity"] = False
    T = 1e-3
    w = o["gamma"] * 3e-6                  # S' = S x (w, 0, 0)
    want = np.array([0.0, math.sin(w * T), math.cos(w * T)])
    o["hmin"] = 1e-12
    errs, steps, drift = [], [], []
    for rtol in (1e-4, 5e-5):
        o["rtol"] = rtol
        # a first step of 1e-9: its k1 = f(t, 0) = 0 then costs only ~1e-10
        st, y, h, na, nr = oracle.integrate(o, 0.0, T, (0, 0, 1), (0, 0, 0), (0, 0, 0), 1e-9, integrator=4)
        assert st == 0
        assert 0.7 < na * rtol / (T * abs(w)) < 1.6, na
        errs.append(np.abs(y - want).max())
        steps.append(na)
        drift.append(abs(np.linalg.norm(y) - 1))
    # the stale first stage (f at the previous step's sixth-stage input) makes
    # the method first order: the error and the |S| drift halve with h
    assert errs[0] < 1e-5 and 1.7 < errs[0] / errs[1] < 2.4, errs
    assert 1.7 < drift[0] / drift[1] < 2.4 and 1.8 < steps[1] / steps[0] < 2.2, (drift, steps)
    # an rtol that would ask for steps below hmin fails with -1
    o["rtol"] = 1e-12
    st, *_ = oracle.integrate(o, 0.0, T, (0, 0, 1), (0, 0, 0), (0, 0, 0), 1e-7, integrator=4)
    assert st == -1


@pytest.mark.parametrize("integ", [2, 3, 4, 5])
def test_nmax_failure(oracle, integ):
    o = turning_opt(oracle, nmax=3)
    st, y, h, na, nr = oracle.integrate(o, 0.0, 1.0, (1, 0, 0), (0, 0, 0), (1, 1, 1), 1e-6, integrator=integ)
    assert st == -2 and na + nr == 3
