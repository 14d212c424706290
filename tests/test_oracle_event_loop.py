"""Pins the oracle's event loop (run_simulation in oracle/st_oracle.c), i.e.
how runSimulationCPU (src/particle.cpp:712-837) composes the individually
pinned kinematics and integrator calls:

* the reference's own end-to-end test (tests/output_tests.cpp:20-82):
  three particles with gas collisions and fully diffuse walls precess freely;
* an axis-aligned neutron whose wall hits, 'N' segments, maxPosStep caps,
  n_bounce, n_steps (src/particle.cpp:802), position and spin phase all
  follow from closed forms in exact arithmetic;
* a restatement of the loop body in Python over the pinned primitives, which
  carries the integrator's truncated last step across the segments of an
  interval (tempH, src/particle.cpp:752,761; integrator.cpp:226-229), keeps it
  across intervals only with keepStepSize (:783-784), NaN-poisons a stopped
  particle (:790-801) and counts every pass (:802) -- compared bit for bit
  with the oracle, with negative controls showing that a tempH reset per
  segment, or a keepStepSize mix-up, would be caught.
"""
import math

import numpy as np
import pytest

from event_loop_cases import AxisCase, check_axis_case, check_output_test, output_test_fields


def make_opt(oracle, **kw):
    o = oracle.default_options()
    for k, v in kw.items():
        o[k] = v.encode() if k == "dist" else v
    return o


def ensemble(oracle, o, n, plant=None):
    E = oracle.Ensemble(o, n)
    if plant is not None:
        pos, v = plant
        E.arr["pos"][0] = pos
        E.arr["v"][0] = v
    return E


def snaps_of(E, targets):
    out = []
    for t in targets:
        E.advance_to(t)
        out.append(E.snapshot())
    return out


# ---------------------------------------------------------------------------
def test_reference_output_test(oracle):
    """tests/output_tests.cpp:20-82: mainAnalysis with N = 3, tf = 0.2,
    ioutInt = 0.05, gas on, diffuse 1.0, no gravity, B0 only: t and S at
    every snapshot equal free precession (relative 1e-9; S.x exactly 0)."""
    f = output_test_fields()
    o = make_opt(oracle, **f)
    E = oracle.Ensemble(o, 3)
    snaps = [E.snapshot()] + snaps_of(E, [(i + 1.0) * 0.05 for i in range(4)])
    check_output_test(snaps, B0=3e-6, gamma=float(o["gamma"]))
    last = snaps[-1]
    # the case exercises the composition: hundreds of gas and wall segments
    assert (last["n_coll"] > 100).all() and (last["n_bounce"] > 0).all()
    assert (last["n_steps"] >= last["n_coll"] + last["n_bounce"]).all()


def test_output_test_planted_mistake_fails(oracle):
    """The same check rejects a run that loses the first millisecond of
    spin precession (the particles start their first segment at t = 0.001;
    their output times still come out at ti*ioutInt): one way a mis-composed
    loop would show up."""
    f = output_test_fields()
    o = make_opt(oracle, **f)
    E = oracle.Ensemble(o, 3)
    E.arr["t"][:] = 0.001
    snaps = [E.snapshot()] + snaps_of(E, [(i + 1.0) * 0.05 for i in range(4)])
    snaps[0]["t"][:] = 0.0
    check_output_test(snaps[:1], B0=3e-6, gamma=float(o["gamma"]))
    with pytest.raises(AssertionError):
        check_output_test(snaps, B0=3e-6, gamma=float(o["gamma"]))


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("cap,x0", [(0.01, 0.0123), (0.1, -0.0211), (0.0031, 0.0301)])
def test_axis_aligned_closed_form(oracle, cap, x0):
    """Wall hits every Lx/vx, 'N' segments at maxPosStep, and the interval
    ends: n_bounce and n_steps per interval, x, v and the spin phase
    gamma*int B dt (gradient and motional field included) from closed
    forms."""
    case = AxisCase(x0=x0, cap=cap)
    o = make_opt(oracle, **case.fields(), tf=0.5, ioutInt=0.05, rtol=1e-12, atol=1e-12, h=1e-3)
    targets = [(i + 1.0) * 0.05 for i in range(10)]
    E = ensemble(oracle, o, 1, case.start())
    exp = check_axis_case(case, targets, snaps_of(E, targets))
    # wall hits and interval-end segments always occur; caps when the cap is
    # shorter than the gap between wall hits (Lx/vx = 0.014 s)
    assert exp[-1]["n_bounce"] > 20 and exp[-1]["n_steps"] >= exp[-1]["n_bounce"] + 10
    if cap < 0.014:
        assert exp[-1]["n_steps"] > exp[-1]["n_bounce"] + 10


def test_axis_closed_form_sensitivity():
    """The closed form sees the terms a mis-composed segment would get
    wrong: the motional field's sign (a segment integrated with the
    post-collision velocity flips it) moves the spin by far more than 1e-9,
    and 'N' segments at the cap make n_steps exceed the wall hits plus the
    interval ends in every interval."""
    case = AxisCase(x0=0.0123, cap=0.01)
    targets = [(i + 1.0) * 0.05 for i in range(10)]
    exp = case.expected(targets)
    bad = AxisCase(x0=0.0123, cap=0.01, Ey=-7.5e6).expected(targets)
    assert max(abs(b["S"][i] - e["S"][i]) for b, e in zip(bad, exp) for i in range(2)) > 1e-5
    prev = (0, 0)
    for k, e in enumerate(exp):
        assert e["n_steps"] - prev[1] > e["n_bounce"] - prev[0] + 1, (k, e)
        prev = (e["n_bounce"], e["n_steps"])


# ---------------------------------------------------------------------------
# The loop body restated over the pinned primitives
# ---------------------------------------------------------------------------
class Lane:
    """One particle's state as runSimulationCPU loads it (src/particle.cpp:724-746)."""

    def __init__(self, E, j=0):
        a = E.arr
        self.S = np.array(list(a["S"][j]))
        self.v = np.array(list(a["v"][j]))
        self.pos = np.array(list(a["pos"][j]))
        self.v_old = np.array(list(a["v_old"][j]))
        self.pos_old = np.array(list(a["pos_old"][j]))
        self.t, self.t_old = float(a["t"][j]), float(a["t_old"][j])
        self.next_gas, self.h = float(a["next_gas"][j]), float(a["h"][j])
        self.rng = a["state"][j:j + 1].copy().reshape(())
        self.n_bounce, self.n_coll, self.n_steps = int(a["n_bounce"][j]), int(a["n_coll"][j]), int(a["n_steps"][j])
        self.failure, self.stop = int(a["failure"][j]), bool(a["stop"][j])
        self.coll_type, self.wall_hit = chr(a["coll_type"][j]), chr(a["wall_hit"][j])
        self.acc = self.rej = self.seg = 0


def loop_interval(oracle, o, L: Lane, next_t: float, integrator=0, reset_temph=False):
    """src/particle.cpp:748-803 for one particle: dt = nextTOut - t,
    tempH = h; per pass: collision time, move, sanity, one integrator call
    from tempH (which the call leaves at its truncated last step), h = tempH
    under keepStepSize, stop on a negative status, NaN-poison when stopped,
    n_steps += 1.  reset_temph restarts every segment from h (a misreading,
    for the negative control)."""
    tf = next_t
    finished = False
    tempH = L.h
    while not finished and not L.stop:
        c = oracle.calc_next_collision_time(o, L.t, tf, L.v, L.pos, L.next_gas, L.rng, L.n_bounce, L.n_coll)
        L.next_gas, L.n_bounce, L.n_coll = c["next_gas"], c["n_bounce"], c["n_coll"]
        L.coll_type, L.wall_hit, finished = c["coll_type"], c["wall_hit"], c["finished"]
        u = oracle.update_position_and_velocity(o, L.t, c["dt"], L.pos, L.v, L.coll_type, L.wall_hit, L.rng)
        L.t_old, L.t = u["t_old"], u["t"]
        L.pos_old, L.pos, L.v_old, L.v = u["pos_old"], u["pos"], u["v_old"], u["v"]
        L.stop = L.stop or u["stop"]
        stop, fail = oracle.sanity_check(o, L.t_old, L.t, L.pos_old, L.pos, L.v, L.v_old)
        if fail:
            L.stop, L.failure = True, fail
        if L.failure == 0:
            if reset_temph:
                tempH = L.h
            status, y, tempH, na, nr = oracle.integrate(o, L.t_old, L.t, L.S, L.pos_old, L.v_old, tempH,
                                                        integrator)
            L.S = y
            L.acc += na
            L.rej += nr
            L.seg += 1
            if o["keepStepSize"]:
                L.h = tempH
            if status < 0:
                L.stop, L.failure = True, status
        if L.stop:
            L.t = math.nan
            L.pos = np.full(3, math.nan)
            L.v = np.full(3, math.nan)
            L.S = np.full(3, math.nan)
        L.n_steps += 1


def same_bits(a, b):
    return np.asarray(a, dtype=np.float64).tobytes() == np.asarray(b, dtype=np.float64).tobytes()


def assert_lane_equals(E, L: Lane, j=0):
    a = E.arr
    for k in ("S", "v", "pos", "v_old", "pos_old"):
        assert same_bits(list(a[k][j]), getattr(L, k)), (k, list(a[k][j]), getattr(L, k))
    for k in ("t", "t_old", "h", "next_gas"):
        assert same_bits(a[k][j], getattr(L, k)), (k, a[k][j], getattr(L, k))
    assert (int(a["n_bounce"][j]), int(a["n_coll"][j]), int(a["n_steps"][j])) == (L.n_bounce, L.n_coll, L.n_steps)
    assert (int(a["failure"][j]), bool(a["stop"][j])) == (L.failure, L.stop)
    for f in ("x", "y", "z", "w", "spare", "hasSpare"):
        assert a["state"][f][j] == L.rng[f], f
    n_ode = E.n_ode.reshape(-1, 3)[j]
    assert (int(n_ode[0]), int(n_ode[1]), int(n_ode[2])) == (L.acc, L.rej, L.seg)


def axis_opt(oracle, case, **kw):
    return make_opt(oracle, **case.fields(), tf=0.2, ioutInt=0.05, rtol=1e-12, atol=1e-12, h=1e-3, **kw)


def he3_opt(oracle, **kw):
    """One He-3 atom with gravity, gas collisions (tc = 2e-3 s) and 5%
    diffuse walls: every kind of segment and RNG draw."""
    base = dict(L=(0.07, 0.1, 0.4), B0=(3e-6, 0.0, 0.0), E=(0.0, 0.0, 7.5e6), Gx=(0.0, 0.0, 1e-8),
                yi=(1.0, 0.0, 0.0), m=5.008e-27, gamma=-2.038e8, dist="M", T=0.45, diffuse=0.05,
                gas_coll=True, gravity=True, tc=2e-3, sqrtKT_m=math.sqrt(1.380649e-23 * 0.45 / 5.008e-27),
                rtol=1e-12, atol=1e-12, h=1e-3, tf=0.2, ioutInt=0.05, seed=1)
    base.update(kw)
    return make_opt(oracle, **base)


CASES = ["axis", "he3"]


def build_case(oracle, which, **kw):
    if which == "axis":
        case = AxisCase(x0=0.0123, cap=0.01)
        o = axis_opt(oracle, case, **kw)
        return o, ensemble(oracle, o, 1, case.start())
    o = he3_opt(oracle, **kw)
    return o, ensemble(oracle, o, 1)


@pytest.mark.parametrize("which", CASES)
@pytest.mark.parametrize("integrator", [0, 1])
def test_loop_composition_bitwise(oracle, which, integrator):
    """Two intervals: the oracle's run_simulation equals the restated loop
    bit for bit -- spin, step size, kinematics, RNG, counters."""
    o, E = build_case(oracle, which, integratorType=integrator)
    L = Lane(E)
    for t in (0.05, 0.1):
        E.advance_to(t)
        loop_interval(oracle, o, L, t, integrator)
        assert_lane_equals(E, L)
    assert L.seg > 4 and L.n_steps == L.seg


@pytest.mark.parametrize("which", CASES)
def test_temph_carry_is_observable(oracle, which):
    """Negative control: restarting every segment from h instead of the
    previous segment's truncated last step changes the spin bits and the
    step counts, so the bitwise comparison above pins the carry."""
    o, E = build_case(oracle, which)
    L, R = Lane(E), Lane(E)
    loop_interval(oracle, o, L, 0.05)
    loop_interval(oracle, o, R, 0.05, reset_temph=True)
    assert not same_bits(L.S, R.S)
    assert (L.acc, L.rej) != (R.acc, R.rej)


@pytest.mark.parametrize("which", CASES)
@pytest.mark.parametrize("keep", [False, True])
def test_keep_step_size(oracle, which, keep):
    """keepStepSize (src/particle.cpp:783-784): h becomes the last segment's
    tempH and the next interval starts from it; otherwise h stays the
    initial step.  Both interval sequences are checked bit for bit, and the
    two settings really differ."""
    o, E = build_case(oracle, which, keepStepSize=keep)
    L = Lane(E)
    h0 = L.h
    for t in (0.05, 0.1, 0.15):
        E.advance_to(t)
        loop_interval(oracle, o, L, t)
        assert_lane_equals(E, L)
        if keep:
            assert L.h != h0
        else:
            assert float(E.arr["h"][0]) == h0
    # the other setting gives other bits
    o2, E2 = build_case(oracle, which, keepStepSize=not keep)
    for t in (0.05, 0.1, 0.15):
        E2.advance_to(t)
    assert not same_bits(list(E2.arr["S"][0]), L.S)


@pytest.mark.parametrize("integrator,code", [(0, -1), (1, -2)])
def test_stop_poisons_and_freezes(oracle, integrator, code):
    """A failing integrator call (nmax exceeded: integrateDOP -1,
    integrateRK45 -2) stops the particle at that pass: failure = status,
    t/pos/v/S NaN, the pass still counted in n_steps, pos_old/v_old/t_old
    left at the failing segment's start; later intervals change nothing."""
    case = AxisCase(x0=0.0123, cap=0.01)
    o = axis_opt(oracle, case, integratorType=integrator, nmax=3)
    E = ensemble(oracle, o, 1, case.start())
    L = Lane(E)
    E.advance_to(0.05)
    loop_interval(oracle, o, L, 0.05, integrator)
    assert_lane_equals(E, L)
    a = E.arr
    assert int(a["failure"][0]) == code and bool(a["stop"][0])
    assert math.isnan(a["t"][0]) and all(math.isnan(x) for x in list(a["pos"][0]) + list(a["v"][0]) + list(a["S"][0]))
    # the first segment failed: one bounce counted, one pass, its start kept
    assert int(a["n_steps"][0]) == 1 and int(a["n_bounce"][0]) == 1
    assert list(a["pos_old"][0]) == list(case.start()[0]) and list(a["v_old"][0]) == list(case.start()[1])
    assert float(a["t_old"][0]) == 0.0
    before = {k: a[k].tobytes() for k in a}
    E.advance_to(0.1)
    for k in a:
        if k not in ("tf", "dt"):
            assert a[k].tobytes() == before[k], k
