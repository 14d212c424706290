"""Hand-built cases for the event loop of runSimulationCPU
(src/particle.cpp:712-837), with their expected results worked out without
the loop itself: from closed forms in exact rational arithmetic, or by
composing single integrator calls (whose own pins live in
test_oracle_integrators.py) the way the reference's loop composes them.

Shared by the oracle pins (test_oracle_event_loop.py) and the GPU parity
tests (test_gpu_event_loop.py).  Not a test module itself.
"""
from __future__ import annotations

import math
from fractions import Fraction as Fr

import numpy as np

K_BOLTZMANN = 1.380649e-23   # src/simulation.cpp:9
C_LIGHT = 299792458.0


# ---------------------------------------------------------------------------
# The reference's own end-to-end test, tests/output_tests.cpp:20-82
# ---------------------------------------------------------------------------
def output_test_fields() -> dict:
    """The options tests/output_tests.cpp:24-40 sets on top of the defaults
    of include/options.h:10-61 (gas collisions on, diffuse = 1.0 walls,
    'C' speeds of 5 m/s, m = 1.20239e-26 kg all stay at their defaults),
    plus the two derived values mainAnalysis adds (src/simulation.cpp:9-11)."""
    m, T = 1.20239e-26, 0.4
    return dict(B0=(3e-6, 0.0, 0.0), E=(0.0, 0.0, 0.0), T=T, tf=0.2, t0=0.0, rtol=1e-12, atol=1e-12,
                h=1e-5, yi=(0.0, 0.0, 1.0), gravity=False, ioutInt=0.05, integratorType=0,
                numParticles=3,
                tc=1.6e-4 * m / (K_BOLTZMANN * math.pow(T, 8.0)),
                sqrtKT_m=math.sqrt(K_BOLTZMANN * T / m))


def check_output_test(snaps, ioutInt=0.05, B0=3e-6, gamma=-2.038e8, tol=1e-9):
    """tests/output_tests.cpp:70-79 on snapshots 0..4: every particle's time
    is ti*ioutInt and its spin is (0, sin(B0*gamma*t), cos(B0*gamma*t)) --
    free precession about B0 from (0, 0, 1), whatever the particle's gas and
    wall collisions.  The reference compares with a relative tolerance of
    1e-9 (utf::tolerance); S.x is compared with 0, i.e. exactly."""
    for ti, s in enumerate(snaps):
        t_ref = ti * ioutInt
        ph = B0 * gamma * t_ref
        for n in range(len(s["t"])):
            t = float(s["t"][n])
            assert abs(t - t_ref) <= tol * abs(t_ref), (ti, n, t, t_ref)
            sx, sy, sz = (float(s["S"][n][k]) for k in range(3))
            assert sx == 0.0, (ti, n, sx)
            for got, want in ((sy, math.sin(ph)), (sz, math.cos(ph))):
                assert abs(got - want) <= tol * max(abs(got), abs(want)), (ti, n, got, want)


# ---------------------------------------------------------------------------
# An axis-aligned neutron between the x walls: closed-form segments
# ---------------------------------------------------------------------------
class AxisCase:
    """One neutron moving along x only (v = (+vx, 0, 0)) in a cell of width
    Lx, no gravity, no gas, specular walls, so its wall hits come every Lx/vx
    after the first at (Lx/2 - x0)/vx.  The reference's segments then follow
    from the breakpoints alone (src/particle.cpp:121-257): between two
    consecutive breakpoints (wall hits and output times) a gap g is covered by
    ceil(g / maxPosStep) segments -- 'N' segments of maxPosStep, the last one
    ending on the breakpoint -- and n_steps counts every segment
    (src/particle.cpp:802).  The field is B = (0, 0, B0z + g*x) + v x E / c^2
    with E = (0, Ey, 0), so the spin precesses about z by
    theta(t) = gamma*(B0z*t + g*int x dt + Ey/c^2*(x(t) - x0)) from
    S = (1, 0, 0): S = (cos theta, -sin theta, 0)."""

    def __init__(self, x0=0.0123, vx=5.0, Lx=0.07, cap=0.01, B0z=1e-6, g=1e-8, Ey=7.5e6,
                 gamma=-1.832e8):
        self.x0, self.vx, self.Lx, self.cap = Fr(x0), Fr(vx), Fr(Lx), Fr(cap)
        self.B0z, self.g, self.Ey, self.gamma = B0z, g, Ey, gamma

    def fields(self) -> dict:
        return dict(L=(float(self.Lx), 0.1, 0.4), B0=(0.0, 0.0, self.B0z), E=(0.0, self.Ey, 0.0),
                    Gx=(0.0, 0.0, 0.0), Gy=(0.0, 0.0, 0.0), Gz=(self.g, 0.0, 0.0), yi=(1.0, 0.0, 0.0),
                    gravity=False, gas_coll=False, diffuse=0.0, maxPosStep=float(self.cap),
                    m=1.675e-27, gamma=self.gamma, V=float(self.vx), dist="C")

    def start(self):
        """pos, v of the planted particle."""
        return (float(self.x0), 0.0, 0.0), (float(self.vx), 0.0, 0.0)

    def x_of(self, t: Fr):
        """Position and velocity sign at time t (closed form): the unfolded
        coordinate x0 + vx*t folded into [-Lx/2, Lx/2]."""
        h = self.Lx / 2
        per = 2 * self.Lx
        w = (self.x0 + self.vx * t + h) % per
        if w <= self.Lx:
            return w - h, 1
        return h - (w - self.Lx), -1

    def bounce_times(self, T: Fr):
        """Wall-hit times in (0, T]."""
        t1 = (self.Lx / 2 - self.x0) / self.vx
        P = self.Lx / self.vx
        out, k = [], 0
        while t1 + k * P <= T:
            out.append(t1 + k * P)
            k += 1
        return out

    def counts(self, targets):
        """(n_bounce, n_steps) after each output time: ceil(gap / cap)
        segments per gap between consecutive breakpoints."""
        T = [Fr(t) for t in targets]
        bps = sorted(set(self.bounce_times(T[-1])) | set(T))
        res, nb, ns, prev = [], 0, 0, Fr(0)
        bset = set(self.bounce_times(T[-1]))
        for b in bps:
            gap = b - prev
            ns += -((-gap) // self.cap)          # ceil
            if b in bset:
                nb += 1
            if b in T:
                res.append((nb, ns))
            prev = b
        return res

    def int_x(self, T: Fr):
        """int_0^T x dt, piecewise over the wall hits (x is linear between)."""
        pts = [Fr(0)] + self.bounce_times(T) + [T]
        tot = Fr(0)
        for a, b in zip(pts[:-1], pts[1:]):
            if b <= a:
                continue
            xa, sa = self.x_of(a)
            # on (a, b) the velocity sign is that just after a
            mid = (a + b) / 2
            _, s = self.x_of(mid)
            xb = xa + s * self.vx * (b - a)
            tot += (xa + xb) / 2 * (b - a)
        return tot

    def expected(self, targets):
        """Per output time: dict(t, x, vx, S, n_bounce, n_steps)."""
        out = []
        for (nb, ns), t in zip(self.counts(targets), targets):
            T = Fr(t)
            x, s = self.x_of(T)
            ix = self.int_x(T)
            theta = self.gamma * (self.B0z * t + self.g * float(ix)
                                  + self.Ey / C_LIGHT ** 2 * float(x - self.x0))
            out.append(dict(t=t, x=float(x), vx=float(s * self.vx),
                            S=(math.cos(theta), -math.sin(theta), 0.0), n_bounce=nb, n_steps=ns))
        return out


def check_axis_case(case: AxisCase, targets, snaps, tol_S=1e-9):
    """Compares per-interval snapshots of the one planted particle with the
    closed form: counts exactly, t to rounding, x to 1e-12, v exactly, S to
    tol_S."""
    exp = case.expected(targets)
    for k, (e, s) in enumerate(zip(exp, snaps)):
        assert int(s["n_bounce"][0]) == e["n_bounce"], (k, int(s["n_bounce"][0]), e)
        assert int(s["n_steps"][0]) == e["n_steps"], (k, int(s["n_steps"][0]), e)
        assert abs(float(s["t"][0]) - e["t"]) < 1e-13, (k, float(s["t"][0]), e["t"])
        p, v, S = (np.array([float(s[f][0][c]) for c in range(3)]) for f in ("pos", "v", "S"))
        assert abs(p[0] - e["x"]) < 1e-12 and p[1] == 0.0 and p[2] == 0.0, (k, p, e["x"])
        assert v[0] == e["vx"] and v[1] == 0.0 and v[2] == 0.0, (k, v, e["vx"])
        dS = np.abs(S - np.array(e["S"])).max()
        assert dS < tol_S, (k, dS, S, e["S"])
    return exp
