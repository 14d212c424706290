"""Pins the oracle's kinematics and initialisation (src/particle.cpp:121-444,
638-710) against closed forms, invariants and the reference's own tests."""
import math

import numpy as np
import pytest

from spintrack import configs
from conftest import to_oracle_opt

G = 9.81


def base_opt(oracle, **kw):
    o = oracle.default_options()
    o["L"] = (0.07, 0.1, 0.4)
    o["gas_coll"] = False
    o["diffuse"] = 0.0
    o["maxPosStep"] = 10.0
    for k, v in kw.items():
        o[k] = v
    return o


def test_hit_time_no_gravity(oracle):
    o = base_opt(oracle, gravity=False)
    s = oracle.rng_seeded(0, 0)
    c = oracle.calc_next_collision_time(o, 0.0, 5.0, (1.0, 0.5, 0.25), (0.0, 0.0, 0.0), 1e300, s)
    # x wall at 0.035 after 0.035 s; y at 0.05/0.5 = 0.1 s; z at 0.2/0.25 = 0.8 s
    assert c["wall_hit"] == "x" and c["coll_type"] == "W"
    assert c["dt"] == 0.035 / 1.0 and c["n_bounce"] == 1


def test_hit_time_quadratic_under_gravity(oracle):
    """Falling from y0 with vy<0: y0 + vy t - g t^2/2 = -Ly/2."""
    o = base_opt(oracle, gravity=True)
    s = oracle.rng_seeded(0, 0)
    y0, vy = 0.01, -0.3
    c = oracle.calc_next_collision_time(o, 0.0, 5.0, (1e-9, vy, 1e-9), (0.0, y0, 0.0), 1e300, s)
    d = y0 + 0.05
    t_exact = (vy + math.sqrt(vy * vy + 2 * G * d)) / G
    assert c["wall_hit"] == "y"
    assert abs(c["dt"] - t_exact) < 1e-15
    # rising but not reaching the ceiling: up, then down to the floor
    vy = 0.5
    c = oracle.calc_next_collision_time(o, 0.0, 5.0, (1e-9, vy, 1e-9), (0.0, y0, 0.0), 1e300, s)
    t_exact = (vy + math.sqrt(vy * vy + 2 * G * d)) / G
    assert abs(c["dt"] - t_exact) < 1e-14
    # rising fast enough to hit the ceiling at +Ly/2
    vy = 2.0
    c = oracle.calc_next_collision_time(o, 0.0, 5.0, (1e-9, vy, 1e-9), (0.0, y0, 0.0), 1e300, s)
    h = 0.05 - y0
    t_exact = (vy - math.sqrt(vy * vy - 2 * G * h)) / G
    assert abs(c["dt"] - t_exact) < 1e-15


def test_gas_vs_wall_and_interval_cap(oracle):
    o = base_opt(oracle, gravity=False, maxPosStep=0.1, tc=1e-3)
    s = oracle.rng_seeded(0, 0)
    # gas collision before the wall
    c = oracle.calc_next_collision_time(o, 0.0, 5.0, (1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.01, s)
    assert c["coll_type"] == "G" and c["dt"] == 0.01 and c["n_coll"] == 1 and c["next_gas"] > 0.01
    # nothing before the interval end: 'N' and finished
    c = oracle.calc_next_collision_time(o, 0.0, 0.02, (1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1e300, s)
    assert c["coll_type"] == "N" and c["finished"] and c["dt"] == 0.02
    # capped by maxPosStep, not finished
    o2 = base_opt(oracle, gravity=False, maxPosStep=0.01)
    c = oracle.calc_next_collision_time(o2, 0.0, 1.0, (0.1, 0.0, 0.0), (0.0, 0.0, 0.0), 1e300, s)
    assert c["coll_type"] == "N" and not c["finished"] and c["dt"] == 0.01


def test_corner_and_edge_ties(oracle):
    o = base_opt(oracle, gravity=False, L=(0.1, 0.1, 0.1))
    s = oracle.rng_seeded(0, 0)
    c = oracle.calc_next_collision_time(o, 0.0, 5.0, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), 1e300, s)
    assert c["wall_hit"] == "a"
    c = oracle.calc_next_collision_time(o, 0.0, 5.0, (1.0, 1.0, 0.5), (0.0, 0.0, 0.0), 1e300, s)
    assert c["wall_hit"] == "b"
    u = oracle.update_position_and_velocity(o, 0.0, c["dt"], (0.0, 0.0, 0.0), (1.0, 1.0, 0.5), "W", "b", s)
    assert tuple(u["v"]) == (-1.0, -1.0, 0.5)
    assert u["pos"][0] == 0.05 and u["pos"][1] == 0.05


def test_specular_energy_invariant(oracle):
    """1/2 |v|^2 + g (y + Ly/2) is conserved by flights and specular bounces
    (scripts/PhysicsSanityChecks/EnergyConservationCheck.ipynb)."""
    o = base_opt(oracle, gravity=True)
    s = oracle.rng_seeded(0, 0)
    pos = np.array([0.01, 0.02, -0.1])
    v = np.array([1.3, 2.1, -0.7])
    t = 0.0
    e0 = 0.5 * v @ v + G * (pos[1] + 0.05)
    for _ in range(3000):
        c = oracle.calc_next_collision_time(o, t, 100.0, v, pos, 1e300, s)
        u = oracle.update_position_and_velocity(o, t, c["dt"], pos, v, c["coll_type"], c["wall_hit"], s)
        t, pos, v = u["t"], u["pos"], u["v"]
        assert np.all(np.abs(pos) <= np.array([0.035, 0.05, 0.2]) + 1e-15)
    e1 = 0.5 * v @ v + G * (pos[1] + 0.05)
    assert abs(e1 / e0 - 1) < 1e-10


def test_diffuse_preserves_speed_and_is_lambertian(oracle):
    o = base_opt(oracle, gravity=False, diffuse=1.0)
    s = oracle.rng_seeded(5, 5)
    cosang = []
    for _ in range(20000):
        u = oracle.update_position_and_velocity(o, 0.0, 0.0, (0.035, 0.0, 0.0), (2.0, 1.0, -1.0), "W", "x", s)
        v = u["v"]
        assert abs(np.linalg.norm(v) - math.sqrt(6.0)) < 1e-14
        assert v[0] < 0          # re-emitted into the cell
        cosang.append(-v[0] / np.linalg.norm(v))
    cosang = np.array(cosang)
    # Lambert: P(cos <= c) = c^2  =>  E[cos] = 2/3, E[cos^2] = 1/2
    assert abs(cosang.mean() - 2 / 3) < 0.01 and abs((cosang ** 2).mean() - 0.5) < 0.01


def test_sanity_check_codes(oracle):
    o = base_opt(oracle)
    ok = ((0, 0, 0), (0, 0, 0), (1, 1, 1), (1, 1, 1))
    assert oracle.sanity_check(o, 0.0, 1.0, *ok) == (False, 0)
    assert oracle.sanity_check(o, 0.0, 1.0, (0, 0, 0), (0.04, 0, 0), (1, 1, 1), (1, 1, 1)) == (True, 1)
    assert oracle.sanity_check(o, 0.0, math.inf, *ok) == (True, 2)
    assert oracle.sanity_check(o, 0.0, 1.0, (math.nan, 0, 0), (0, 0, 0), (1, 1, 1), (1, 1, 1)) == (True, 3)
    assert oracle.sanity_check(o, 0.0, 1.0, (0, 0, 0), (0, 0, 0), (math.nan, 1, 1), (1, 1, 1)) == (True, 4)


def test_interpolation_reference_case(oracle):
    """tests/particle_validation.cpp:44-80 of the reference."""
    o = base_opt(oracle, gravity=False)
    p, v = oracle.interpolate(o, 1.0, 0.0, (1, 2, 3), (1, 2, 3))
    assert np.allclose(p, (2, 4, 6), atol=1e-9) and np.allclose(v, (1, 2, 3), atol=1e-9)
    o["gravity"] = True
    p, v = oracle.interpolate(o, 1.0, 0.0, (1, 2, 3), (1, 2, 3))
    g = np.array([0, -9.81, 0])
    assert np.allclose(p, np.array([1, 2, 3]) + np.array([1, 2, 3]) + 0.5 * g, atol=1e-9)
    assert np.allclose(v, np.array([1, 2, 3]) + g, atol=1e-9)


def test_init_maxwell_boltzmann_reference_case(oracle):
    """tests/particle_validation.cpp:14-42: 100000 particles, 'M', T=0.4,
    m=1e-26; mean velocity < 2e-2 sigma, rms speed within 2%."""
    k = 1.380649e-23
    o = oracle.default_options()
    o["gravity"] = False
    o["T"] = 0.4
    o["m"] = 1e-26
    o["dist"] = b"M"
    o["numParticles"] = 100000
    o["sqrtKT_m"] = math.sqrt(k * 0.4 / 1e-26)
    E = oracle.Ensemble(o, 100000)
    v = oracle.xyz(E.arr["v"])
    sig = math.sqrt(k * 0.4 / 1e-26)
    assert np.all(np.abs(v.mean(0) / sig) < 2e-2)
    assert abs(math.sqrt((v ** 2).sum(1).mean()) / math.sqrt(3) / sig - 1) < 2e-2
    pos = oracle.xyz(E.arr["pos"])
    assert np.all(np.abs(pos) <= np.array([0.035, 0.05, 0.2]))


def test_init_isotropic_fixed_speed(oracle):
    opt = configs.neutron_cell(n=20000)
    E = oracle.Ensemble(to_oracle_opt(opt), 20000)
    v = oracle.xyz(E.arr["v"])
    assert np.allclose(np.linalg.norm(v, axis=1), 5.0, rtol=1e-14)
    assert np.all(np.abs(v.mean(0)) < 0.1)
    assert np.allclose((v ** 2).mean(0), 25.0 / 3, rtol=0.05)


def test_he3_run_invariants(oracle):
    """A short gas + diffuse + gravity run: no failures, |S| = 1 to 1e-12,
    and the gas collision rate matches 1/tc."""
    opt = configs.he3_thermal(n=400, tf=1.0, ioutInt=0.5)
    E = oracle.Ensemble(to_oracle_opt(opt), 400)
    E.advance_to(0.5, prev_t=0.0)
    E.advance_to(1.0, prev_t=0.5)
    s = E.snapshot()
    assert np.all(s["failure"] == 0)
    assert np.all(np.abs(np.linalg.norm(oracle.xyz(s["S"]), axis=1) - 1) < 1e-12)
    assert abs(s["n_coll"].mean() / (1.0 / 2e-3) - 1) < 0.03
    assert np.abs(s["t"] - 1.0).max() < 1e-15
