"""Named workloads used by the tests and the benchmark.

Axis conventions follow the reference, which is what parity is defined
against: gravity acts along -y (include/options.h:7, src/particle.cpp:271-280)
and the user pulse hook is (0, 0, a*cos(w*t)) unless ``pulse_axis`` is
changed (src/integrator.cpp:40-42).  The gas free time tc normally follows
the reference formula tc = 1.6e-4*m/(k*T^8) (src/simulation.cpp:10); a
workload that needs a given mean free time sets ``tc`` explicitly.
"""
from __future__ import annotations

import math

from ._lib import Options
from .options import K_BOLTZMANN, default_options, set_fields


def _common_fields(opt: Options) -> Options:
    # cell, fields and integrator settings shared by the neutron / He-3 runs
    return set_fields(opt, L=(0.07, 0.1, 0.4), B0=(3e-6, 0.0, 0.0), E=(0.0, 0.0, 7.5e6),
                      Gx=(0.0, 0.0, 1e-8), Gy=(0.0, 0.0, 0.0), Gz=(0.0, 0.0, 0.0),
                      yi=(1.0, 0.0, 0.0), integratorType=0, rtol=1e-12, atol=1e-12, h=1e-3,
                      hmax=1.0, hmin=1e-8, t0=0.0, a=0.0, w=0.0, gravity=True)


def neutron_cell(n: int = 1000, tf: float = 1.0, ioutInt: float = 0.05, seed: int = 0) -> Options:
    """1000 neutrons at 5 m/s (isotropic), specular walls, no gas, 1 s,
    snapshot every 0.05 s, DOP853 at rtol = atol = 1e-12."""
    o = _common_fields(default_options())
    set_fields(o, numParticles=n, m=1.675e-27, gamma=-1.832e8, dist="C", V=5.0, diffuse=0.0,
               gas_coll=False, tf=tf, ioutInt=ioutInt, seed=seed, T=0.45)
    o.tc = 1.6e-4 * o.m / (K_BOLTZMANN * o.T ** 8)
    o.sqrtKT_m = math.sqrt(K_BOLTZMANN * o.T / o.m)
    return o


def he3_thermal(n: int = 1048576, tf: float = 10.0, ioutInt: float = 0.5, seed: int = 1,
                integrator: int = 0, tc: float = 2e-3) -> Options:
    """He-3 atoms with Maxwell-Boltzmann velocities at 0.45 K, gas
    collisions with mean free time tc (2e-3 s) and thermal re-emission, 5%
    diffuse walls, gravity, same cell and fields; 10 s, output every 0.5 s."""
    o = _common_fields(default_options())
    set_fields(o, numParticles=n, m=5.008e-27, gamma=-2.038e8, dist="M", T=0.45, diffuse=0.05,
               gas_coll=True, tf=tf, ioutInt=ioutInt, seed=seed, integratorType=integrator)
    o.sqrtKT_m = math.sqrt(K_BOLTZMANN * o.T / o.m)
    o.tc = tc
    return o


def dressing(n: int = 262144, species: str = "n", tf: float = 2.0, ioutInt: float = 0.1,
             seed: int = 3, tc: float = 2e-3) -> Options:
    """Spin dressing: B0 = (1e-6,0,0) T plus a*cos(w*t) with a = 1e-5 T and
    w = 2*pi*3000 rad/s (applied along y via pulse_axis=1), Maxwell-Boltzmann
    at 0.45 K, gas on, 2% diffuse walls, gravity; 2 s, output every 0.1 s."""
    o = _common_fields(default_options())
    if species == "n":
        set_fields(o, m=1.675e-27, gamma=-1.832e8)
    else:
        set_fields(o, m=5.008e-27, gamma=-2.038e8)
    set_fields(o, numParticles=n, B0=(1e-6, 0.0, 0.0), a=1e-5, w=2.0 * math.pi * 3000.0,
               dist="M", T=0.45, gas_coll=True, diffuse=0.02, tf=tf, ioutInt=ioutInt, seed=seed)
    o.sqrtKT_m = math.sqrt(K_BOLTZMANN * o.T / o.m)
    o.tc = tc
    return o


def batched_sweep(n_cfg: int = 64, n_per: int = 16384, seed: int = 2, tf: float = 20.0,
                  ioutInt: float = 1.0, g_max: float = 5e-8, dur=(1.0, 20.0),
                  diffuse: float = 0.1) -> tuple[list, "np.ndarray"]:
    """Field-configuration sweep for one batched launch: ``n_cfg``
    configurations of ``n_per`` neutrons (5 m/s isotropic, specular or
    diffuse walls, no gas, gravity, B0 = (3e-6,0,0) T, DOP853 at
    rtol = atol = 1e-12).  Per configuration:

    * the nine gradient components dB_i/dx_j (rows Gx, Gy, Gz) are uniform in
      [-g_max, g_max], drawn config-major from stream (seed, 0);
    * E = (0, 0, +7.5e6) V/m for even, (0, 0, -7.5e6) V/m for odd configurations;
    * diffuse = ``diffuse`` for the second half of the configurations, 0 for the first;
    * particles start with S = (0, 1, 0), transverse to B0, so each
      configuration's precession phase atan2(Sz, Sy) is defined;
    * configuration c uses seed ``seed + c*n_per``, so particle j of it is
      global stream seed + c*n_per + j.

    Returns (configurations, stop_at): particle i (global index) stops at
    dur[0] + (dur[1]-dur[0])*u with u the first draw of stream (seed, 1 + i).
    """
    import numpy as np

    from .rng import Streams, uniforms

    u = uniforms(seed, 0, 9 * n_cfg).reshape(n_cfg, 9)
    G = g_max * (2.0 * u - 1.0)
    cfgs = []
    for c in range(n_cfg):
        o = neutron_cell(n=n_per, tf=tf, ioutInt=ioutInt, seed=seed + c * n_per)
        set_fields(o, Gx=tuple(G[c, 0:3]), Gy=tuple(G[c, 3:6]), Gz=tuple(G[c, 6:9]),
                   E=(0.0, 0.0, 7.5e6 if c % 2 == 0 else -7.5e6),
                   diffuse=diffuse if c >= n_cfg // 2 else 0.0, yi=(0.0, 1.0, 0.0))
        cfgs.append(o)
    n = n_cfg * n_per
    s = Streams(seed, 1 + np.arange(n, dtype=np.uint64))
    stop_at = cfgs[0].t0 + dur[0] + (dur[1] - dur[0]) * s.uniform()
    return cfgs, stop_at


def dressing_cotracked(n: int = 262144, tf: float = 2.0, ioutInt: float = 0.1, seed: int = 3,
                       tc: float = 2e-3) -> list:
    """The spin-dressing workload with neutrons and He-3 co-tracked: two
    configurations of one batch (species = configuration), each with ``n``
    particles; the He-3 population continues the neutron population's seed
    range (seed + n)."""
    return [dressing(n, "n", tf, ioutInt, seed, tc), dressing(n, "he3", tf, ioutInt, seed + n, tc)]


def adjoint_case(n: int = 131072, tf: float = 50.0, ioutInt: float = 0.5, diffuse: float = 0.01,
                 seed: int = 4) -> Options:
    """Differentiation workload: neutrons in the cell (5 m/s isotropic,
    gravity, no gas) with diffuse = 0.01 walls, B0 = (3e-6,0,0) T,
    E = (0,0,7.5e6) V/m, and the nine gradient coefficients dB_i/dx_j
    uniform in [-1e-8, 1e-8] T/m (stream (seed, 0), rows Gx, Gy, Gz); 50 s
    with an output every 0.5 s.  The spins start transverse to B0,
    S = (0, 1, 0), so the precession phase atan2(Sz, Sy) of the ensemble is
    defined."""
    from .rng import uniforms

    o = neutron_cell(n=n, tf=tf, ioutInt=ioutInt, seed=seed)
    g = 1e-8 * (2.0 * uniforms(seed, 0, 9) - 1.0)
    set_fields(o, yi=(0.0, 1.0, 0.0), diffuse=diffuse, Gx=tuple(g[0:3]), Gy=tuple(g[3:6]), Gz=tuple(g[6:9]))
    return o
