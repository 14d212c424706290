"""integratorType 6 (the spectral / Floquet path) on the GPU against the
oracle: the per-particle Goertzel aggregators the sampling kernel leaves
after each interval (SpectrumAggregator, src/floquet.cpp:111-147, fed by
integrateSpectrum, src/integrator.cpp:730-743), the run's summed covariance
spectrum in both sum modes (src/particle.cpp:824-834 and :893-897), the
Floquet decomposition and the final Bloch vector (floquet_integrate,
src/integrator.cpp:869-889).

Specular neutrons have deterministic trajectories, so without test noise
every sample is computed from the same operands in the same order as the
oracle: aggregators, sample counts and spectra are compared to 1e-12
relative (they agree bit for bit in practice).  With the test noise on, the
device sine differs from glibc's in the last bits (1e-9).  A size-free
invariant holds at any size: a particle that never stops samples exactly
the grid points k*h inside (t0, t_end].
"""
import math

import numpy as np
import pytest

import spintrack as st
from spintrack import configs
from conftest import to_oracle_opt

pytestmark = pytest.mark.gpu


def spectral_neutrons(n=256, tf=0.2, h=1e-4, noise=False, seed=0):
    opt = configs.neutron_cell(n=n, tf=tf, ioutInt=tf / 2, seed=seed)
    opt.integratorType = 6
    opt.a = 3e-5
    opt.w = 2 * math.pi * 1000.0
    opt.h = h
    opt.Gy = st.Coords(2e-9, 0.0, 5e-9)
    if noise:
        opt.noiseAmplitudes = st.Coords(1e-9, 2e-9, 0.5e-9)
        opt.noiseFrequencies = st.Coords(2 * math.pi * 300.0, 2 * math.pi * 700.0, -2 * math.pi * 450.0)
    return opt


def run_gpu(opt, mode, persistent=False, variant=st.VARIANT_EXACT):
    targets = st.output_times(opt)
    with st.Particle(opt, spectrum_sum=mode, variant=variant) as p:
        p.initParticles()
        if persistent:
            p.run(targets)
        else:
            for t in targets:
                p.runSimulation(t)
        return dict(snap=p.snapshot(), agg=p.spectrumAggregators(), spec=p.spectrum(), b_end=p.floquetResults(),
                    fd=p.floquetInfo(), counters=p.particle_counters())


def run_oracle(oracle, opt, mode):
    E = oracle.Ensemble(to_oracle_opt(opt), opt.numParticles, spectrum_sum=mode)
    prev = opt.t0
    for t in st.output_times(opt):
        E.advance_to(t, prev_t=prev)
        prev = t
    return E


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def check_against_oracle(oracle, opt, mode, g, E, tol):
    # the Floquet decomposition is host work in both: identical
    assert np.array_equal(g["fd"]["frequencies"], E.fd["frequencies"])
    assert np.array_equal(g["fd"]["propagators"], E.fd["propagators"])
    assert g["fd"]["dt"] == E.fd["dt"]
    # trajectories (specular: deterministic)
    r = E.snapshot()
    for k in ("n_bounce", "n_coll", "n_steps", "failure"):
        assert np.array_equal(g["snap"][k], r[k]), k
    assert np.abs(st.xyz(g["snap"]["pos"]) - oracle.xyz(r["pos"])).max() < 1e-12
    # S is not integrated: it stays yi
    assert np.array_equal(st.xyz(g["snap"]["S"]), oracle.xyz(r["S"]))
    # per-particle aggregators of the last interval
    a = g["agg"]
    assert np.array_equal(a["n_samples"].astype(np.int64), E.agg["n_samples"].astype(np.int64))
    assert a["n_samples"].min() > 0
    for k in ("s1", "s2", "im"):
        assert rel(a[k], E.agg[k]) < tol, (k, rel(a[k], E.agg[k]))
    # the run's summed spectrum
    sp = g["spec"]
    assert sp["n_samples"] == E.cspec_n
    assert rel(sp["raw"], E.cspec) < (tol if mode == 0 else max(tol, 1e-11)), rel(sp["raw"], E.cspec)
    assert np.array_equal(sp["frequencies"], E.fd["frequencies"]) and sp["dt"] == E.fd["dt"]
    # the final Bloch vector
    b_ref = E.floquet_result()
    assert np.all(np.isfinite(g["b_end"]))
    assert np.abs(g["b_end"] - b_ref).max() < 1e-9, (g["b_end"], b_ref)
    # every interval's samples are counted in the per-particle "accepted" counter
    assert np.array_equal(g["counters"]["accepted"], E.samples_total)


# the fast variant (noise polynomial, FMA filters) agrees to rounding: its
# filter states carry relative differences of ~1e-13 after 1000 samples
VARIANT_TOL = {st.VARIANT_EXACT: 1e-12, st.VARIANT_FAST: 1e-9}


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
@pytest.mark.parametrize("mode", [0, 1])
def test_neutrons_match_oracle(oracle, mode, variant):
    opt = spectral_neutrons()
    print(f"[spectral] neutrons n={opt.numParticles} mode={mode} variant={variant}", flush=True)
    g = run_gpu(opt, mode, variant=variant)
    E = run_oracle(oracle, opt, mode)
    check_against_oracle(oracle, opt, mode, g, E, tol=VARIANT_TOL[variant])
    exact = all(np.array_equal(g["agg"][k], E.agg[k]) for k in ("s1", "s2", "im"))
    print(f"  aggregators bit-identical: {exact}; b_end {g['b_end']}", flush=True)


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
def test_neutrons_with_test_noise(oracle, variant):
    opt = spectral_neutrons(noise=True)
    print(f"[spectral] neutrons with test noise variant={variant}", flush=True)
    g = run_gpu(opt, 1, variant=variant)
    E = run_oracle(oracle, opt, 1)
    check_against_oracle(oracle, opt, 1, g, E, tol=1e-9)


def test_run_equals_advance():
    opt = spectral_neutrons(n=300, tf=0.3)
    opt.ioutInt = 0.1
    print("[spectral] st_run vs st_advance_to", flush=True)
    a = run_gpu(opt, 1, persistent=False)
    b = run_gpu(opt, 1, persistent=True)
    for k in ("s1", "s2", "im", "n_samples"):
        assert np.array_equal(a["agg"][k], b["agg"][k]), k
    assert np.array_equal(a["spec"]["raw"], b["spec"]["raw"])
    assert np.array_equal(a["b_end"], b["b_end"])
    # particle::spinMean: S is not integrated, so the mean stays yi
    with st.Particle(opt) as p:
        p.initParticles()
        p.runSimulation(0.1)
        assert np.allclose(p.spinMean(), [opt.yi.x, opt.yi.y, opt.yi.z], rtol=0, atol=1e-15)


def test_sample_grid_invariant():
    """Each particle samples exactly the multiples of h in (t_prev, t_end]
    of the last interval, however its segments fall."""
    opt = spectral_neutrons(n=4096, tf=0.2, h=3e-5)
    print(f"[spectral] grid invariant n={opt.numParticles}", flush=True)
    g = run_gpu(opt, 1)
    t_prev, t_end = st.output_times(opt)[-2:]
    t = g["snap"]["t"]
    ok = np.isfinite(t)
    assert ok.mean() > 0.99
    want = np.floor(t[ok] / opt.h) - math.floor(t_prev / opt.h)
    got = g["agg"]["n_samples"][ok].astype(np.float64)
    # t_end itself may sit a rounding step off the grid
    assert np.abs(got - want).max() <= 1
    assert np.mean(got == want) > 0.999
    # mode 1 sums every particle's samples of every interval
    assert g["spec"]["n_samples"] == int(g["counters"]["accepted"].sum())


def test_he3_gas_and_diffuse(oracle):
    """Thermal He-3 with gas collisions and diffuse walls: integer
    trajectory state exact; the aggregators of particles whose trajectory
    floats agree (all but the libm-sensitive few) match to 1e-12."""
    opt = configs.he3_thermal(n=256, tf=0.1, ioutInt=0.05, integrator=6)
    opt.a = 3e-5
    opt.w = 2 * math.pi * 1000.0
    opt.h = 1e-4
    print("[spectral] he3 gas+diffuse", flush=True)
    g = run_gpu(opt, 1)
    E = run_oracle(oracle, opt, 1)
    r = E.snapshot()
    for k in ("n_bounce", "n_coll", "n_steps", "failure"):
        assert np.array_equal(g["snap"][k], r[k]), k
    dp = np.abs(st.xyz(g["snap"]["pos"]) - oracle.xyz(r["pos"])).max(1)
    same = (dp < 1e-12) & (g["agg"]["n_samples"] == E.agg["n_samples"])
    assert same.mean() > 0.95, same.mean()
    for k in ("s1", "s2", "im"):
        d = np.abs(g["agg"][k][same] - E.agg[k][same]).max()
        assert d <= 1e-12 * np.abs(E.agg[k][same]).max(), (k, d)
    assert abs(g["spec"]["n_samples"] - E.cspec_n) <= 0.01 * E.cspec_n


def test_partial_tile_and_stop_times(oracle):
    """200 particles (the last 64-particle tile of the filter state is
    partial) with uneven per-particle end times: particles past their end
    time sample nothing in later intervals, and their filters are cleared
    like every other particle's (src/particle.cpp:746-747)."""
    opt = spectral_neutrons(n=200, tf=0.3)
    opt.ioutInt = 0.1
    targets = st.output_times(opt)
    stop_at = np.random.default_rng(7).uniform(0.05, 0.35, opt.numParticles)
    print("[spectral] n=200 with stop times", flush=True)
    for mode in (0, 1):
        with st.Particle(opt, spectrum_sum=mode) as p:
            p.initParticles()
            p.set_stop_times(stop_at)
            for t in targets:
                p.runSimulation(t)
            agg, sp, snap = p.spectrumAggregators(), p.spectrum(), p.snapshot()
        E = oracle.Ensemble(to_oracle_opt(opt), opt.numParticles, spectrum_sum=mode)
        prev = opt.t0
        for t in targets:
            E.advance_to(t, stop_at=stop_at, prev_t=prev)
            prev = t
        r = E.snapshot()
        assert np.array_equal(snap["n_steps"], r["n_steps"])
        assert np.array_equal(agg["n_samples"].astype(np.int64), E.agg["n_samples"].astype(np.int64))
        assert np.all(agg["n_samples"][stop_at <= targets[-2]] == 0)
        for k in ("s1", "s2", "im"):
            assert rel(agg[k], E.agg[k]) < 1e-12, k
        assert sp["n_samples"] == E.cspec_n
        assert rel(sp["raw"], E.cspec) < 1e-11


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
def test_no_gravity_pulse_axis_and_step_switches(oracle, variant):
    """Gravity off (the fast variant's noise polynomial loses its quadratic
    term), the drive along y (it enters the Floquet decomposition only) and
    keepStepSize / fixedStepSize set (no-ops here: the sampling step never
    changes)."""
    opt = spectral_neutrons(n=128)
    opt.gravity = 0
    opt.keepStepSize = 1
    opt.fixedStepSize = 1
    print(f"[spectral] no gravity, pulse on y, step switches, variant={variant}", flush=True)
    targets = st.output_times(opt)
    with st.Particle(opt, spectrum_sum=1, variant=variant, pulse_axis=1) as p:
        p.initParticles()
        for t in targets:
            p.runSimulation(t)
        g = dict(snap=p.snapshot(), agg=p.spectrumAggregators(), spec=p.spectrum(), b_end=p.floquetResults(),
                 fd=p.floquetInfo(), counters=p.particle_counters())
    E = oracle.Ensemble(to_oracle_opt(opt), opt.numParticles, spectrum_sum=1, pulse_axis=1)
    prev = opt.t0
    for t in targets:
        E.advance_to(t, prev_t=prev)
        prev = t
    check_against_oracle(oracle, opt, 1, g, E, tol=VARIANT_TOL[variant])

