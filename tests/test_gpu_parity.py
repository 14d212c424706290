"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bar (the tolerances are the ones stated for this path):
  * integer state bit-exact: bounce, gas-collision and event counts, failure
    codes, the RNG words, integrated-segment counts (identical event
    sequences);
  * spin and position within 1e-9 absolute for runs without gas collisions
    or diffuse walls (the exact variant agrees to ~1e-14 there);
  * with gas collisions and diffuse walls, libm last-bit differences (ocml on
    the GPU, glibc on the CPU: log in the exponential/normal draws, sin/cos/
    acos in diffuse re-emission) are amplified by grazing wall hits under
    gravity; there the spread must stay within that of the reference
    algorithm against itself with a 1-ulp perturbation (ulp_control), and
    every traced segment, replayed through the oracle's integrator from the
    GPU's own inputs, must agree to 1e-13 (chaos-free local parity).
"""
import math

import numpy as np
import pytest

import spintrack as st
from spintrack import configs
from conftest import to_oracle_opt

pytestmark = pytest.mark.gpu

TOL = 1e-9
PIPE, FUSED = st._lib.ENGINE_PIPELINE, st._lib.ENGINE_FUSED


def run_gpu(opt, targets, variant=st.VARIANT_EXACT, persistent=False, pulse_axis=2, stop_at=None,
            engine=PIPE, seg_buffer_mb=0):
    p = st.Particle(opt, pulse_axis=pulse_axis, variant=variant, engine=engine, seg_buffer_mb=seg_buffer_mb)
    p.initParticles()
    if stop_at is not None:
        p.set_stop_times(stop_at)
    if persistent:
        p.run(targets)
    else:
        for t in targets:
            p.runSimulation(t)
    snap = p.snapshot()
    state = p.get_state()
    cnt = p.particle_counters()
    p.close()
    return snap, state, cnt


def run_oracle(O, opt, targets, pulse_axis=2, stop_at=None):
    E = O.Ensemble(to_oracle_opt(opt), opt.numParticles, pulse_axis=pulse_axis)
    prev = opt.t0
    for t in targets:
        E.advance_to(t, stop_at=stop_at, prev_t=prev)
        prev = t
    return E


def check_integer_state(O, g, gstate, gcnt, E):
    r = E.snapshot()
    for k in ("n_bounce", "n_coll", "n_steps", "failure"):
        bad = np.nonzero(g[k] != r[k])[0]
        assert bad.size == 0, f"{k} differs for {bad.size} particles, first {bad[:5]}"
    for k in ("x", "y", "z", "w"):
        assert np.array_equal(gstate[f"rng_{k}"], E.arr["state"][k]), f"rng word {k} differs"
    assert np.array_equal(gstate["hasSpare"].astype(bool), E.arr["state"]["hasSpare"])
    assert np.array_equal(gcnt["segments"], E.n_ode.reshape(-1, 3)[:, 2]), "segment counts differ"
    return r


def compare(O, g, gstate, gcnt, E, tol=TOL, step_slack=0):
    r = check_integer_state(O, g, gstate, gcnt, E)
    n_ode = E.n_ode.reshape(-1, 3)
    out = {}
    # accepted/rejected trial counts follow the step-size controller, which
    # sees last-bit libm differences (pow in the controller, log in the
    # initial normal draws); reported, and bounded loosely (step_slack more
    # steps for integrators whose error estimate sits at round-off level)
    da = np.abs(gcnt["accepted"].astype(np.int64) - n_ode[:, 0].astype(np.int64))
    out["steps_identical"] = float(np.mean(da == 0))
    assert ((da - step_slack) / np.maximum(1, n_ode[:, 0])).max() < 0.05
    for k in ("S", "pos", "v"):
        a, b = st.xyz(g[k]), O.xyz(r[k])
        assert np.array_equal(np.isfinite(a), np.isfinite(b)), f"{k} NaN pattern"
        d = np.abs(a - b)
        fin = np.isfinite(d)
        mx = d[fin].max() if fin.any() else 0.0
        assert mx < tol, f"max |d{k}| = {mx:.3e}"
        out[k] = mx
    assert np.nanmax(np.abs(g["t"] - r["t"])) < 1e-12 if np.isfinite(g["t"]).any() else True
    return out


def ulp_control(O, opt, targets, pulse_axis=2):
    """The reference algorithm against itself with each particle's first gas
    time moved by one ulp; returns per-particle max |dS| and max |dpos|."""
    runs = []
    for perturb in (False, True):
        E = O.Ensemble(to_oracle_opt(opt), opt.numParticles, pulse_axis=pulse_axis)
        if perturb:
            E.arr["next_gas"][:] = np.nextafter(E.arr["next_gas"], np.inf)
        prev = opt.t0
        for t in targets:
            E.advance_to(t, prev_t=prev)
            prev = t
        runs.append(E.snapshot())
    a, b = runs
    return (np.abs(O.xyz(a["S"]) - O.xyz(b["S"])).max(1), np.abs(O.xyz(a["pos"]) - O.xyz(b["pos"])).max(1))


def compare_chaotic(O, opt, targets, g, gs, gc, E, pulse_axis=2):
    r = check_integer_state(O, g, gs, gc, E)
    dS = np.abs(st.xyz(g["S"]) - O.xyz(r["S"])).max(1)
    dp = np.abs(st.xyz(g["pos"]) - O.xyz(r["pos"])).max(1)
    cS, cp = ulp_control(O, opt, targets, pulse_axis)
    out = dict(frac_dS_gt_1e9=float(np.mean(dS > TOL)), ctrl_frac_dS=float(np.mean(cS > TOL)),
               frac_dpos_gt_1e9=float(np.mean(dp > TOL)), ctrl_frac_dpos=float(np.mean(cp > TOL)),
               median_dS=float(np.median(dS)), max_dS=float(dS.max()), ctrl_max_dS=float(cS.max()))
    assert out["frac_dS_gt_1e9"] <= 2 * out["ctrl_frac_dS"] + 0.005, out
    assert out["frac_dpos_gt_1e9"] <= 2 * out["ctrl_frac_dpos"] + 0.005, out
    assert out["median_dS"] < 1e-12, out
    return out


# --------------------------------------------------------------------------
@pytest.mark.parametrize("engine", [PIPE, FUSED])
def test_neutron_cell_exact_advance(oracle, engine):
    """1000 neutrons, 1 s, 20 output intervals, exact variant, st_advance_to."""
    opt = configs.neutron_cell()
    targets = st.output_times(opt)
    g, gs, gc = run_gpu(opt, targets, engine=engine)
    E = run_oracle(oracle, opt, targets)
    print(f"neutron_cell exact engine={engine}:", compare(oracle, g, gs, gc, E))


def test_neutron_cell_fast(oracle):
    opt = configs.neutron_cell(n=500)
    targets = st.output_times(opt)
    g, gs, gc = run_gpu(opt, targets, variant=st.VARIANT_FAST, persistent=True)
    E = run_oracle(oracle, opt, targets)
    print("neutron_cell fast:", compare(oracle, g, gs, gc, E))


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
def test_engines_and_paths_agree_bitwise(variant):
    """Pipeline (with a tiny staging buffer, so runs cross many chunk
    boundaries) and fused engines, persistent and per-interval paths: the
    same per-particle state, bit for bit."""
    opt = configs.he3_thermal(n=777, tf=0.5, ioutInt=0.125)
    targets = st.output_times(opt)
    ref = None
    for engine, persistent, mb in ((FUSED, False, 0), (FUSED, True, 0), (PIPE, False, 1), (PIPE, True, 1),
                                   (PIPE, True, 0)):
        g, s, c = run_gpu(opt, targets, variant=variant, persistent=persistent, engine=engine,
                          seg_buffer_mb=mb)
        if ref is None:
            ref = (s, c)
            continue
        if variant == st.VARIANT_EXACT:
            assert s.tobytes() == ref[0].tobytes(), (engine, persistent, mb)
            for k in c:
                assert np.array_equal(c[k], ref[1][k]), (engine, persistent, mb, k)
        else:
            # the fast variant's kernels are compiled per engine and pulse
            # form, and the compiler may contract an FMA differently in each:
            # trajectories and RNG stay bit-exact, spins agree to round-off
            for name in s.dtype.names:
                if name not in ("S", "h"):
                    assert s[name].tobytes() == ref[0][name].tobytes(), (engine, persistent, mb, name)
            dS = np.abs(st.xyz(s["S"]) - st.xyz(ref[0]["S"]))
            assert np.nanmax(dS, initial=0.0) < 1e-11, (engine, persistent, mb, np.nanmax(dS))
            assert np.array_equal(c["segments"], ref[1]["segments"])
            assert np.mean(c["accepted"] == ref[1]["accepted"]) > 0.99


def test_he3_thermal_exact(oracle):
    """He-3, Maxwell-Boltzmann at 0.45 K, gas collisions, 5% diffuse walls."""
    opt = configs.he3_thermal(n=2000, tf=1.0, ioutInt=0.25)
    targets = st.output_times(opt)
    g, gs, gc = run_gpu(opt, targets)
    E = run_oracle(oracle, opt, targets)
    assert g["n_coll"].sum() > 0 and g["n_bounce"].sum() > 0
    print("he3 exact:", compare_chaotic(oracle, opt, targets, g, gs, gc, E))


def test_he3_thermal_fast_persistent(oracle):
    opt = configs.he3_thermal(n=1500, tf=1.0, ioutInt=0.5)
    targets = st.output_times(opt)
    g, gs, gc = run_gpu(opt, targets, variant=st.VARIANT_FAST, persistent=True)
    E = run_oracle(oracle, opt, targets)
    print("he3 fast persistent:", compare_chaotic(oracle, opt, targets, g, gs, gc, E))


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
def test_he3_segment_replay(oracle, variant):
    """Every segment the GPU integrated for a traced particle, replayed
    through the oracle's integrateDOP from the GPU's own segment inputs,
    ends at the GPU's spin within 1e-13 (exact) / 1e-11 (fast)."""
    opt = configs.he3_thermal(n=2000, tf=1.0, ioutInt=0.25)
    o = to_oracle_opt(opt)
    for pid in (3, 1271):
        p = st.Particle(opt, variant=variant)
        p.initParticles()
        p.set_trace(pid, 20000)
        for t in st.output_times(opt):
            p.runSimulation(t)
        tr = p.get_trace()
        p.close()
        starts, ends = tr[tr[:, 0] == 1], tr[tr[:, 0] == 2]
        assert len(starts) == len(ends) > 500
        worst = 0.0
        for s, e in zip(starts, ends):
            stt, y, h, na, nr = oracle.integrate(o, s[1], s[2], s[9:12], s[3:6], s[6:9], s[12])
            assert stt == 1
            worst = max(worst, np.abs(y - e[9:12]).max())
        print(f"variant {variant} particle {pid}: {len(starts)} segments, worst |dS| {worst:.2e}")
        assert worst < (1e-13 if variant == st.VARIANT_EXACT else 1e-11)


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
def test_overflow_pages_match_deep_chunks(variant):
    """Staging budgets that limit the chunk depth, for 1024 neutrons over 2 s
    (~150 records each).  7 MiB: depth 112 and a pool of 112 overflow pages
    of 256 records -- the particles that reach the depth with less than
    twice the chunk's time left to go take pages until the pool runs out, the
    rest continue in later chunks.  12 MiB: depth 192 and 192 pages, enough for
    every particle that qualifies.  Per-particle results are those of one
    deep chunk without a pool, bit for bit, in no more chunks than the same
    depth without a pool."""
    opt = configs.neutron_cell(n=1024, tf=2.0, ioutInt=0.5)
    opt.integratorType = 2 if variant == st.VARIANT_FAST else 0
    targets = st.output_times(opt)

    def run(**kw):
        p = st.Particle(opt, variant=variant, **kw)
        p.initParticles()
        p.kernel_timing(True)
        p.run(targets)
        out = p.snapshot(), p.get_state(), p.particle_counters(), p.kernel_times()
        p.close()
        return out

    g_deep, s_deep, c_deep, k_deep = run(chunk_records=4096)
    assert k_deep["kin_launches"] == 1
    assert np.mean(c_deep["segments"] > 112) > 0.5   # many particles need more than the small depth
    for mb in (7, 12):
        g, s_, c, k = run(seg_buffer_mb=mb)
        assert s_.tobytes() == s_deep.tobytes(), mb
        for name in c_deep:
            assert np.array_equal(c[name], c_deep[name]), (mb, name)
        for name in g_deep:
            assert np.asarray(g[name]).tobytes() == np.asarray(g_deep[name]).tobytes(), (mb, name)
        print(f"{mb} MiB: kinematics launches {k['kin_launches']}", flush=True)
        # the same depth without a pool needs at least as many chunks
        same = run(chunk_records=112 if mb == 7 else 192)[3]["kin_launches"]
        print(f"   same depth without a pool: {same}", flush=True)
        assert k["kin_launches"] <= same


@pytest.mark.parametrize("engine", [PIPE, FUSED])
def test_integrator_failure_nmax(oracle, engine):
    """nmax = 4: segments needing more than 5 trial steps stop the particle
    with failureState -1 (src/integrator.cpp:129-131, src/particle.cpp:596-612).
    The pipeline must rewind its kinematics to the failing event."""
    opt = configs.he3_thermal(n=300, tf=0.3, ioutInt=0.1)
    opt.nmax = 4
    targets = st.output_times(opt)
    g, gs, gc = run_gpu(opt, targets, engine=engine, seg_buffer_mb=1)
    E = run_oracle(oracle, opt, targets)
    r = E.snapshot()
    assert (r["failure"] == -1).sum() > 20 and (r["failure"] == 0).sum() > 0
    check_integer_state(oracle, g, gs, gc, E)
    fin = r["failure"] == 0
    assert np.all(np.isnan(st.xyz(g["S"])[~fin]))
    assert np.all(np.isnan(g["t"][~fin]))


def test_rk45_exact(oracle):
    """integratorType 1 with the stale-k1 quirk, neutrons and He-3."""
    for opt in (configs.neutron_cell(n=256, tf=0.02, ioutInt=0.01),
                configs.he3_thermal(n=256, tf=0.02, ioutInt=0.01, integrator=1)):
        opt.integratorType = 1
        targets = st.output_times(opt)
        g, gs, gc = run_gpu(opt, targets)
        E = run_oracle(oracle, opt, targets)
        res = compare(oracle, g, gs, gc, E) if opt.gas_coll == 0 else compare_chaotic(
            oracle, opt, targets, g, gs, gc, E)
        print("rk45:", res)


def test_pulse_along_y(oracle):
    """Dressing pulse a*cos(w*t) along y (pulse hook on another axis)."""
    # output times off the 3 kHz pulse's zero phase, where the dressed spin
    # is furthest from yi
    opt = configs.dressing(n=300, species="n", tf=0.05, ioutInt=0.01241)
    opt.gas_coll = 0
    opt.diffuse = 0.0
    targets = st.output_times(opt)
    for variant in (st.VARIANT_EXACT, st.VARIANT_FAST):
        g, gs, gc = run_gpu(opt, targets, variant=variant, pulse_axis=1)
        E = run_oracle(oracle, opt, targets, pulse_axis=1)
        print(f"dressing variant {variant}:", compare(oracle, g, gs, gc, E, tol=TOL))
        # the pulse dominates: the spin must have moved far from yi
        assert np.abs(st.xyz(g["S"])[:, 0] - 1).max() > 1e-3


def test_stop_times_and_t0(oracle):
    """Per-particle end times (uneven durations) and a non-zero t0."""
    opt = configs.neutron_cell(n=400, tf=1.0, ioutInt=0.25)
    opt.t0 = 0.1
    rng = np.random.default_rng(2)
    stop_at = rng.uniform(0.1, 1.0, 400)
    targets = st.output_times(opt)
    g, gs, gc = run_gpu(opt, targets, stop_at=stop_at, persistent=True)
    E = run_oracle(oracle, opt, targets, stop_at=stop_at)
    print("stop times:", compare(oracle, g, gs, gc, E))
    assert np.allclose(g["t"], np.minimum(stop_at, targets[-1]), atol=1e-12)


def test_snapshot_blocks_match_snapshots():
    """st_run's on-GPU snapshot blocks == st_snapshot after each st_advance_to."""
    opt = configs.he3_thermal(n=301, tf=0.4, ioutInt=0.1)
    targets = st.output_times(opt)
    with st.Particle(opt, variant=st.VARIANT_FAST) as p:
        p.initParticles()
        bb = p.block_bytes()
        blocks = st._lib.DeviceBuffer((len(targets) + 1) * bb)
        p.snapshot_block_device(blocks.ptr)
        p.run(targets, blocks.ptr + bb)
        host = p.copy_to_host(blocks.ptr, blocks.nbytes)
        blocks.free()
    dt = st.snapshot_block_dtype(301)
    with st.Particle(opt, variant=st.VARIANT_FAST) as q:
        q.initParticles()
        for k in range(len(targets) + 1):
            if k:
                q.runSimulation(targets[k - 1])
            s = q.snapshot()
            b = np.frombuffer(host[k * bb:(k + 1) * bb].tobytes(), dtype=dt)[0]
            for f in ("t", "pos", "v", "S", "failure", "n_coll", "n_bounce", "n_steps"):
                assert b[f].tobytes() == s[f].tobytes(), (k, f)


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("kind", ["neutrons_specular", "he3_gas_diffuse"])
def test_trajectory_only_matches_oracle(oracle, kind, persistent):
    """integratorType -1 (trajectory only, src/particle.cpp's loop with no
    spin integration): the kinematics kernel keeps no segment records and
    runs each call in one chunk.  Integer state and RNG words bit-exact; the
    spin stays the initial one; positions within 1e-9 without gas and diffuse
    walls, and within the reference's own 1-ulp spread with them."""
    if kind == "neutrons_specular":
        opt = configs.neutron_cell(n=2048, tf=1.0, ioutInt=0.25, seed=3)
    else:
        opt = configs.he3_thermal(n=2048, tf=0.5, ioutInt=0.125, seed=5)
    opt.integratorType = -1
    targets = st.output_times(opt)
    g, gs, gc = run_gpu(opt, targets, persistent=persistent)
    E = run_oracle(oracle, opt, targets)
    r = E.snapshot()
    for k in ("n_bounce", "n_coll", "n_steps", "failure"):
        bad = np.nonzero(g[k] != r[k])[0]
        assert bad.size == 0, f"{k} differs for {bad.size} particles, first {bad[:5]}"
    for k in ("x", "y", "z", "w"):
        assert np.array_equal(gs[f"rng_{k}"], E.arr["state"][k]), f"rng word {k} differs"
    assert np.all(gc["segments"] == 0) and np.all(gc["accepted"] == 0)
    gS, oS = st.xyz(g["S"]), oracle.xyz(r["S"])
    assert np.array_equal(gS, oS, equal_nan=True), "spin differs"
    live = np.isfinite(gS).all(1)
    assert live.mean() > 0.99 and np.all(gS[live] == np.array([opt.yi.x, opt.yi.y, opt.yi.z])), "spin moved"
    assert np.nanmax(np.abs(g["t"] - r["t"])) < 1e-12
    dp = np.abs(st.xyz(g["pos"]) - oracle.xyz(r["pos"])).max(1)
    if kind == "neutrons_specular":
        assert dp.max() < TOL, dp.max()
    else:
        _, cp = ulp_control(oracle, opt, targets)
        assert np.mean(dp > TOL) <= 2 * np.mean(cp > TOL) + 0.005, (np.mean(dp > TOL), np.mean(cp > TOL))
