"""GPU parity of the event loop's composition and of the reference switches
on the hot path, through the C ABI:

* the reference's own end-to-end test (tests/output_tests.cpp:20-82) and the
  closed-form axis-aligned neutron (wall hits, 'N' caps, n_steps, spin phase),
  checked against their closed forms and against the oracle;
* keepStepSize and fixedStepSize (src/particle.cpp:783-784,
  src/integrator.cpp:197-246, :275-317) for DOP853 and RK45 forward states;
* RK45's h < hmin failure (-1, src/integrator.cpp:282-286) and DOP853 /
  RK45 nmax failures with the kinematics replay to the failing event;
* the spin-dressing workload as stated (neutrons and He-3, gas collisions
  with tc = 2e-3 s, 2% diffuse walls, gravity, 1e-5 T cos(2 pi 3000 t) along
  y, DOP853 at rtol = atol = 1e-12), plain per species and co-tracked in one
  batch, against the oracle.
"""
import numpy as np
import pytest

import spintrack as st
from spintrack import configs
from spintrack.options import default_options, set_fields
from conftest import to_oracle_opt
from event_loop_cases import AxisCase, check_axis_case, check_output_test, output_test_fields
from test_gpu_parity import (PIPE, FUSED, TOL, check_integer_state, compare, compare_chaotic, run_gpu,
                             run_oracle)

pytestmark = pytest.mark.gpu


def opts(**kw):
    return set_fields(default_options(), **kw)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
@pytest.mark.parametrize("engine,persistent", [(PIPE, False), (PIPE, True), (FUSED, False)])
def test_reference_output_test(oracle, variant, engine, persistent):
    """tests/output_tests.cpp:20-82 on the GPU: three particles with gas
    collisions and fully diffuse walls precess freely about B0 (t and S at
    every snapshot within 1e-9 relative of the closed form, S.x exactly 0);
    event counts and RNG words equal the oracle's."""
    opt = opts(**output_test_fields())
    targets = [(i + 1.0) * 0.05 for i in range(4)]
    print(f"[event-loop] output test variant={variant} engine={engine} persistent={persistent}", flush=True)
    p = st.Particle(opt, variant=variant, engine=engine)
    p.initParticles()
    snaps = [p.snapshot()]
    if persistent:
        p.run(targets)
        snaps.append(p.snapshot())
        check_output_test(snaps[:1], gamma=opt.gamma)
        # the persistent run has only the last snapshot on the host here
        s = snaps[-1]
        ph = 3e-6 * opt.gamma * 0.2
        assert np.all(s["S"]["x"] == 0.0)
        assert np.abs(s["S"]["y"] - np.sin(ph)).max() < 1e-9 and np.abs(s["S"]["z"] - np.cos(ph)).max() < 1e-9
        assert np.abs(s["t"] - 0.2).max() < 1e-15
    else:
        for t in targets:
            p.runSimulation(t)
            snaps.append(p.snapshot())
        check_output_test(snaps, gamma=opt.gamma)
    gs, gc = p.get_state(), p.particle_counters()
    p.close()
    E = run_oracle(oracle, opt, targets)
    check_integer_state(oracle, snaps[-1], gs, gc, E)
    assert (snaps[-1]["n_coll"] > 100).all()


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
@pytest.mark.parametrize("engine", [PIPE, FUSED])
@pytest.mark.parametrize("cap,x0", [(0.01, 0.0123), (0.0031, 0.0301)])
def test_axis_aligned_closed_form(oracle, variant, engine, cap, x0):
    """One planted neutron moving along x (st_set_state): per interval
    n_bounce and n_steps exactly, x to 1e-12, v exactly and the spin phase
    to 1e-9 from the closed form; and the whole state against the oracle."""
    case = AxisCase(x0=x0, cap=cap)
    opt = opts(**case.fields(), tf=0.5, ioutInt=0.05, rtol=1e-12, atol=1e-12, h=1e-3, numParticles=1)
    targets = [(i + 1.0) * 0.05 for i in range(10)]
    print(f"[event-loop] axis case cap={cap} variant={variant} engine={engine}", flush=True)
    p = st.Particle(opt, variant=variant, engine=engine)
    p.initParticles()
    s0 = p.get_state()
    pos, v = case.start()
    s0["pos"][0] = pos
    s0["v"][0] = v
    p.set_state(s0)
    snaps = []
    for t in targets:
        p.runSimulation(t)
        snaps.append(p.snapshot())
    gs, gc = p.get_state(), p.particle_counters()
    p.close()
    check_axis_case(case, targets, snaps)
    E = oracle.Ensemble(to_oracle_opt(opt), 1)
    E.arr["pos"][0] = pos
    E.arr["v"][0] = v
    for t in targets:
        E.advance_to(t)
    print("  ", compare(oracle, snaps[-1], gs, gc, E))
    assert float(gs["h"][0]) == float(E.arr["h"][0])


# ---------------------------------------------------------------------------
def neutrons(n=256, tf=0.2, integ=0, yi=(0.0, 0.6, 0.8), **kw):
    opt = configs.neutron_cell(n=n, tf=tf, ioutInt=tf / 4)
    opt.integratorType = integ
    return set_fields(opt, yi=yi, **kw)


def he3(n=256, tf=0.1, integ=0, **kw):
    opt = configs.he3_thermal(n=n, tf=tf, ioutInt=tf / 4, integrator=integ)
    opt.yi.x, opt.yi.y, opt.yi.z = 0.0, 1.0, 0.0
    return set_fields(opt, **kw)


def traced_segments(opt, pid, targets, variant):
    """Per-interval run with one particle traced: (interval index, t_old, t,
    pos_old, v_old, S in, h in, S out, h out, accepted, rejected) per
    integrated segment, and the final state."""
    p = st.Particle(opt, variant=variant)
    p.initParticles()
    p.set_trace(pid, 20000)
    for t in targets:
        p.runSimulation(t)
    tr = p.get_trace()
    gs = p.get_state()
    p.close()
    starts, ends = tr[tr[:, 0] == 1], tr[tr[:, 0] == 2]
    assert len(starts) == len(ends) > 0
    segs = []
    for s, e in zip(starts, ends):
        k = int(np.searchsorted(np.asarray(targets), s[2] - 1e-15))
        segs.append(dict(k=k, t_old=s[1], t=s[2], p0=s[3:6], v0=s[6:9], S_in=s[9:12], h_in=s[12],
                         S_out=e[9:12], h_out=e[12], acc=int(e[13] - s[13]), rej=int(e[14] - s[14])))
    return segs, gs


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
@pytest.mark.parametrize("integ", [0, 1])
@pytest.mark.parametrize("mode", ["keep", "fixed", "keep+fixed"])
def test_step_size_switches(oracle, integ, mode, variant):
    """keepStepSize (h = tempH after every segment, carried into the next
    interval) and fixedStepSize (every trial accepted at h; DOP853 carries
    each segment's truncated last step, so hmin = h restores it), forward
    states of DOP853 and RK45, per-interval and persistent, against the
    oracle.  The step size itself is compared through its composition: a
    traced particle's segments each start from the previous segment's
    truncated last step, across output times only under keepStepSize, and
    every segment replayed through the oracle's integrator from the GPU's
    inputs gives its spin and trial counts.  (h is not compared with the
    oracle's directly: the DOP853 error estimate is a cancellation with
    ~1e-4 relative precision, so one-ulp differences between the device's
    and glibc's pow in the controller move later step sizes by ~1e-5
    relative while the spin stays within 1e-14.)"""
    kw = {}
    if "keep" in mode:
        kw["keepStepSize"] = True
    if "fixed" in mode:
        kw.update(fixedStepSize=True, h=2e-4, hmin=2e-4)
    tf = 0.2 if integ == 0 else 0.02
    opt = neutrons(tf=tf, integ=integ, **kw)
    targets = st.output_times(opt)
    print(f"[event-loop] {mode} integrator={integ} variant={variant}", flush=True)
    E = run_oracle(oracle, opt, targets)
    for persistent in (False, True):
        g, gs, gc = run_gpu(opt, targets, variant=variant, persistent=persistent)
        res = compare(oracle, g, gs, gc, E)
        assert np.all(g["failure"] == 0)
        if "keep" not in mode:
            assert np.all(gs["h"] == opt.h)
        if "fixed" in mode and integ == 0:
            # no rejections with fixed steps
            assert gc["rejected"].sum() == 0
        print(f"  persistent={persistent}: {res}", flush=True)
    o = to_oracle_opt(opt)
    mismatched = total = 0
    for pid in (0, 77):
        segs, gs = traced_segments(opt, pid, targets, variant)
        assert len(segs) > 4 * (len(targets) if integ == 0 else 1)
        for i, sg in enumerate(segs):
            if i == 0 or (sg["k"] != segs[i - 1]["k"] and "keep" not in mode):
                assert sg["h_in"] == opt.h, (i, sg["h_in"])
            else:
                assert sg["h_in"] == segs[i - 1]["h_out"], (i, sg["h_in"], segs[i - 1]["h_out"])
            stt, y, h, na, nr = oracle.integrate(o, sg["t_old"], sg["t"], sg["S_in"], sg["p0"], sg["v0"],
                                                 sg["h_in"], integ)
            assert stt == (1 if integ == 0 else 0)
            tol = 1e-10 if integ == 1 else (1e-13 if variant == st.VARIANT_EXACT else 1e-11)
            assert np.abs(y - sg["S_out"]).max() < tol, (i, np.abs(y - sg["S_out"]).max())
            if "fixed" in mode:
                # fixed steps: the same trials, and the same truncated last step
                assert (na, nr) == (sg["acc"], sg["rej"]), (i, na, nr, sg)
                assert abs(h - sg["h_out"]) <= 1e-12 * abs(h), (i, h, sg["h_out"])
            else:
                mismatched += (na, nr) != (sg["acc"], sg["rej"])
                total += 1
        assert float(gs["h"][pid]) == (segs[-1]["h_out"] if "keep" in mode else opt.h)
    # adaptive: accept/reject decisions follow the controller's last bits,
    # so an occasional segment may take one trial more or less
    assert mismatched <= 0.05 * max(total, 1), (mismatched, total)


def test_keep_step_size_he3_chaotic(oracle):
    """keepStepSize on He-3 with gas collisions and diffuse walls (the
    chaotic case: compared through the 1-ulp control)."""
    opt = he3(keepStepSize=True)
    targets = st.output_times(opt)
    g, gs, gc = run_gpu(opt, targets, persistent=True)
    E = run_oracle(oracle, opt, targets)
    print("[event-loop] keep he3:", compare_chaotic(oracle, opt, targets, g, gs, gc, E), flush=True)


@pytest.mark.parametrize("hmin", [1e-6, 3e-6])
@pytest.mark.parametrize("engine,persistent", [(PIPE, False), (PIPE, True), (FUSED, False)])
def test_rk45_hmin_failure(oracle, hmin, engine, persistent):
    """RK45 with a step floor above what rtol = 1e-12 needs: trials with
    h < hmin return -1 (src/integrator.cpp:282-286), the particle stops with
    NaN state at that segment; the pipeline replays its kinematics to the
    failing event (counters, RNG, pos_old/v_old as the oracle's)."""
    # spins along B0: with transverse spins the first RK45 trial (stale k1
    # = 0) already fails everywhere; here particles fail after 1-10 segments
    opt = neutrons(tf=0.1, integ=1, hmin=hmin, yi=(1.0, 0.0, 0.0))
    targets = st.output_times(opt)
    print(f"[event-loop] rk45 hmin={hmin} engine={engine} persistent={persistent}", flush=True)
    g, gs, gc = run_gpu(opt, targets, engine=engine, persistent=persistent, seg_buffer_mb=1)
    E = run_oracle(oracle, opt, targets)
    r = check_integer_state(oracle, g, gs, gc, E)
    bad = r["failure"] == -1
    assert bad.sum() > 20 and (~bad).sum() > 5
    assert np.all(np.isnan(st.xyz(g["S"])[bad])) and np.all(np.isnan(g["t"][bad]))
    for k in ("pos_old", "v_old"):
        d = np.abs(st.xyz(gs[k]) - oracle.xyz(E.arr[k]))
        assert d.max() < 1e-12, k
    assert np.abs(gs["t_old"] - E.arr["t_old"]).max() < 1e-12
    res = compare(oracle, g, gs, gc, E)
    print("  ", res, flush=True)


# ---------------------------------------------------------------------------
# Spin dressing as stated
# ---------------------------------------------------------------------------
# output times off the 3 kHz pulse's zero phase, where the dressed spin is
# furthest from yi
DRESS_TF, DRESS_IOUT = 0.0817, 0.02041


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
@pytest.mark.parametrize("species", ["n", "he3"])
def test_dressing_as_stated(oracle, species, variant):
    """Spin dressing: B0 = (1e-6, 0, 0) T plus 1e-5 T cos(2 pi 3000 t) along
    y, Maxwell-Boltzmann at 0.45 K, gas collisions (tc = 2e-3 s), 2% diffuse
    walls, gravity, DOP853 at rtol = atol = 1e-12; 512 particles, 0.08 s."""
    opt = configs.dressing(n=512, species=species, tf=DRESS_TF, ioutInt=DRESS_IOUT)
    assert opt.gas_coll and opt.diffuse == 0.02 and opt.gravity and opt.tc == 2e-3 and opt.rtol == 1e-12
    targets = st.output_times(opt)
    print(f"[event-loop] dressing {species} variant={variant}", flush=True)
    g, gs, gc = run_gpu(opt, targets, variant=variant, persistent=True, pulse_axis=1)
    E = run_oracle(oracle, opt, targets, pulse_axis=1)
    assert g["n_coll"].sum() > 1000 and g["n_bounce"].sum() > 1000
    res = compare_chaotic(oracle, opt, targets, g, gs, gc, E, pulse_axis=1)
    # the pulse dominates: spins moved far from yi
    assert np.abs(st.xyz(g["S"])[:, 0] - 1).max() > 1e-3
    print("  ", res, flush=True)


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
def test_dressing_cotracked_batch_vs_oracle(oracle, variant):
    """Neutrons and He-3 co-tracked in one batched handle: each species'
    particles against that species' own oracle run."""
    cfgs = configs.dressing_cotracked(n=384, tf=DRESS_TF, ioutInt=DRESS_IOUT)
    targets = st.output_times(cfgs[0])
    print(f"[event-loop] dressing co-tracked batch variant={variant}", flush=True)
    p = st.Particle.batch(cfgs, pulse_axis=1, variant=variant)
    p.initParticles()
    p.run(targets)
    g, gs, gc = p.snapshot(), p.get_state(), p.particle_counters()
    off = p.layout()
    p.close()
    for c, opt in enumerate(cfgs):
        lo, hi = int(off[c]), int(off[c + 1])
        gk = {k: v[lo:hi] for k, v in g.items()}
        ck = {k: v[lo:hi] for k, v in gc.items()}
        E = run_oracle(oracle, opt, targets, pulse_axis=1)
        res = compare_chaotic(oracle, opt, targets, gk, gs[lo:hi], ck, E, pulse_axis=1)
        print(f"  configuration {c}: {res}", flush=True)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("engine,persistent", [(PIPE, False), (PIPE, True), (FUSED, False)])
def test_event_watchdog(oracle, engine, persistent):
    """st_ext.max_events (no counterpart in the reference): a particle whose
    interval needs more events is stopped with failureState -9 at the event
    past the bound, with NaN state like any other stop; 0xffffffff removes
    the bound (the same run as the default bound of 1e8)."""
    opt = configs.he3_thermal(n=256, tf=0.1, ioutInt=0.05)
    targets = st.output_times(opt)
    print(f"[event-loop] watchdog engine={engine} persistent={persistent}", flush=True)
    p = st.Particle(opt, engine=engine, max_events=10)
    p.initParticles()
    if persistent:
        p.run(targets)
    else:
        for t in targets:
            p.runSimulation(t)
    g = p.snapshot()
    p.close()
    # He-3 with tc = 2e-3 s takes ~25 gas collisions per 0.05 s interval
    assert np.all(g["failure"] == st._lib.FAIL_WATCHDOG)
    assert np.all(g["n_steps"] == 11) and np.all(np.isnan(g["t"])) and np.all(np.isnan(st.xyz(g["S"])))
    # the reference's loop takes more than 11 events in the first interval
    E = oracle.Ensemble(to_oracle_opt(opt), opt.numParticles)
    E.advance_to(targets[0])
    assert np.all(E.arr["n_steps"] > 11)
    g2, gs2, gc2 = run_gpu(opt, targets, engine=engine, persistent=persistent)
    p = st.Particle(opt, engine=engine, max_events=0xffffffff)
    p.initParticles()
    for t in targets:
        p.runSimulation(t)
    g3 = p.snapshot()
    p.close()
    assert np.all(g3["failure"] == 0)
    for k in ("n_steps", "n_bounce", "n_coll"):
        assert np.array_equal(g3[k], g2[k]), k
