"""integratorTypes 2-5 on the GPU against the oracle: integrateMagnusCFET
(2), integrateRK45Quaternion (3), integrateRKF45 (4) and
integrateRKF45Quaternion (5) of src/integrator.cpp:339-723, in both variants,
on specular neutrons (deterministic trajectories) and thermal He-3 with gas
collisions; the fixed-step quirk (such calls end only at nmax, -2) with the
kinematics replay of the failing particles; batched handles."""
import numpy as np
import pytest

import spintrack as st
from spintrack import configs
from test_gpu_parity import TOL, check_integer_state, compare, compare_chaotic, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

# rtol per type such that steps stay above hmin: RKF45's "error" is the
# size of the step's increment and RKF45Q's the sixth rotation factor, so
# both need loose tolerances (src/integrator.cpp:398, :587-593)
CASES = {2: (1e-12, 0.1), 3: (1e-12, 0.1), 4: (1e-4, 0.01), 5: (1e-4, 0.05)}
NAMES = {2: "magnus", 3: "rk45q", 4: "rkf45", 5: "rkf45q"}


def neutrons(integ, n=256):
    rtol, tf = CASES[integ]
    opt = configs.neutron_cell(n=n, tf=tf, ioutInt=tf / 2)
    opt.integratorType = integ
    opt.rtol = opt.atol = rtol
    opt.yi.x, opt.yi.y, opt.yi.z = 0.0, 0.6, 0.8       # transverse to B0
    return opt


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
@pytest.mark.parametrize("integ", [2, 3, 4, 5])
def test_neutrons_match_oracle(oracle, integ, variant):
    opt = neutrons(integ)
    targets = st.output_times(opt)
    print(f"[integrators] {NAMES[integ]} variant={variant} n={opt.numParticles}", flush=True)
    for persistent in (False, True):
        g, gs, gc = run_gpu(opt, targets, variant=variant, persistent=persistent)
        E = run_oracle(oracle, opt, targets)
        # Magnus and RK45Q: on this cell's weak gradients the embedded error
        # is at round-off level, so the controller follows last-bit libm
        # differences and a particle's step count may differ by a few
        res = compare(oracle, g, gs, gc, E, tol=TOL, step_slack=3 if integ in (2, 3) else 0)
        assert np.all(g["failure"] == 0)
        # the spins turned: the comparison is not of a resting state
        assert np.abs(st.xyz(g["S"])[:, 1] - 0.6).max() > 0.1
        print(f"  persistent={persistent}: {res}", flush=True)


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
@pytest.mark.parametrize("integ", [2, 3])
def test_he3_with_gas_matches_oracle(oracle, integ, variant):
    opt = configs.he3_thermal(n=256, tf=0.1, ioutInt=0.05, integrator=integ)
    opt.yi.x, opt.yi.y, opt.yi.z = 0.0, 1.0, 0.0
    targets = st.output_times(opt)
    print(f"[integrators] he3 {NAMES[integ]} variant={variant}", flush=True)
    g, gs, gc = run_gpu(opt, targets, variant=variant)
    E = run_oracle(oracle, opt, targets)
    print(" ", compare_chaotic(oracle, opt, targets, g, gs, gc, E), flush=True)


@pytest.mark.parametrize("integ", [2, 3, 4, 5])
def test_fixed_step_ends_at_nmax(oracle, integ):
    # with fixedStepSize these four never set their stop flag: each call runs
    # nmax steps of h past the segment and returns -2, stopping the particle
    # (NaN state); the pipeline replays the kinematics to the failing event
    opt = neutrons(integ, n=128)
    opt.fixedStepSize = True
    opt.h = 1e-4
    opt.nmax = 40
    targets = st.output_times(opt)
    print(f"[integrators] fixed-step {NAMES[integ]}", flush=True)
    g, gs, gc = run_gpu(opt, targets, seg_buffer_mb=1)
    E = run_oracle(oracle, opt, targets)
    r = check_integer_state(oracle, g, gs, gc, E)
    assert np.all(r["failure"] == -2)
    assert np.all(np.isnan(st.xyz(g["S"]))) and np.all(np.isnan(g["t"]))


@pytest.mark.parametrize("integ,variant", [(2, st.VARIANT_EXACT), (5, st.VARIANT_FAST)])
def test_batched_handles(integ, variant):
    from test_gpu_batch import check_batch_equals_plain
    cfgs, stop_at = configs.batched_sweep(n_cfg=3, n_per=128, tf=CASES[integ][1], ioutInt=CASES[integ][1] / 2,
                                          dur=(0.5 * CASES[integ][1], CASES[integ][1]))
    for c in cfgs:
        c.integratorType = integ
        c.rtol = c.atol = CASES[integ][0]
        c.yi.x, c.yi.y, c.yi.z = 0.0, 0.6, 0.8
    check_batch_equals_plain(cfgs, st.output_times(cfgs[0]), stop_at, variant)
