"""GPU parity with non-default field hooks: the library and the oracle both
built with tests/fields/quadratic_hooks.h (a gradient with a quadratic
term, a pulse rotating about the pulse axis) -- the counterpart of a user
editing the reference's pulse()/grad() (src/integrator.cpp:40-46) and
recompiling.  The exact variant runs the hooks as written and must match the
oracle to the stated bar; the fast variant and the adjoint, which need the
linear/cos field, are refused."""
import math
import os

import numpy as np
import pytest

import spintrack as st
from spintrack import configs
from conftest import to_oracle_opt
from test_gpu_parity import PIPE, FUSED, compare, run_oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_FIELDTEST = os.path.join(ROOT, "oracle", "liboracle_fieldtest.so")
LIB_FIELDTEST = os.path.join(ROOT, "spin-tracking_amd", "libspintrack_fieldtest.so")


def fields_opt(n=256, tf=0.2, integ=0):
    opt = configs.neutron_cell(n=n, tf=tf, ioutInt=tf / 4)
    opt.integratorType = integ
    opt.a = 1e-7
    opt.w = 2 * math.pi * 50
    opt.yi.x, opt.yi.y, opt.yi.z = 0.0, 0.6, 0.8
    return opt


def run_lib(opt, targets, library, engine=PIPE, persistent=False, pulse_axis=2):
    p = st.Particle(opt, engine=engine, pulse_axis=pulse_axis, library=library)
    p.initParticles()
    if persistent:
        p.run(targets)
    else:
        for t in targets:
            p.runSimulation(t)
    out = p.snapshot(), p.get_state(), p.particle_counters()
    p.close()
    return out


@pytest.mark.parametrize("integ", [0, 2])
@pytest.mark.parametrize("engine,persistent", [(PIPE, False), (PIPE, True), (FUSED, False)])
def test_custom_hooks_match_oracle(oracle, integ, engine, persistent):
    if integ == 2 and engine == FUSED:
        pytest.skip("the fused engine carries DOP853 and RK45 only")
    opt = fields_opt(integ=integ)
    targets = st.output_times(opt)
    print(f"[fields] integrator={integ} engine={engine} persistent={persistent}", flush=True)
    g, gs, gc = run_lib(opt, targets, LIB_FIELDTEST, engine, persistent, pulse_axis=1)
    with oracle.library(ORACLE_FIELDTEST):
        E = run_oracle(oracle, opt, targets, pulse_axis=1)
    res = compare(oracle, g, gs, gc, E, step_slack=3 if integ == 2 else 0)
    print("  ", res, flush=True)
    # the hooks matter: the default library's spins differ
    d, _, _ = run_lib(opt, targets, None, engine, persistent, pulse_axis=1)
    assert np.abs(st.xyz(d["S"]) - st.xyz(g["S"])).max() > 1e-6


def test_custom_hooks_refuse_fast_and_adjoint():
    opt = fields_opt(n=64, tf=0.1)
    with pytest.raises(st._lib.SpinTrackError, match="fast variant"):
        st.Particle(opt, variant=st.VARIANT_FAST, library=LIB_FIELDTEST)
    p = st.Particle(opt, library=LIB_FIELDTEST)
    p.initParticles()
    p.set_checkpointing(True)
    targets = st.output_times(opt)
    bb = p.block_bytes()
    blocks = st._lib.DeviceBuffer((len(targets) + 1) * bb)
    p.snapshot_block_device(blocks.ptr)
    p.run(targets, blocks.ptr + bb)
    with pytest.raises(st._lib.SpinTrackError, match="linear-gradient"):
        p.adjoint(blocks.ptr + bb, np.zeros((1, len(targets), 3)))
    blocks.free()
    p.close()
