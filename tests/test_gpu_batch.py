"""Batched handles (st_create_batch) on the GPU: every configuration of a
batch reproduces the plain single-configuration run of the same options --
state, RNG, counters and every snapshot block -- bit for bit in the exact
variant, and with bit-exact trajectories, RNG and counters and spins within
1e-12 in the fast variant, so the batched form inherits the plain path's
parity with the oracle."""
import numpy as np
import pytest

import spintrack as st
from spintrack import _lib, configs
from spintrack.particle import snapshot_block_dtype
from conftest import to_oracle_opt

pytestmark = pytest.mark.gpu


def run(handle, targets, stop_at):
    handle.initParticles()
    if stop_at is not None:
        handle.set_stop_times(stop_at)
    bb = handle.block_bytes()
    buf = _lib.DeviceBuffer(bb * len(targets))
    handle.run(targets, buf.ptr)
    raw = handle.copy_to_host(buf.ptr, bb * len(targets))
    buf.free()
    blocks = [raw[k * bb:(k + 1) * bb].view(snapshot_block_dtype(handle.n))[0] for k in range(len(targets))]
    return handle.get_state(), blocks, handle.particle_counters()


def block_slice(block, a, b):
    return {k: block[k][a:b] for k in block.dtype.names}


def same_bits(x, y):
    return np.asarray(x).tobytes() == np.asarray(y).tobytes()


def check_batch_equals_plain(cfgs, targets, stop_at, variant, pulse_axis=2):
    print(f"[batch] {len(cfgs)} configs, n={[c.numParticles for c in cfgs]}, variant={variant}", flush=True)
    with st.Particle.batch(cfgs, variant=variant, pulse_axis=pulse_axis) as b:
        off = b.layout()
        bs, bblocks, bcnt = run(b, targets, stop_at)
    assert list(off) == list(np.cumsum([0] + [c.numParticles for c in cfgs]))
    for c, cfg in enumerate(cfgs):
        lo, hi = int(off[c]), int(off[c + 1])
        with st.Particle(cfg, variant=variant, pulse_axis=pulse_axis) as p:
            ps, pblocks, pcnt = run(p, targets, None if stop_at is None else stop_at[lo:hi])
        if variant == st.VARIANT_EXACT:
            assert same_bits(bs[lo:hi], ps), f"config {c}: state differs"
        else:
            # the fast variant lets the compiler choose FMA contraction per
            # kernel instantiation, so the batched and plain spin kernels may
            # round an operation differently: the trajectory, RNG and counters
            # stay bit-exact, spins agree to integration round-off
            for name in ps.dtype.names:
                if name != "S":
                    assert same_bits(bs[lo:hi][name], ps[name]), f"config {c}: {name} differs"
            dS = np.abs(st.xyz(bs[lo:hi]["S"]) - st.xyz(ps["S"]))
            assert np.nanmax(dS, initial=0.0) < 1e-12, f"config {c}: max |dS| {np.nanmax(dS)}"
        for k in range(len(targets)):
            bb = block_slice(bblocks[k], lo, hi)
            for name in bb:
                if variant == st.VARIANT_EXACT or name != "S":
                    assert same_bits(bb[name], pblocks[k][name]), f"config {c} block {k} field {name}"
                else:
                    d = np.abs(st.xyz(bb[name]) - st.xyz(pblocks[k][name]))
                    assert np.nanmax(d, initial=0.0) < 1e-12, f"config {c} block {k} S"
        for name in ("accepted", "rejected", "segments"):
            assert np.array_equal(bcnt[name][lo:hi], pcnt[name]), (c, name)
    return bs, off


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
def test_sweep_batch_equals_plain_runs(variant):
    # six configurations of the field sweep (random gradients, +-Ez, specular
    # and diffuse walls) with uneven per-particle stop times
    cfgs, stop_at = configs.batched_sweep(n_cfg=6, n_per=192, tf=1.0, ioutInt=0.25, dur=(0.3, 1.0))
    targets = st.output_times(cfgs[0])
    bs, off = check_batch_equals_plain(cfgs, targets, stop_at, variant)
    # the configurations really differ: final spins are not identical
    assert not same_bits(bs["S"][off[0]:off[1]], bs["S"][off[1]:off[2]])
    # the stops took effect: every particle ends at its own stop time
    assert np.allclose(bs["t"], stop_at, rtol=0, atol=1e-9)


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
def test_cotracked_species_batch_equals_plain_runs(variant):
    # neutrons and He-3 in one batch: per-configuration gamma, thermal speed
    # and seed, gas collisions, pulse along y
    cfgs = configs.dressing_cotracked(n=160, tf=0.02, ioutInt=0.005)
    for c in cfgs:
        c.rtol = c.atol = 1e-10
    targets = st.output_times(cfgs[0])
    check_batch_equals_plain(cfgs, targets, None, variant, pulse_axis=1)


def test_batch_configuration_matches_oracle(oracle):
    # a batched configuration other than the first, directly against the oracle
    cfgs, _ = configs.batched_sweep(n_cfg=4, n_per=96, tf=0.5, ioutInt=0.25)
    targets = st.output_times(cfgs[0])
    with st.Particle.batch(cfgs, variant=st.VARIANT_EXACT) as b:
        b.initParticles()
        b.run(targets)
        s = b.snapshot()
        off = b.layout()
    c = 1   # specular walls (configurations 2 and 3 are diffuse)
    E = oracle.Ensemble(to_oracle_opt(cfgs[c]), cfgs[c].numParticles)
    prev = cfgs[c].t0
    for t in targets:
        E.advance_to(t, prev_t=prev)
        prev = t
    o = E.snapshot()
    lo, hi = int(off[c]), int(off[c + 1])
    assert np.array_equal(s["n_bounce"][lo:hi], o["n_bounce"])
    assert np.array_equal(s["n_steps"][lo:hi], o["n_steps"])
    dS = np.abs(st.xyz(s["S"][lo:hi]) - oracle.xyz(o["S"])).max()
    dp = np.abs(st.xyz(s["pos"][lo:hi]) - oracle.xyz(o["pos"])).max()
    print(f"[batch-oracle] max |dS| {dS:.3g}, max |dpos| {dp:.3g}", flush=True)
    assert dS < 1e-9 and dp < 1e-9
