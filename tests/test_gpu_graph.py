"""The graph-captured chunk loop (st_ext.graph_chunks): K pipeline chunks
captured once per call and replayed until every particle is done.  Runs
that need many chunks (short chunk_records) and therefore several replays,
with eager work (st_advance_to, snapshots, synchronisations) between the
graph runs, give states, snapshot blocks and counters identical to the
host-driven chunk loop; so do runs with integrator failures, whose
kinematics replay (k_fixup) reads its count on the device, and batched
handles."""
import numpy as np
import pytest

import spintrack as st
from spintrack import configs
from spintrack.particle import snapshot_block_dtype

pytestmark = pytest.mark.gpu


def drive(opt_or_cfgs, graph_chunks, variant, chunk_records=16, batch=False, stop_at=None):
    kw = dict(variant=variant, chunk_records=chunk_records, graph_chunks=graph_chunks)
    p = st.Particle.batch(opt_or_cfgs, **kw) if batch else st.Particle(opt_or_cfgs, **kw)
    opt = opt_or_cfgs[0] if batch else opt_or_cfgs
    targets = st.output_times(opt)
    with p:
        p.initParticles()
        if stop_at is not None:
            p.set_stop_times(stop_at)
        bb = p.block_bytes()
        buf = st._lib.DeviceBuffer(bb * len(targets))
        half = len(targets) // 2
        # graph (or host) runs, with eager calls and synchronisations between
        p.run(targets[:half], buf.ptr)
        p.synchronize()
        mid = p.snapshot()
        p.runSimulation(targets[half])
        p.snapshot_block_device(buf.ptr + half * bb)
        p.synchronize()
        p.run(targets[half + 1:], buf.ptr + (half + 1) * bb)
        raw = p.copy_to_host(buf.ptr, bb * len(targets))
        buf.free()
        state = p.get_state()
        cnt = p.particle_counters()
    blocks = [raw[k * bb:(k + 1) * bb].tobytes() for k in range(len(targets))]
    return mid, state, blocks, cnt


def same(a, b):
    ma, sa, ba, ca = a
    mb, sb, bb, cb = b
    for f in ma:
        assert np.asarray(ma[f]).tobytes() == np.asarray(mb[f]).tobytes(), f
    assert sa.tobytes() == sb.tobytes()
    for k, (x, y) in enumerate(zip(ba, bb)):
        assert x == y, f"block {k}"
    for f in ("accepted", "rejected", "segments"):
        assert np.array_equal(ca[f], cb[f]), f


@pytest.mark.parametrize("variant", [st.VARIANT_EXACT, st.VARIANT_FAST])
def test_graph_chunks_equal_host_loop(variant):
    opt = configs.he3_thermal(n=3000, tf=0.2, ioutInt=0.05)
    print(f"[graph] he3 n=3000 variant={variant}", flush=True)
    ref = drive(opt, 0, variant)
    for K in (2, 5):
        same(drive(opt, K, variant), ref)


def test_graph_chunks_with_integrator_failures(oracle):
    opt = configs.he3_thermal(n=600, tf=0.2, ioutInt=0.05)
    opt.nmax = 4
    print("[graph] nmax failures", flush=True)
    ref = drive(opt, 0, st.VARIANT_EXACT)
    got = drive(opt, 4, st.VARIANT_EXACT)
    same(got, ref)
    fails = np.frombuffer(got[2][-1], dtype=snapshot_block_dtype(600))[0]["failure"]
    assert (fails == -1).sum() > 20


def test_graph_chunks_batched():
    cfgs, stop_at = configs.batched_sweep(n_cfg=4, n_per=512, tf=0.4, ioutInt=0.1, dur=(0.15, 0.4))
    print("[graph] batched sweep 4 x 512", flush=True)
    ref = drive(cfgs, 0, st.VARIANT_FAST, batch=True, stop_at=stop_at)
    same(drive(cfgs, 6, st.VARIANT_FAST, batch=True, stop_at=stop_at), ref)
