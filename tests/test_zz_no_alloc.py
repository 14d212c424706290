"""st_run_noalloc (the st_run form that allocates no device memory) against
st_run through the C ABI: the same per-particle state and snapshot blocks,
bit for bit, on two calls in a row that share the library's static scratch
array, the second on a smaller ensemble with fewer output times; sizes past
the scratch's maximum are refused."""
import numpy as np
import pytest

import spintrack as st
from spintrack import configs

pytestmark = pytest.mark.gpu


def run(opt, targets, noalloc, variant, engine, graph_chunks=0):
    p = st.Particle(opt, variant=variant, engine=engine, graph_chunks=graph_chunks)
    p.initParticles()
    bb = p.block_bytes()
    blocks = st._lib.DeviceBuffer(len(targets) * bb)
    (p.run_noalloc if noalloc else p.run)(targets, blocks.ptr)
    host = p.copy_to_host(blocks.ptr, blocks.nbytes)
    blocks.free()
    state, cnt = p.get_state(), p.particle_counters()
    p.close()
    return host, state, cnt


@pytest.mark.parametrize("variant,engine,graph", [(st.VARIANT_FAST, st._lib.ENGINE_PIPELINE, 0),
                                                  (st.VARIANT_EXACT, st._lib.ENGINE_PIPELINE, 4),
                                                  (st.VARIANT_EXACT, st._lib.ENGINE_FUSED, 0)])
def test_noalloc_equals_st_run(variant, engine, graph):
    # first call: 1500 He-3 atoms, 6 output times; second: 300 atoms, 2 times
    for n, tf, iout in ((1500, 0.3, 0.05), (300, 0.1, 0.05)):
        opt = configs.he3_thermal(n=n, tf=tf, ioutInt=iout)
        targets = st.output_times(opt)
        print(f"[noalloc] n={n} targets={len(targets)} variant={variant} engine={engine} graph={graph}", flush=True)
        ref = run(opt, targets, False, variant, engine, graph)
        got = run(opt, targets, True, variant, engine, graph)
        assert got[0].tobytes() == ref[0].tobytes(), "snapshot blocks differ"
        assert got[1].tobytes() == ref[1].tobytes(), "state differs"
        for k in ref[2]:
            assert np.array_equal(got[2][k], ref[2][k]), k


def test_noalloc_limits():
    opt = configs.neutron_cell(n=64, tf=0.1)
    p = st.Particle(opt)
    p.initParticles()
    too_many = 0.1 + 1e-4 * np.arange(1, st._lib.NOALLOC_MAX_TARGETS + 2)
    with pytest.raises(st._lib.SpinTrackError, match="at most 4096"):
        p.run_noalloc(too_many)
    p.set_checkpointing(True)
    with pytest.raises(st._lib.SpinTrackError, match="checkpoint"):
        p.run_noalloc([0.05])
    p.set_checkpointing(False)
    # the largest accepted size runs
    p.run_noalloc(0.1 + 1e-5 * np.arange(1, st._lib.NOALLOC_MAX_TARGETS + 1))
    assert np.all(p.snapshot()["failure"] == 0)
    p.close()
