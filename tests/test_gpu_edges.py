"""Edge cases of the GPU path against the oracle: ensemble sizes that do not
fill a wave or a workgroup, a large ragged ensemble checked on a strided
sample, and a zero-length interval (an advance to the current time)."""
import numpy as np
import pytest

import spintrack as st
from spintrack import configs
from conftest import to_oracle_opt
from test_gpu_parity import TOL, compare, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 63, 65, 257])
def test_ragged_ensemble_sizes(oracle, n):
    opt = configs.neutron_cell(n=n, tf=0.3, ioutInt=0.1)
    targets = st.output_times(opt)
    print(f"[edges] n={n}", flush=True)
    for persistent in (False, True):
        g, gs, gc = run_gpu(opt, targets, persistent=persistent)
        E = run_oracle(oracle, opt, targets)
        compare(oracle, g, gs, gc, E, tol=TOL)


def test_large_ragged_ensemble_strided_sample(oracle):
    # 2^21 + 5 neutrons; particles 0, k, 2k, ... checked against the oracle
    # seeded as the same global indices
    n, stride = (1 << 21) + 5, 4099
    opt = configs.neutron_cell(n=n, tf=0.05, ioutInt=0.025)
    targets = st.output_times(opt)
    print(f"[edges] n={n}", flush=True)
    p = st.Particle(opt, variant=st.VARIANT_EXACT)
    p.initParticles()
    p.run(targets)
    s = p.snapshot()
    p.close()
    idx = np.arange(0, n, stride)
    o = np.frombuffer(bytes(opt), dtype=oracle.OPTIONS_DTYPE)[0].copy()
    o["numParticles"] = idx.size
    E = oracle.Ensemble(o, idx.size, first=0, stride=stride)
    prev = opt.t0
    for t in targets:
        E.advance_to(t, prev_t=prev)
        prev = t
    r = E.snapshot()
    assert np.array_equal(s["n_bounce"][idx], r["n_bounce"])
    assert np.array_equal(s["n_steps"][idx], r["n_steps"])
    assert np.abs(st.xyz(s["S"][idx]) - oracle.xyz(r["S"])).max() < TOL
    assert np.abs(st.xyz(s["pos"][idx]) - oracle.xyz(r["pos"])).max() < TOL
    assert np.all(s["failure"] == 0)
    # the last particle (ragged tail) moved like the others
    assert s["n_steps"][-1] > 0 and np.isfinite(st.xyz(s["S"][-1:])).all()


def test_zero_length_interval(oracle):
    # advancing to the time the ensemble already has runs one empty segment,
    # as the reference's loop does; GPU and oracle agree
    opt = configs.neutron_cell(n=128, tf=0.2, ioutInt=0.1)
    targets = [0.1, 0.1, 0.2]
    g, gs, gc = run_gpu(opt, targets)
    E = oracle.Ensemble(to_oracle_opt(opt), opt.numParticles)
    prev = opt.t0
    for t in targets:
        E.advance_to(t, prev_t=prev)
        prev = t
    compare(oracle, g, gs, gc, E, tol=TOL)


@pytest.mark.parametrize("n", [70001, 256, 255])
def test_snapshot_block_device_ragged(n):
    # k_snapshot stages the coordinate fields of each 256-particle group in
    # LDS; a ragged last group, odd N (4-byte aligned counter arrays) and
    # exact multiples all give the reference's block byte for byte
    opt = configs.he3_thermal(n=n, tf=0.05, ioutInt=0.05)
    print(f"[edges] snapshot block n={n}", flush=True)
    with st.Particle(opt, variant=st.VARIANT_FAST) as p:
        p.initParticles()
        p.runSimulation(0.05)
        bb = p.block_bytes()
        assert bb == 108 * n
        buf = st._lib.DeviceBuffer(bb + 64)
        p.snapshot_block_device(buf.ptr)
        raw = p.copy_to_host(buf.ptr, bb + 64)
        buf.free()
        s = p.snapshot()
    b = np.frombuffer(raw[:bb].tobytes(), dtype=st.snapshot_block_dtype(n))[0]
    for f in ("t", "pos", "v", "S", "failure", "n_coll", "n_bounce", "n_steps"):
        assert b[f].tobytes() == s[f].tobytes(), f
