"""On-device ensemble analysis against the oracle's restatement of the
reference's analysis: the block sums (st_block_sums) and the per-particle
phase statistics (st_phase_stats) of real snapshot blocks."""
import numpy as np
import pytest

import analysis as OA
import spintrack as st
from spintrack import _lib, configs
from spintrack.particle import snapshot_block_dtype

pytestmark = pytest.mark.gpu


def run_blocks(cfgs, targets):
    b = st.Particle.batch(cfgs, variant=st.VARIANT_FAST)
    b.initParticles()
    bb = b.block_bytes()
    buf = _lib.DeviceBuffer(bb * len(targets))
    b.run(targets, buf.ptr)
    raw = b.copy_to_host(buf.ptr, bb * len(targets))
    blocks = [raw[k * bb:(k + 1) * bb].view(snapshot_block_dtype(b.n))[0] for k in range(len(targets))]
    return b, buf, blocks


def test_phase_stats_match_the_reference_analysis():
    # two configurations: neutrons with spins transverse to B0, the second
    # with a large gradient that dephases the ensemble (spread phases, cuts)
    cfgs, _ = configs.batched_sweep(n_cfg=2, n_per=4096, tf=2.0, ioutInt=0.25)
    cfgs[1].Gx = _lib.Coords(0.0, 4e-7, 6e-7)
    targets = st.output_times(cfgs[0])
    K = len(targets)
    b, buf, blocks = run_blocks(cfgs, targets)
    off = b.layout()
    phi = _lib.DeviceBuffer(8 * b.n * K)
    stats = b.phase_stats(buf.ptr, K, phi.ptr)
    dev_phi = b.copy_to_host(phi.ptr, 8 * b.n * K).view(np.float64).reshape(K, b.n)
    sums = b.block_sums(buf.ptr, K)
    S = np.stack([st.xyz(blk["S"]) for blk in blocks])            # (K, N, 3)
    fails = blocks[-1]["failure"]
    for c in range(2):
        lo, hi = int(off[c]), int(off[c + 1])
        # the device phases are atan2(Sz, Sy) of the selected particles
        ref_phis = OA.select(S[:, lo:hi], fails[lo:hi])
        got = dev_phi[:, lo:hi]
        got = got[:, ~np.isnan(got[-1])]
        assert got.shape == ref_phis.shape
        assert np.max(np.abs(got - ref_phis)) < 1e-15 * 4
        # the reductions, on the same phases
        means, stds, kept, refs = OA.analyze(got)
        print(f"[phase] cfg {c}: means {means[[0, -1]]}, stds {stds[[0, -1]]}, kept {kept[[0, -1]]}", flush=True)
        assert np.array_equal(stats[:, c, 4], refs)
        assert np.array_equal(stats[:, c, 2], kept)
        assert np.all(stats[:, c, 3] == got.shape[1])
        assert np.allclose(stats[:, c, 0], means, rtol=0, atol=1e-12)
        assert np.allclose(stats[:, c, 1], stds, rtol=1e-10, atol=1e-14)
        # block sums: counts and mean spins
        m = S[:, lo:hi].mean(axis=1)
        assert np.all(sums[:, c, 0] == hi - lo)
        assert np.allclose(sums[:, c, 1:] / sums[:, c, :1], m, rtol=0, atol=1e-14)
    # the dephased configuration really spreads out
    assert stats[-1, 1, 1] > 10 * stats[-1, 0, 1]
    phi.free()
    buf.free()
    b.close()
