"""Host side of the batched form, without a GPU: argument checks of
st_create_batch (they run before any device call), the host-side stream
generator against the oracle's RNG, and the sweep generator's contract."""
import ctypes as C

import numpy as np

import spintrack as st
from spintrack import _lib, configs, rng


def test_stream_generator_matches_oracle(oracle):
    for seed, index in ((0, 0), (2, 0), (2, 1), (7, 2 ** 20 - 1), (0xFFFFFFFF, 3)):
        s = oracle.rng_seeded(seed, index)
        want = oracle.xoshiro(s, 200)
        g = rng.Streams(seed, [index])
        got = np.array([g.next_u64()[0] for _ in range(200)], dtype=np.uint64)
        assert np.array_equal(got, want), (seed, index)
    # vectorised over streams == one stream at a time
    idx = np.arange(50, dtype=np.uint64) * 977
    v = rng.Streams(5, idx).uniform()
    for k in (0, 17, 49):
        assert v[k] == oracle.uniform(oracle.rng_seeded(5, int(idx[k])), 1)[0]


def test_sweep_contract():
    cfgs, stop = configs.batched_sweep(n_cfg=8, n_per=64)
    assert len(cfgs) == 8 and stop.shape == (512,)
    G = np.array([[*c.Gx, *c.Gy, *c.Gz] for c in cfgs])
    assert np.all(np.abs(G) <= 5e-8) and np.unique(G).size == G.size
    # drawn config-major from stream (2, 0)
    assert np.array_equal(G.ravel(), 5e-8 * (2.0 * rng.uniforms(2, 0, 72) - 1.0))
    assert [c.E.z for c in cfgs] == [7.5e6, -7.5e6] * 4
    assert [c.diffuse for c in cfgs] == [0.0] * 4 + [0.1] * 4
    assert [c.seed for c in cfgs] == [2 + 64 * c for c in range(8)]
    assert stop.min() >= 1.0 and stop.max() <= 20.0
    assert abs(stop.mean() - 10.5) < 1.0


def test_create_batch_rejects_mismatched_run_options():
    lib = st.load()
    a = configs.neutron_cell(n=8)
    b = configs.neutron_cell(n=8)
    b.L = _lib.Coords(0.07, 0.1, 0.5)
    arr = (_lib.Options * 2)(a, b)
    h = C.c_void_p()
    assert lib.st_create_batch(arr, 2, None, C.byref(h)) == _lib.ST_E_ARG
    assert b"'L'" in lib.st_last_error(None)
    b = configs.neutron_cell(n=8)
    b.rtol = 1e-10
    arr = (_lib.Options * 2)(a, b)
    assert lib.st_create_batch(arr, 2, None, C.byref(h)) == _lib.ST_E_ARG
    assert b"rtol" in lib.st_last_error(None)
    assert lib.st_create_batch(arr, 0, None, C.byref(h)) == _lib.ST_E_ARG
    fused = _lib.Ext(2, 0, 0, _lib.ENGINE_FUSED, 0)
    assert lib.st_create_batch(arr, 1, C.byref(fused), C.byref(h)) == _lib.ST_E_UNSUPPORTED
