"""The field-hook surface (include/spintrack_fields.h), the counterpart of the
reference's user-edited pulse()/grad() (src/integrator.cpp:40-46,
README.md "Modification of Field Gradients and Magnetic Fields").

CPU checks: the oracle built with the test hooks of tests/fields/
quadratic_hooks.h evaluates exactly those fields in its right-hand side
(against a numpy restatement of the test hooks, and unlike the default
build); the library built with them loads, exports every symbol, names its
hooks, and refuses the fast variant (whose per-segment polynomial needs a
linear gradient and a cos pulse) before touching the GPU.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle as O
import spintrack as st
from spintrack import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_FIELDTEST = os.path.join(ROOT, "oracle", "liboracle_fieldtest.so")
LIB_FIELDTEST = os.path.join(ROOT, "spin-tracking_amd", "libspintrack_fieldtest.so")
Q = 2.0e-6   # ST_TEST_Q of the test hooks


def test_hooks(r, G, t, a, w, axis):
    """numpy restatement of tests/fields/quadratic_hooks.h"""
    x, y, z = r
    g = np.array(G).reshape(3, 3) @ np.array(r) + Q * np.array([z * z - x * x, x * y, x * z])
    c, s = a * math.cos(w * t), a * math.sin(w * t)
    p = np.array([0.0 if axis == 0 else c, 0.0 if axis == 1 else (c if axis == 0 else s),
                  0.0 if axis == 2 else s])
    return g, p


test_hooks.__test__ = False   # a helper, not a test


def opt_with_fields(oracle):
    o = oracle.default_options()
    o["B0"] = (3e-6, 1e-7, -2e-7)
    o["E"] = (1e6, -2e6, 7.5e6)
    o["Gx"] = (1e-8, 2e-8, -3e-8)
    o["Gy"] = (4e-9, -5e-9, 6e-9)
    o["Gz"] = (-7e-9, 8e-9, 9e-9)
    o["a"] = 1e-7
    o["w"] = 2 * math.pi * 50
    o["gravity"] = True
    return o


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_oracle_uses_the_hooks(oracle, axis):
    """findCrossTerm of the field-test oracle = gamma*(pulse + B0 + v x E/c^2
    + grad) with the test hooks, at positions along a gravity segment."""
    if not os.path.exists(ORACLE_FIELDTEST):
        pytest.skip("liboracle_fieldtest.so not built (__graft_entry__.build())")
    o = opt_with_fields(oracle)
    rng = np.random.default_rng(axis)
    with oracle.library(ORACLE_FIELDTEST):
        for _ in range(20):
            t0 = rng.uniform(0, 1)
            t = t0 + rng.uniform(0, 0.01)
            p0 = rng.uniform(-0.03, 0.03, 3)
            v0 = rng.uniform(-5, 5, 3)
            got = oracle.cross_term(o, t, t0, p0, v0, pulse_axis=axis)
            p, v = oracle.interpolate(o, t, t0, p0, v0)
            g, pl = test_hooks(p, list(o["Gx"]) + list(o["Gy"]) + list(o["Gz"]), t, o["a"], o["w"], axis)
            B = pl + np.array(o["B0"]) + np.cross(v, np.array(o["E"])) / 299792458.0 ** 2 + g
            want = o["gamma"] * B
            assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
        # the noise term's gradient is the hook's too (findNoiseTerm, :72-80)
        p0, v0 = np.array([0.02, -0.01, 0.1]), np.array([1.0, 2.0, -3.0])
        n = oracle.noise_term(o, 0.5, 0.49, p0, v0)
        p, v = oracle.interpolate(o, 0.5, 0.49, p0, v0)
        g, _ = test_hooks(p, list(o["Gx"]) + list(o["Gy"]) + list(o["Gz"]), 0.5, 0, 0, axis)
        want = o["gamma"] * (np.cross(v, np.array(o["E"])) / 299792458.0 ** 2 + g)
        assert np.abs(n - want).max() <= 1e-13 * np.abs(want).max()
    # the default build evaluates other fields
    d = oracle.cross_term(o, 0.5, 0.49, p0, v0, pulse_axis=axis)
    with oracle.library(ORACLE_FIELDTEST):
        f = oracle.cross_term(o, 0.5, 0.49, p0, v0, pulse_axis=axis)
    assert np.abs(d - f).max() > 1e-3 * np.abs(d).max()


def test_fieldtest_library_exports_and_refuses_fast():
    if not os.path.exists(LIB_FIELDTEST):
        pytest.skip("libspintrack_fieldtest.so not built (__graft_entry__.build())")
    L = _lib.open_library(LIB_FIELDTEST)
    for name in _lib.SIGNATURES:
        assert hasattr(L, name), name
    v = L.st_version().decode()
    assert "quadratic_hooks.h" in v and ",fast" not in v, v
    assert "fields=default" in _lib.load().st_version().decode()
    opt = st.configs.neutron_cell(n=16, tf=0.1)
    h = C.c_void_p()
    x = _lib.Ext(2, _lib.VARIANT_FAST)
    rc = L.st_create(C.byref(opt), C.byref(x), C.byref(h))
    assert rc == _lib.ST_E_UNSUPPORTED and not h.value
    assert b"fast variant" in L.st_last_error(None)
