"""The C ABI boundary and the host-side mirror, without a GPU: the library
loads and exports every symbol include/spintrack.h declares; the options
block has the reference's layout; the params.txt grammar and the binary
output format behave as the reference's parser and logger do."""
import ctypes as C
import io
import os
import re

import numpy as np

import spintrack as st
from spintrack import _lib, configs, logger, options
from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "spintrack.h")).read()
    return sorted(set(re.findall(r"ST_API\s+[\w\s\*]+?\b(st_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = st.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"
    assert lib.st_version().startswith(b"spintrack gfx950")


def test_error_path_without_device_call():
    lib = st.load()
    o = st.default_options()
    o.numParticles = 0
    h = C.c_void_p()
    assert lib.st_create(C.byref(o), None, C.byref(h)) == _lib.ST_E_ARG
    assert b"numParticles" in lib.st_last_error(None)
    o.numParticles = 10
    # integratorType 7 does not exist
    o.integratorType = 7
    assert lib.st_create(C.byref(o), None, C.byref(h)) == _lib.ST_E_UNSUPPORTED
    assert b"integratorType 7" in lib.st_last_error(None)
    # integratorType 6 (the spectral path) needs a drive frequency and a
    # sampling step that keeps the sample count bounded; it runs on plain
    # pipeline handles with spectrum_sum 0 or 1
    o.integratorType = 6
    o.w, o.h = 0.0, 1e-4
    assert lib.st_create(C.byref(o), None, C.byref(h)) == _lib.ST_E_ARG
    assert b"drive frequency" in lib.st_last_error(None)
    o.w, o.h = 6283.0, 1e-13
    assert lib.st_create(C.byref(o), None, C.byref(h)) == _lib.ST_E_ARG
    assert b"samples every h" in lib.st_last_error(None)
    o.h = 1e-4
    x6 = _lib.Ext()
    x6.pulse_axis, x6.spectrum_sum = 2, 2
    assert lib.st_create(C.byref(o), C.byref(x6), C.byref(h)) == _lib.ST_E_ARG
    assert b"spectrum_sum" in lib.st_last_error(None)
    x6.spectrum_sum, x6.engine = 1, _lib.ENGINE_FUSED
    assert lib.st_create(C.byref(o), C.byref(x6), C.byref(h)) == _lib.ST_E_UNSUPPORTED
    x6.engine = _lib.ENGINE_PIPELINE
    assert lib.st_create_batch(C.byref(o), 1, C.byref(x6), C.byref(h)) == _lib.ST_E_UNSUPPORTED
    assert b"plain handles" in lib.st_last_error(None)
    sp = _lib.Spectrum()
    assert lib.st_spectrum_get(None, C.byref(sp)) == _lib.ST_E_ARG
    # types 2-5 run on the pipeline engine; the fused engine refuses them
    # before touching the device
    o.integratorType = 2
    x = _lib.Ext()
    x.pulse_axis, x.variant, x.engine = 2, _lib.VARIANT_EXACT, _lib.ENGINE_FUSED
    assert lib.st_create(C.byref(o), C.byref(x), C.byref(h)) == _lib.ST_E_UNSUPPORTED
    assert b"pipeline engine" in lib.st_last_error(None)
    # the graph-captured chunk loop: 0..64 chunks, pipeline engine only
    o.integratorType = 0
    x.engine, x.graph_chunks = _lib.ENGINE_PIPELINE, 65
    assert lib.st_create(C.byref(o), C.byref(x), C.byref(h)) == _lib.ST_E_ARG
    assert b"graph_chunks" in lib.st_last_error(None)
    x.engine, x.graph_chunks = _lib.ENGINE_FUSED, 4
    assert lib.st_create(C.byref(o), C.byref(x), C.byref(h)) == _lib.ST_E_UNSUPPORTED
    assert lib.st_advance_to(None, 1.0) == _lib.ST_E_ARG


def test_options_layout_matches_oracle_and_reference(oracle):
    assert C.sizeof(_lib.Options) == 440 == oracle.lib().or_sizeof_options()
    for name, off in (("m", 168), ("diffuse", 352), ("nmax", 360), ("seed", 364), ("integratorType", 368),
                      ("numParticles", 372), ("iout", 380), ("dist", 384), ("gas_coll", 385),
                      ("keepStepSize", 388), ("noiseAmplitudes", 392), ("noiseFrequencies", 416)):
        assert getattr(_lib.Options, name).offset == off, name
        assert oracle.OPTIONS_DTYPE.fields[name][1] == off, name
    d = st.default_options()
    assert bytes(d) == oracle.default_options().tobytes()


def test_params_grammar():
    text = """dist, M
T, 0.45
gas_coll, TRUE
diffuse, 0.05
B0, 3.0e-6, 0, 0
numParticles, 1024
seed, -1
output, a
justonefield
fixedStepSize, yes
"""
    log = io.StringIO()
    o = options.parse_params(text, log)
    assert o.dist == b"M" and o.T == 0.45 and o.gas_coll == 1 and o.diffuse == 0.05
    assert tuple(o.B0) == (3e-6, 0.0, 0.0) and o.numParticles == 1024
    assert o.seed == 0xFFFFFFFF            # std::stoi then unsigned
    assert o.fixedStepSize == 0            # only "true" parses as true
    out = log.getvalue()
    assert "Unrecognized Option: output" in out and "Misunderstood input line" in out
    assert options.parse_params("").tf == 10.0      # defaults


def test_derive_and_output_times():
    o = st.default_options()
    o.m, o.T = 5.008e-27, 0.45
    options.derive(o)
    assert abs(o.tc / (1.6e-4 * 5.008e-27 / (1.380649e-23 * 0.45 ** 8)) - 1) < 1e-15
    o.t0, o.tf, o.ioutInt = 0.0, 1.0, 0.05
    ts = options.output_times(o)
    assert len(ts) == 20 and ts[-1] == 20 * 0.05
    o.t0 = 0.3     # the reference ignores t0 in nextTime
    assert options.output_times(o)[0] == 0.05


def test_binary_logger_roundtrip(tmp_path):
    o = configs.neutron_cell(n=7)
    path = str(tmp_path / "out.bin")
    rng = np.random.default_rng(0)
    snaps = []
    with logger.BinaryLogger(o, path) as lg:
        for _ in range(3):
            s = dict(t=rng.random(7), pos=np.zeros(7, st.COORDS), v=np.zeros(7, st.COORDS),
                     S=np.zeros(7, st.COORDS), failure=np.arange(7, dtype=np.int32),
                     n_coll=np.arange(7, dtype=np.uint64), n_bounce=np.arange(7, dtype=np.uint64) * 2,
                     n_steps=np.arange(7, dtype=np.uint64) * 3)
            s["S"]["x"] = rng.random(7)
            lg.writeSnapshot(*[s[k] for k in ("t", "pos", "v", "S", "failure", "n_coll", "n_bounce", "n_steps")])
            snaps.append(s)
    assert os.path.getsize(path) == 440 + 3 * 108 * 7
    o2, data = logger.read_output(path)
    assert bytes(o2) == bytes(o)
    for k in range(3):
        assert np.array_equal(data["t"][k], snaps[k]["t"])
        assert np.array_equal(data["S"][k]["x"], snaps[k]["S"]["x"])
        assert np.array_equal(data["n_steps"][k], snaps[k]["n_steps"])


def test_snapshot_block_dtype_matches_writer():
    dt = st.snapshot_block_dtype(5)
    assert dt.itemsize == 540
    assert dt.fields["n_coll"][1] == 84 * 5
