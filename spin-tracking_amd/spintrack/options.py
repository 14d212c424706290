"""The reference's parameter block and its ``params.txt`` grammar.

* ``default_options()``   -- include/options.h:11-60 defaults.
* ``parse_params(text)``  -- src/optionsParser.cpp:40-197: one ``key, v1[, v2, v3]``
  per line, whitespace removed, unknown keys ignored with a message, a line
  with fewer than two elements ignored with a message, booleans read with
  ``std::boolalpha`` after lower-casing (anything but "true" is false).
* ``derive(opt)``         -- src/simulation.cpp:9-11: tc = 1.6e-4*m/(k*T^8) and
  sqrtKT_m = sqrt(k*T/m), recomputed on every run.  ``tc_override`` replaces the
  formula's tc when given (the reference has no such key).
"""
from __future__ import annotations

import ctypes as C
import math
import sys
from typing import Iterable

from ._lib import Coords, Options

K_BOLTZMANN = 1.380649e-23   # src/simulation.cpp:9

# Per-species constants the reference's Python helper writes into params
# files (scripts/nedmSim.py:109-149).
NEUTRON = dict(gamma=-1.832471850e8, m=1.674927e-27)
HELIUM3 = dict(gamma=-2.037947093e8, m=1.20200437865e-26)


def default_options() -> Options:
    o = Options()
    o.B0 = Coords(3e-6, 0.0, 0.0)
    o.E = Coords(75e5, 0.0, 0.0)
    o.L = Coords(0.07, 0.1, 0.4)
    o.yi = Coords(0.0, 0.0, 1.0)
    o.Gx = Coords(0.0, 0.0, 0.0)
    o.Gy = Coords(0.0, 0.0, 0.0)
    o.Gz = Coords(0.0, 0.0, 0.0)
    o.m = 1.20239e-26
    o.t0 = 0.0
    o.tf = 10.0
    o.rtol = 1e-12
    o.atol = 1e-12
    o.beta = 0.0
    o.uround = 1e-16
    o.safe = 0.9
    o.fac1 = 0.333
    o.fac2 = 6.0
    o.hmax = 1.0
    o.hmin = 1.0e-8
    o.h = 0.001
    o.T = 4.2
    o.sqrtKT_m = 0.0
    o.tc = 0.0
    o.gamma = -2.038e8
    o.V = 5.0
    o.a = 0.0
    o.w = 0.0
    o.swapStepSize = 1.0 - 4     # sic, include/options.h:39
    o.maxPosStep = 0.1
    o.ioutInt = 0.05
    o.diffuse = 1.0
    o.nmax = 10000000
    o.seed = 0
    o.integratorType = 0
    o.numParticles = 1000
    o.numPerGPUBlock = 128
    o.iout = 2
    o.dist = b"C"
    o.gas_coll = 1
    o.gravity = 1
    o.fixedStepSize = 0
    o.keepStepSize = 0
    o.noiseAmplitudes = Coords(0.0, 0.0, 0.0)
    o.noiseFrequencies = Coords(0.0, 0.0, 0.0)
    return o


def _to_bool(s: str) -> int:
    # std::istringstream >> std::boolalpha after tolower (src/optionsParser.cpp:16-22)
    return 1 if s.lower() == "true" else 0


def _stoi(s: str) -> int:
    # std::stoi: leading integer prefix, throws on none
    t = s.strip()
    j = 0
    if j < len(t) and t[j] in "+-":
        j += 1
    while j < len(t) and t[j].isdigit():
        j += 1
    if j == 0 or (j == 1 and t[0] in "+-"):
        raise ValueError(f"stoi: no conversion for {s!r}")
    return int(t[:j])


def _stod(s: str) -> float:
    # std::stod: longest valid floating prefix
    t = s.strip()
    for end in range(len(t), 0, -1):
        try:
            return float(t[:end])
        except ValueError:
            continue
    raise ValueError(f"stod: no conversion for {s!r}")


_COORD_KEYS = {"L", "yi", "Gx", "Gy", "Gz", "B0", "E"}
_DOUBLE_KEYS = {"m", "T", "diffuse", "gamma", "V", "a", "w", "t0", "tf", "rtol", "atol", "beta",
                "uround", "safe", "fac1", "fac2", "hmax", "hmin", "h", "maxPosStep",
                "swapStepSize", "ioutInt"}
_BOOL_KEYS = {"gas_coll", "gravity", "fixedStepSize", "keepStepSize"}
_INT_KEYS = {"nmax", "seed", "integratorType", "numParticles", "numPerGPUBlock", "iout"}


def parse_line(opt: Options, line: str, log=sys.stdout) -> None:
    """src/optionsParser.cpp:40-175"""
    parts = ["".join(p.split()) for p in line.split(",")]
    if len(parts) < 2:
        print("Misunderstood input line in options file", file=log)
        print("\t" + line, file=log)
        return
    key = parts[0]
    if key in _COORD_KEYS:
        setattr(opt, key, Coords(_stod(parts[1]), _stod(parts[2]), _stod(parts[3])))
    elif key == "dist":
        opt.dist = parts[1][:1].encode() if parts[1] else b"\0"
    elif key in _DOUBLE_KEYS:
        setattr(opt, key, _stod(parts[1]))
    elif key in _BOOL_KEYS:
        setattr(opt, key, _to_bool(parts[1]))
    elif key in _INT_KEYS:
        v = _stoi(parts[1])
        if key in ("nmax", "seed"):
            v &= 0xFFFFFFFF
        setattr(opt, key, v)
    else:
        print(f"Unrecognized Option: {key}", file=log)
        print("Line Ignored", file=log)


def parse_params(text: str, log=sys.stdout) -> Options:
    """src/optionsParser.cpp:177-197 on the text of a params file."""
    opt = default_options()
    for line in text.splitlines():
        parse_line(opt, line, log)
    return opt


def read_params(path: str, log=sys.stdout) -> Options:
    """A missing file yields the defaults, as in the reference."""
    try:
        with open(path) as f:
            return parse_params(f.read(), log)
    except OSError:
        return default_options()


def derive(opt: Options, tc_override: float | None = None) -> Options:
    """src/simulation.cpp:9-11 (in place; returns opt)."""
    opt.tc = 1.6e-4 * opt.m / (K_BOLTZMANN * math.pow(opt.T, 8.0))
    opt.sqrtKT_m = math.sqrt(K_BOLTZMANN * opt.T / opt.m)
    if tc_override is not None:
        opt.tc = float(tc_override)
    return opt


def to_bytes(opt: Options) -> bytes:
    return C.string_at(C.addressof(opt), C.sizeof(opt))


def from_bytes(b: bytes) -> Options:
    if len(b) != C.sizeof(Options):
        raise ValueError("options block must be 440 bytes")
    return Options.from_buffer_copy(b)


def n_intervals(opt: Options) -> int:
    """int(floor((tf - t0)/ioutInt)) output intervals (src/simulation.cpp:79)."""
    return int(math.floor(float(opt.tf - opt.t0) / opt.ioutInt))


def output_times(opt: Options) -> list:
    """nextTime = (i+1)*ioutInt -- ignores t0 (src/simulation.cpp:85)."""
    return [(float(i) + 1.0) * opt.ioutInt for i in range(n_intervals(opt))]


def set_fields(opt: Options, **kw) -> Options:
    """Convenience: set fields by name (tuples for coords, str for dist)."""
    for k, v in kw.items():
        if k in _COORD_KEYS or k in ("noiseAmplitudes", "noiseFrequencies"):
            setattr(opt, k, Coords(*map(float, v)))
        elif k == "dist":
            opt.dist = v.encode() if isinstance(v, str) else v
        elif k in _BOOL_KEYS:
            setattr(opt, k, 1 if v else 0)
        else:
            setattr(opt, k, v)
    return opt


def iter_fields(opt: Options) -> Iterable:
    for name, _ in Options._fields_:
        yield name, getattr(opt, name)
