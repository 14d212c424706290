"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's cpu_baseline leg, always as the checker and never as the
thing measured or shipped.  The oracle restates the reference CPU path
(src/particle.cpp:24-837, src/integrator.cpp:20-723, src/quaternion.cpp:128-164 of the reference
repository) in plain C; see st_oracle.h for the line-by-line map.

The structured numpy dtypes below mirror the C structs byte for byte:
``OPTIONS_DTYPE`` is the reference's 440-byte ``options`` block
(include/options.h:10-61) and ``RNG_DTYPE`` its 48-byte ``rngState``
(include/particle.h:52-59).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

COORDS = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8")])

OPTIONS_DTYPE = np.dtype({
    "names": ["B0", "E", "L", "yi", "Gx", "Gy", "Gz",
              "m", "t0", "tf", "rtol", "atol", "beta", "uround", "safe", "fac1",
              "fac2", "hmax", "hmin", "h", "T", "sqrtKT_m", "tc", "gamma", "V",
              "a", "w", "swapStepSize", "maxPosStep", "ioutInt", "diffuse",
              "nmax", "seed", "integratorType", "numParticles", "numPerGPUBlock",
              "iout", "dist", "gas_coll", "gravity", "fixedStepSize",
              "keepStepSize", "noiseAmplitudes", "noiseFrequencies"],
    "formats": [("<f8", 3)] * 7 + ["<f8"] * 24 + ["<u4", "<u4"] + ["<i4"] * 4
               + ["S1", "?", "?", "?", "?"] + [("<f8", 3)] * 2,
    "offsets": [0, 24, 48, 72, 96, 120, 144] + [168 + 8 * i for i in range(24)]
               + [360, 364, 368, 372, 376, 380, 384, 385, 386, 387, 388, 392, 416],
    "itemsize": 440,
})

RNG_DTYPE = np.dtype({
    "names": ["x", "y", "z", "w", "spare", "hasSpare"],
    "formats": ["<u8", "<u8", "<u8", "<u8", "<f8", "?"],
    "offsets": [0, 8, 16, 24, 32, 40],
    "itemsize": 48,
})


def default_options() -> np.ndarray:
    """The reference defaults (include/options.h:11-60) as a 0-d record."""
    o = np.zeros((), dtype=OPTIONS_DTYPE)
    o["B0"] = (3e-6, 0.0, 0.0)
    o["E"] = (75e5, 0.0, 0.0)
    o["L"] = (0.07, 0.1, 0.4)
    o["yi"] = (0.0, 0.0, 1.0)
    for k, v in dict(m=1.20239e-26, t0=0.0, tf=10.0, rtol=1e-12, atol=1e-12,
                     beta=0.0, uround=1e-16, safe=0.9, fac1=0.333, fac2=6.0,
                     hmax=1.0, hmin=1.0e-8, h=0.001, T=4.2, sqrtKT_m=0.0, tc=0.0,
                     gamma=-2.038e8, V=5.0, a=0.0, w=0.0, swapStepSize=1.0 - 4,
                     maxPosStep=0.1, ioutInt=0.05, diffuse=1.0).items():
        o[k] = v
    o["nmax"] = 10000000
    o["seed"] = 0
    o["integratorType"] = 0
    o["numParticles"] = 1000
    o["numPerGPUBlock"] = 128
    o["iout"] = 2
    o["dist"] = b"C"
    o["gas_coll"] = True
    o["gravity"] = True
    o["fixedStepSize"] = False
    o["keepStepSize"] = False
    return o


# SpectrumAggregator state of the C oracle (or_specagg)
NW = 17
SPECAGG_DTYPE = np.dtype({
    "names": ["s1", "s2", "im", "n_samples"],
    "formats": [("<f8", (NW, 3)), ("<f8", (NW, 3)), ("<f8", (NW, 3)), "<i4"],
    "offsets": [0, NW * 24, 2 * NW * 24, 3 * NW * 24],
    "itemsize": 3 * NW * 24 + 8,
})


class _Coords(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class _Ext(C.Structure):
    _fields_ = [("pulse_axis", C.c_int)]


class _SpecCfg(C.Structure):
    _fields_ = [("w", C.c_double * NW), ("dt", C.c_double)]


class _Arrays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ["S", "v", "v_old", "pos", "pos_old", "t", "t_old", "tf", "dt",
                 "next_gas", "h", "state", "n_bounce", "n_coll", "n_steps",
                 "partID", "failure", "stop", "coll_type", "wall_hit"]]


ARRAY_FIELDS = {
    "S": COORDS, "v": COORDS, "v_old": COORDS, "pos": COORDS, "pos_old": COORDS,
    "t": np.float64, "t_old": np.float64, "tf": np.float64, "dt": np.float64,
    "next_gas": np.float64, "h": np.float64, "state": RNG_DTYPE,
    "n_bounce": np.uint64, "n_coll": np.uint64, "n_steps": np.uint64,
    "partID": np.uint32, "failure": np.int32, "stop": np.bool_,
    "coll_type": np.uint8, "wall_hit": np.uint8,
}


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}
_current = LIB_PATH

# integratorType -> oracle function
INTEGRATORS = ("or_integrate_dop", "or_integrate_rk45", "or_integrate_magnus",
               "or_integrate_rk45q", "or_integrate_rkf45", "or_integrate_rkf45q")


@contextlib.contextmanager
def library(path: str):
    """Runs the oracle functions (and Ensembles created) inside the block
    with another build of the oracle, e.g. one compiled with other field
    hooks (oracle/Makefile FIELDS=...)."""
    global _current
    old, _current = _current, path
    try:
        yield lib()
    finally:
        _current = old


def lib():
    path = _current
    if path not in _libs:
        if not os.path.exists(path):
            if path != LIB_PATH:
                raise FileNotFoundError(path)
            build()
        L = C.CDLL(path)
        vp, d, u32, i32 = C.c_void_p, C.c_double, C.c_uint32, C.c_int
        L.or_xoshiro.restype = C.c_uint64
        L.or_xoshiro.argtypes = [vp]
        L.or_uniform.restype = d
        L.or_uniform.argtypes = [vp]
        L.or_normal.restype = d
        L.or_normal.argtypes = [vp, d, d]
        L.or_maxboltz.restype = d
        L.or_maxboltz.argtypes = [vp, d]
        L.or_exponential.restype = d
        L.or_exponential.argtypes = [vp, d]
        L.or_seed.restype = None
        L.or_seed.argtypes = [vp, u32, u32]
        L.or_calc_next_collision_time.restype = None
        L.or_calc_next_collision_time.argtypes = [d, d, _Coords, _Coords, vp, vp, vp, vp,
                                                  vp, vp, vp, vp, vp]
        L.or_update_position_and_velocity.restype = None
        L.or_update_position_and_velocity.argtypes = [vp] * 12
        L.or_sanity_check.restype = None
        L.or_sanity_check.argtypes = [vp] * 9
        L.or_interpolate.restype = None
        L.or_interpolate.argtypes = [d, d, _Coords, _Coords, vp, vp, vp]
        L.or_cross_term.restype = _Coords
        L.or_cross_term.argtypes = [d, vp, vp, d, _Coords, _Coords]
        L.or_bloch.restype = _Coords
        L.or_bloch.argtypes = [d, _Coords, vp, vp, d, _Coords, _Coords]
        for nm in INTEGRATORS:
            f = getattr(L, nm)
            f.restype = i32
            f.argtypes = [d, d, vp, _Coords, _Coords, vp, vp, vp, vp, vp]
        L.or_init_particles.restype = None
        L.or_init_particles.argtypes = [vp, vp, i32, u32, u32]
        L.or_run_simulation.restype = None
        L.or_run_simulation.argtypes = [vp, vp, vp, i32, d, vp, d, vp, i32]
        L.or_run_simulation_spec.restype = None
        L.or_run_simulation_spec.argtypes = [vp, vp, vp, i32, d, vp, d, vp, i32, vp, vp]
        L.or_noise_term.restype = _Coords
        L.or_noise_term.argtypes = [d, vp, d, _Coords, _Coords]
        L.or_spectrum_update.restype = None
        L.or_spectrum_update.argtypes = [vp, vp, _Coords]
        L.or_integrate_spectrum.restype = i32
        L.or_integrate_spectrum.argtypes = [d, d, vp, vp, _Coords, _Coords, vp, d]
        L.or_covariances.restype = None
        L.or_covariances.argtypes = [vp, i32, vp, vp]
        L.or_integrate_hamiltonian.restype = i32
        L.or_integrate_hamiltonian.argtypes = [d, d, vp, vp, vp, d]
        L.or_sizeof_specagg.restype = i32
        L.or_sizeof_options.restype = i32
        L.or_sizeof_rng.restype = i32
        _libs[path] = L
    return _libs[path]


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _c3(v) -> _Coords:
    return _Coords(float(v[0]), float(v[1]), float(v[2]))


def _ext(pulse_axis: int = 2) -> _Ext:
    return _Ext(int(pulse_axis))


# ---------------------------------------------------------------- RNG
def rng_seeded(seed: int, index: int) -> np.ndarray:
    s = np.zeros((), dtype=RNG_DTYPE)
    lib().or_seed(_ptr(s), seed, index)
    return s


def xoshiro(state: np.ndarray, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    f = lib().or_xoshiro
    p = _ptr(state)
    for i in range(n):
        out[i] = f(p)
    return out


def uniform(state: np.ndarray, n: int) -> np.ndarray:
    f = lib().or_uniform
    p = _ptr(state)
    return np.array([f(p) for _ in range(n)])


def normal(state: np.ndarray, n: int, mean=0.0, std=1.0) -> np.ndarray:
    f = lib().or_normal
    p = _ptr(state)
    return np.array([f(p, mean, std) for _ in range(n)])


def exponential(state: np.ndarray, n: int, tc: float) -> np.ndarray:
    f = lib().or_exponential
    p = _ptr(state)
    return np.array([f(p, tc) for _ in range(n)])


# ---------------------------------------------------------------- kinematics
def calc_next_collision_time(opt, t, tf, v, pos, next_gas, state, n_bounce=0, n_coll=0):
    ng = np.array([next_gas], dtype=np.float64)
    dt = np.zeros(1)
    ct = np.zeros(1, dtype=np.uint8)
    nb = np.array([n_bounce], dtype=np.uint64)
    nc = np.array([n_coll], dtype=np.uint64)
    fin = np.zeros(1, dtype=np.bool_)
    wh = np.zeros(1, dtype=np.uint8)
    lib().or_calc_next_collision_time(t, tf, _c3(v), _c3(pos), _ptr(ng), _ptr(dt), _ptr(ct),
                                      _ptr(nb), _ptr(nc), _ptr(fin), _ptr(wh), _ptr(state),
                                      _ptr(opt))
    return dict(next_gas=ng[0], dt=dt[0], coll_type=chr(ct[0]), n_bounce=int(nb[0]),
                n_coll=int(nc[0]), finished=bool(fin[0]), wall_hit=chr(wh[0]))


def update_position_and_velocity(opt, t, dt, pos, v, coll_type, wall_hit, state):
    t_old = np.zeros(1)
    tt = np.array([t], dtype=np.float64)
    d = np.array([dt], dtype=np.float64)
    po = np.zeros(3)
    p = np.array(pos, dtype=np.float64)
    vo = np.zeros(3)
    vv = np.array(v, dtype=np.float64)
    ct = np.array([ord(coll_type)], dtype=np.uint8)
    wh = np.array([ord(wall_hit)], dtype=np.uint8)
    stop = np.zeros(1, dtype=np.bool_)
    lib().or_update_position_and_velocity(_ptr(t_old), _ptr(tt), _ptr(d), _ptr(po), _ptr(p),
                                          _ptr(vo), _ptr(vv), _ptr(ct), _ptr(wh), _ptr(state),
                                          _ptr(stop), _ptr(opt))
    return dict(t_old=t_old[0], t=tt[0], pos_old=po, pos=p, v_old=vo, v=vv, stop=bool(stop[0]))


def sanity_check(opt, t_old, t, pos_old, pos, v, v_old):
    a = [np.array([t_old], dtype=np.float64), np.array([t], dtype=np.float64),
         np.array(pos_old, dtype=np.float64), np.array(pos, dtype=np.float64),
         np.array(v, dtype=np.float64), np.array(v_old, dtype=np.float64)]
    stop = np.zeros(1, dtype=np.bool_)
    fail = np.zeros(1, dtype=np.int32)
    lib().or_sanity_check(*[_ptr(x) for x in a], _ptr(stop), _ptr(fail), _ptr(opt))
    return bool(stop[0]), int(fail[0])


# ---------------------------------------------------------------- spin ODE
def interpolate(opt, t, t0, p_old, v_old):
    p = np.zeros(3)
    v = np.zeros(3)
    lib().or_interpolate(t, t0, _c3(p_old), _c3(v_old), _ptr(p), _ptr(v), _ptr(opt))
    return p, v


def bloch(opt, t, y, t0, p_old, v_old, pulse_axis=2):
    e = _ext(pulse_axis)
    r = lib().or_bloch(t, _c3(y), _ptr(opt), C.addressof(e), t0, _c3(p_old), _c3(v_old))
    return np.array([r.x, r.y, r.z])


def cross_term(opt, t, t0, p_old, v_old, pulse_axis=2):
    """findCrossTerm: gamma * B along the segment (src/integrator.cpp:63-70)."""
    e = _ext(pulse_axis)
    r = lib().or_cross_term(t, _ptr(opt), C.addressof(e), t0, _c3(p_old), _c3(v_old))
    return np.array([r.x, r.y, r.z])


def integrate(opt, t0, tf, y, p_old, v_old, h, integrator=0, pulse_axis=2):
    """One reference integrator call of integratorType ``integrator`` (0
    integrateDOP, 1 integrateRK45, 2 integrateMagnusCFET, 3
    integrateRK45Quaternion, 4 integrateRKF45, 5 integrateRKF45Quaternion).
    Returns (status, y_out, h_out, n_accepted, n_rejected)."""
    yy = np.array(y, dtype=np.float64)
    hh = np.array([h], dtype=np.float64)
    na = np.zeros(1, dtype=np.uint64)
    nr = np.zeros(1, dtype=np.uint64)
    e = _ext(pulse_axis)
    f = getattr(lib(), INTEGRATORS[integrator])
    st = f(t0, tf, _ptr(yy), _c3(p_old), _c3(v_old), _ptr(opt), C.addressof(e), _ptr(hh),
           _ptr(na), _ptr(nr))
    return st, yy, hh[0], int(na[0]), int(nr[0])


# ------------------------------------------------------- spectral path
def spec_cfg(freqs, dt: float) -> _SpecCfg:
    """SpectrumAggregator::initialize(freq, dt) (src/floquet.cpp:150-153)."""
    c = _SpecCfg()
    for i in range(NW):
        c.w[i] = float(freqs[i])
    c.dt = float(dt)
    return c


def noise_term(opt, t, t0, p_old, v_old):
    """findNoiseTerm (src/integrator.cpp:72-80)."""
    r = lib().or_noise_term(t, _ptr(opt), t0, _c3(p_old), _c3(v_old))
    return np.array([r.x, r.y, r.z])


def integrate_spectrum(opt, t0, tf, agg: np.ndarray, cfg: _SpecCfg, p_old, v_old, h) -> int:
    """integrateSpectrum (src/integrator.cpp:730-743) into the 0-d aggregator
    record agg; returns the number of samples."""
    return lib().or_integrate_spectrum(t0, tf, _ptr(agg), C.addressof(cfg), _c3(p_old), _c3(v_old),
                                       _ptr(opt), h)


def covariances(agg: np.ndarray, cfg: _SpecCfg) -> np.ndarray:
    """get_covariance_spectrum of every aggregator: (n, NW, 3, 3, 2) float64,
    the last axis (re, im) (src/floquet.cpp:131-147)."""
    agg = np.ascontiguousarray(agg)
    out = np.zeros((agg.size, NW, 3, 3, 2))
    lib().or_covariances(_ptr(agg), agg.size, C.addressof(cfg), _ptr(out))
    return out


def partial_tree_indices(n: int) -> list:
    """The particles the reference's CPU path sums (src/particle.cpp:826-834):
    its in-place tree over ipart = 0, 1, ... adds cspec_array[s] into
    cspec_array[0] for s = 1, 2, 4, ... < n before any other slot is
    reduced, so 0 and the powers of two below n make up the total."""
    idx = [0]
    s = 1
    while s < n:
        idx.append(s)
        s *= 2
    return idx


def sum_spectra(cov: np.ndarray, n_samp: np.ndarray, mode: int):
    """One interval's spectrum sum, in the reference's order: mode 0, the
    CPU path's partial tree (((c0 + c1) + c2) + c4) ...; mode 1, every
    particle in index order (aggregateSpectrum, src/particle.cpp:893-897)."""
    idx = partial_tree_indices(len(cov)) if mode == 0 else list(range(len(cov)))
    tot = cov[idx[0]].copy()
    tn = int(n_samp[idx[0]])
    for j in idx[1:]:
        tot = tot + cov[j]
        tn += int(n_samp[j])
    return tot, tn


# ---------------------------------------------------------------- ensemble
class Ensemble:
    """The reference's particle arrays for n particles, on the host."""

    def __init__(self, opt: np.ndarray, n: int, first: int = 0, stride: int = 1,
                 pulse_axis: int = 2, spectrum_sum: int = 0):
        self.opt = np.array(opt, dtype=OPTIONS_DTYPE).reshape(())
        self.n = int(n)
        self.arr = {k: np.zeros(self.n, dtype=dt) for k, dt in ARRAY_FIELDS.items()}
        self.n_ode = np.zeros(3 * self.n, dtype=np.uint64)
        self._a = _Arrays(**{k: _ptr(v) for k, v in self.arr.items()})
        self._ext = _ext(pulse_axis)
        self._lib = lib()
        self._lib.or_init_particles(_ptr(self.opt), C.addressof(self._a), self.n, first, stride)
        self.spectral = int(self.opt["integratorType"]) == 6
        if self.spectral:
            # particle::initParticles -> initializeSpectra (include/particle.h:241-242,
            # src/particle.cpp:887-891) and SpectrumAggregator::initialize (:520-521)
            import floquet
            self.fd = floquet.diagonalize(self.opt, pulse_axis)
            self.cfg = spec_cfg(self.fd["frequencies"], self.fd["dt"])
            self.agg = np.zeros(self.n, dtype=SPECAGG_DTYPE)
            self.spectrum_sum = int(spectrum_sum)
            self.cspec = np.zeros((NW, 3, 3, 2))   # CovarianceSpectrum::variance, (re, im)
            self.cspec_n = 0
            self.samples_total = np.zeros(self.n, dtype=np.uint64)   # per particle, all intervals

    def advance_to(self, next_t: float, stop_at=None, prev_t: float = 0.0,
                   nthreads: int = 0) -> None:
        sp = 0
        if stop_at is not None:
            self._stop_at = np.ascontiguousarray(stop_at, dtype=np.float64)
            sp = _ptr(self._stop_at)
        if self.spectral:
            self._lib.or_run_simulation_spec(_ptr(self.opt), C.addressof(self._ext), C.addressof(self._a),
                                         self.n, next_t, sp, prev_t, _ptr(self.n_ode), nthreads,
                                         _ptr(self.agg), C.addressof(self.cfg))
            self.samples_total += self.agg["n_samples"].astype(np.uint64)
            tot, tn = sum_spectra(covariances(self.agg, self.cfg), self.agg["n_samples"], self.spectrum_sum)
            self.cspec = self.cspec + tot
            self.cspec_n += tn
            return
        self._lib.or_run_simulation(_ptr(self.opt), C.addressof(self._ext), C.addressof(self._a),
                                self.n, next_t, sp, prev_t, _ptr(self.n_ode), nthreads)

    def floquet_result(self) -> np.ndarray:
        """particle::floquetResults (src/particle.cpp:841-843): b_end."""
        import floquet
        var = self.cspec[..., 0] + 1j * self.cspec[..., 1]
        return floquet.integrate(self.fd, var, self.cspec_n, self.fd["dt"], self.opt)

    def snapshot(self) -> dict:
        """The eight arrays BinaryLogger::writeSnapshot writes (src/logger.cpp:14-25)."""
        a = self.arr
        return dict(t=a["t"].copy(), pos=a["pos"].copy(), v=a["v"].copy(), S=a["S"].copy(),
                    failure=a["failure"].copy(), n_coll=a["n_coll"].copy(),
                    n_bounce=a["n_bounce"].copy(), n_steps=a["n_steps"].copy())


def xyz(a: np.ndarray) -> np.ndarray:
    """COORDS structured array -> (n, 3) float64."""
    return np.stack([a["x"], a["y"], a["z"]], axis=-1)
