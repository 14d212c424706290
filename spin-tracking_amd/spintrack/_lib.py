"""ctypes binding of the gfx950 library ``libspintrack.so`` (include/spintrack.h).

The library is built in-tree by ``spin-tracking_amd/Makefile``.  Loading fails
loudly when it is missing: there is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libspintrack.so")

ST_OK = 0
ST_E_ARG = -1
ST_E_HIP = -2
ST_E_UNSUPPORTED = -3
ST_E_STATE = -4

FAIL_WATCHDOG = -9            # failureState of a particle stopped by the event watchdog
DEFAULT_MAX_EVENTS = 100000000
MAX_PARTICLES = 1 << 29       # particles per handle (32-bit lane offsets)
NOALLOC_MAX_TARGETS = 4096    # output times per st_run_noalloc call

VARIANT_EXACT = 0
VARIANT_FAST = 1


class Coords(C.Structure):
    """include/double3.h:17-21"""
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]

    def __iter__(self):
        return iter((self.x, self.y, self.z))

    def __repr__(self):
        return f"({self.x!r}, {self.y!r}, {self.z!r})"


_DOUBLE_FIELDS = ["m", "t0", "tf", "rtol", "atol", "beta", "uround", "safe", "fac1", "fac2",
                  "hmax", "hmin", "h", "T", "sqrtKT_m", "tc", "gamma", "V", "a", "w",
                  "swapStepSize", "maxPosStep", "ioutInt", "diffuse"]
_COORD_FIELDS = ["B0", "E", "L", "yi", "Gx", "Gy", "Gz"]


class Options(C.Structure):
    """The reference's 440-byte ``options`` block (include/options.h:10-61)."""
    _fields_ = ([(n, Coords) for n in _COORD_FIELDS]
                + [(n, C.c_double) for n in _DOUBLE_FIELDS]
                + [("nmax", C.c_uint32), ("seed", C.c_uint32), ("integratorType", C.c_int32),
                   ("numParticles", C.c_int32), ("numPerGPUBlock", C.c_int32), ("iout", C.c_int32),
                   ("dist", C.c_char), ("gas_coll", C.c_uint8), ("gravity", C.c_uint8),
                   ("fixedStepSize", C.c_uint8), ("keepStepSize", C.c_uint8),
                   ("noiseAmplitudes", Coords), ("noiseFrequencies", Coords)])


assert C.sizeof(Options) == 440, C.sizeof(Options)


ENGINE_PIPELINE = 0
ENGINE_FUSED = 1


class Ext(C.Structure):
    _fields_ = [("pulse_axis", C.c_int32), ("variant", C.c_int32), ("block_size", C.c_int32),
                ("engine", C.c_int32), ("seg_buffer_mb", C.c_int32), ("chunk_records", C.c_int32),
                ("graph_chunks", C.c_int32), ("spectrum_sum", C.c_int32), ("max_events", C.c_uint32),
                ("adjoint_window", C.c_uint32)]


class Rng(C.Structure):
    """include/particle.h:52-59"""
    _fields_ = [("x", C.c_uint64), ("y", C.c_uint64), ("z", C.c_uint64), ("w", C.c_uint64),
                ("spare", C.c_double), ("hasSpare", C.c_uint8), ("pad_", C.c_uint8 * 7)]


assert C.sizeof(Rng) == 48


class ParticleState(C.Structure):
    _fields_ = [("S", Coords), ("v", Coords), ("v_old", Coords), ("pos", Coords),
                ("pos_old", Coords), ("t", C.c_double), ("t_old", C.c_double),
                ("tf", C.c_double), ("next_gas", C.c_double), ("h", C.c_double),
                ("rng", Rng), ("n_bounce", C.c_uint64), ("n_coll", C.c_uint64),
                ("n_steps", C.c_uint64), ("failure", C.c_int32), ("part_id", C.c_uint32),
                ("stop", C.c_uint8), ("coll_type", C.c_uint8), ("wall_hit", C.c_uint8),
                ("pad_", C.c_uint8 * 5)]


NW = 17              # include/floquet.h:34
NK = 11              # include/floquet.h:33
FLOQUET_NPROP = 100
AGG_FIELDS = 9 * NW  # s1[NW][3], s2[NW][3], imag_diag[NW][3] per particle


class Floquet(C.Structure):
    """floquetDiagonalization (include/floquet.h:99-106)."""
    _fields_ = [("f_modes_0", C.c_double * 4), ("f_energies", C.c_double * 4),
                ("frequencies", C.c_double * NW), ("dt", C.c_double), ("n_prop", C.c_int32),
                ("pad_", C.c_int32), ("propagators", (C.c_double * 4) * FLOQUET_NPROP)]


class Spectrum(C.Structure):
    """CovarianceSpectrum (include/floquet.h:45-62), un-normalised."""
    _fields_ = [("frequencies", C.c_double * NW), ("dt", C.c_double), ("n_samples", C.c_uint64),
                ("variance", C.c_double * (NW * 18))]


# every symbol include/spintrack.h declares, with its ctypes signature
_vp, _d, _i32, _u32, _sz = C.c_void_p, C.c_double, C.c_int32, C.c_uint32, C.c_size_t
SIGNATURES = {
    "st_version": (C.c_char_p, []),
    "st_last_error": (C.c_char_p, [_vp]),
    "st_create": (C.c_int, [C.POINTER(Options), C.POINTER(Ext), C.POINTER(_vp)]),
    "st_create_batch": (C.c_int, [_vp, _i32, C.POINTER(Ext), C.POINTER(_vp)]),
    "st_batch_layout": (C.c_int, [_vp, C.POINTER(_i32), _vp]),
    "st_destroy": (None, [_vp]),
    "st_set_stream": (C.c_int, [_vp, _vp]),
    "st_init": (C.c_int, [_vp]),
    "st_init_subset": (C.c_int, [_vp, _u32, _u32]),
    "st_advance_to": (C.c_int, [_vp, _d]),
    "st_snapshot": (C.c_int, [_vp] + [_vp] * 8),
    "st_snapshot_block_bytes": (_sz, [_vp]),
    "st_snapshot_block_device": (C.c_int, [_vp, _vp]),
    "st_run": (C.c_int, [_vp, _i32, _vp, _vp]),
    "st_run_noalloc": (C.c_int, [_vp, _i32, _vp, _vp]),
    "st_set_stop_times": (C.c_int, [_vp, _vp]),
    "st_get_state": (C.c_int, [_vp, _vp]),
    "st_set_state": (C.c_int, [_vp, _vp]),
    "st_work_counters": (C.c_int, [_vp, _vp]),
    "st_particle_counters": (C.c_int, [_vp, _vp, _vp, _vp]),
    "st_synchronize": (C.c_int, [_vp]),
    "st_set_trace": (C.c_int, [_vp, C.c_int64, _u32]),
    "st_get_trace": (C.c_int, [_vp, _vp, _u32, C.POINTER(_u32)]),
    "st_device_alloc": (C.c_int, [_sz, C.POINTER(_vp)]),
    "st_device_free": (C.c_int, [_vp]),
    "st_host_alloc": (C.c_int, [_sz, C.POINTER(_vp)]),
    "st_host_free": (C.c_int, [_vp]),
    "st_memcpy_d2h": (C.c_int, [_vp, _vp, _vp, _sz]),
    "st_memcpy_d2d": (C.c_int, [_vp, _vp, _vp, _sz]),
    "st_timer_start": (C.c_int, [_vp]),
    "st_timer_stop": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "st_kernel_timing": (C.c_int, [_vp, C.c_int]),
    "st_kernel_times": (C.c_int, [_vp, _vp, _vp]),
    "st_set_checkpointing": (C.c_int, [_vp, C.c_int]),
    "st_adjoint": (C.c_int, [_vp, _vp, _vp, _vp]),
    "st_block_sums": (C.c_int, [_vp, _vp, _i32, _vp]),
    "st_phase_stats": (C.c_int, [_vp, _vp, _i32, _vp, _vp]),
    "st_spectrum_get": (C.c_int, [_vp, C.POINTER(Spectrum)]),
    "st_spectrum_aggregators": (C.c_int, [_vp, _vp, _vp]),
    "st_floquet_info": (C.c_int, [_vp, C.POINTER(Floquet)]),
    "st_floquet_result": (C.c_int, [_vp, _vp]),
    "st_floquet_diagonalize": (C.c_int, [C.POINTER(Options), C.POINTER(Ext), C.POINTER(Floquet)]),
    "st_floquet_rates": (C.c_int, [_vp, _vp, _vp, _vp, _i32, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "st_floquet_markov": (C.c_int, [_d, _d, _vp, _vp, _vp]),
    "st_floquet_integrate": (C.c_int, [C.POINTER(Floquet), C.POINTER(Spectrum), C.POINTER(Options), _vp]),
}

# parameter order of st_adjoint's gradient (15 per configuration)
GRAD_PARAMS = ["Gx.x", "Gx.y", "Gx.z", "Gy.x", "Gy.y", "Gy.z", "Gz.x", "Gz.y", "Gz.z",
               "E.x", "E.y", "E.z", "B0.x", "B0.y", "B0.z"]


class DeviceBuffer:
    """Device memory owned through the library (its own HIP runtime)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(load().st_device_alloc(self.nbytes, C.byref(p)))
        self.ptr = int(p.value)

    def free(self) -> None:
        if getattr(self, "ptr", 0):
            load().st_device_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

TRACE_DOUBLES = 20
TRACE_FIELDS = ["tag", "t_old", "t", "px", "py", "pz", "vx", "vy", "vz", "sx", "sy", "sz", "h",
                "accepted", "rejected", "coll_type", "wall_hit", "status", "n_bounce", "n_coll"]

_libs = {}


class SpinTrackError(RuntimeError):
    pass


def open_library(path: str) -> C.CDLL:
    """Loads a build of the library (e.g. one compiled with other field hooks,
    spin-tracking_amd/Makefile FIELDS=... LIBNAME=...); raises if missing."""
    if path not in _libs:
        if not os.path.exists(path):
            raise SpinTrackError(
                f"{path} is missing: build it with `make -C spin-tracking_amd` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _libs[path] = L
    return _libs[path]


def load() -> C.CDLL:
    """Loads libspintrack.so; raises if it has not been built."""
    return open_library(LIB_PATH)


def check(rc: int, handle=None, lib: C.CDLL = None) -> None:
    if rc != ST_OK:
        msg = (lib or load()).st_last_error(handle)
        raise SpinTrackError(f"spintrack error {rc}: {msg.decode() if msg else ''}")
