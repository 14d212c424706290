"""Reader of the reference's HDF5 output layout (src/logger.cpp:40-188),
through the HDF5 C library (ctypes; no h5py needed).

The file holds every option as a scalar attribute of the root group and one
[nsave, numParticles] dataset per snapshot field ("Time", "Position",
"Velocity", "Spin", "Failure State", "Num Collisions", "Num Bounces", "Num
Steps"), each with an int attribute "nt" = rows written.  ``read_hdf5``
returns the same (options, arrays) pair as ``logger.read_output`` does for
the binary format.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .options import Options, default_options

DEFAULT_LIB = "/opt/conda/lib/libhdf5.so"

# dataset name -> (snapshot field, numpy dtype of one element)
DATASETS = {"Time": ("t", np.dtype("<f8")),
            "Position": ("pos", np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8")])),
            "Velocity": ("v", np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8")])),
            "Spin": ("S", np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8")])),
            "Failure State": ("failure", np.dtype("<i4")),
            "Num Collisions": ("n_coll", np.dtype("<u8")),
            "Num Bounces": ("n_bounce", np.dtype("<u8")),
            "Num Steps": ("n_steps", np.dtype("<u8"))}
COORD_ATTRS = {"B0": "B0", "E": "E", "L": "L", "yi": "yi", "Gx": "Gx", "Gy": "Gy", "Gz": "Gz",
               "testNoiseAmp": "noiseAmplitudes", "testNoiseFreq": "noiseFrequencies"}
DOUBLE_ATTRS = ["m", "t0", "tf", "rtol", "atol", "beta", "uround", "safe", "fac1", "fac2", "hmax", "hmin",
                "h", "T", "sqrtKT_m", "tc", "gamma", "V", "a", "w", "swapStepSize", "maxPosStep", "ioutInt",
                "diffuse"]
INT_ATTRS = ["integratorType", "numParticles", "numPerGPUBlock", "iout"]
BOOL_ATTRS = ["gas_coll", "gravity", "fixedStepSize", "keepStepSize"]
UINT_ATTRS = ["nmax", "seed"]

_H5T_COMPOUND = 6
_hid = C.c_int64


class _Lib:
    def __init__(self, path: str):
        L = C.CDLL(path)
        for name, res, args in (
                ("H5open", C.c_int, []),
                ("H5Fopen", _hid, [C.c_char_p, C.c_uint, _hid]),
                ("H5Fclose", C.c_int, [_hid]),
                ("H5Dopen2", _hid, [_hid, C.c_char_p, _hid]),
                ("H5Dread", C.c_int, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
                ("H5Dget_space", _hid, [_hid]),
                ("H5Dclose", C.c_int, [_hid]),
                ("H5Sget_simple_extent_dims", C.c_int, [_hid, C.POINTER(C.c_ulonglong), C.c_void_p]),
                ("H5Sclose", C.c_int, [_hid]),
                ("H5Aopen", _hid, [_hid, C.c_char_p, _hid]),
                ("H5Aread", C.c_int, [_hid, _hid, C.c_void_p]),
                ("H5Aclose", C.c_int, [_hid]),
                ("H5Aexists", C.c_int, [_hid, C.c_char_p]),
                ("H5Tcreate", _hid, [C.c_int, C.c_size_t]),
                ("H5Tinsert", C.c_int, [_hid, C.c_char_p, C.c_size_t, _hid]),
                ("H5Tclose", C.c_int, [_hid])):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        if L.H5open() < 0:
            raise RuntimeError("H5open failed")
        self.L = L
        g = lambda n: _hid.in_dll(L, n).value
        self.double, self.int, self.uint = g("H5T_NATIVE_DOUBLE_g"), g("H5T_NATIVE_INT_g"), g("H5T_NATIVE_UINT_g")
        self.uchar, self.ulong = g("H5T_NATIVE_UCHAR_g"), g("H5T_NATIVE_ULONG_g")
        self.coords = L.H5Tcreate(_H5T_COMPOUND, 24)
        for i, m in enumerate(b"xyz"):
            L.H5Tinsert(self.coords, bytes([m]), 8 * i, self.double)


_lib = None


def hdf5_lib() -> _Lib:
    """The HDF5 C library ($SPINTRACK_HDF5_LIB, else /opt/conda/lib/libhdf5.so)."""
    global _lib
    if _lib is None:
        _lib = _Lib(os.environ.get("SPINTRACK_HDF5_LIB") or DEFAULT_LIB)
    return _lib


def hdf5_available() -> bool:
    try:
        hdf5_lib()
        return True
    except OSError:
        return False


def _attr(H, loc, name: str, htype, np_dtype):
    a = H.L.H5Aopen(loc, name.encode(), 0)
    if a < 0:
        raise KeyError(name)
    buf = np.zeros(1, dtype=np_dtype)
    rc = H.L.H5Aread(a, htype, buf.ctypes.data)
    H.L.H5Aclose(a)
    if rc < 0:
        raise RuntimeError(f"reading attribute {name}")
    return buf[0]


def read_hdf5(path: str):
    """Returns (options, dict of arrays shaped (rows written, N)); the
    options are rebuilt from the root attributes.  A spectral run's b_end
    root attribute comes back under "b_end"."""
    H = hdf5_lib()
    f = H.L.H5Fopen(os.fsencode(path), 0, 0)   # H5F_ACC_RDONLY, H5P_DEFAULT
    if f < 0:
        raise OSError(f"cannot open {path} as HDF5")
    try:
        opt = default_options()
        c3 = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8")])
        for name, field in COORD_ATTRS.items():
            v = _attr(H, f, name, H.coords, c3)
            c = getattr(opt, field)
            c.x, c.y, c.z = float(v["x"]), float(v["y"]), float(v["z"])
        for name in DOUBLE_ATTRS:
            setattr(opt, name, float(_attr(H, f, name, H.double, "<f8")))
        for name in INT_ATTRS:
            setattr(opt, name, int(_attr(H, f, name, H.int, "<i4")))
        for name in BOOL_ATTRS:
            setattr(opt, name, bool(_attr(H, f, name, H.uchar, "u1")))
        opt.dist = bytes([int(_attr(H, f, "dist", H.uchar, "u1"))])
        for name in UINT_ATTRS:
            setattr(opt, name, int(_attr(H, f, name, H.uint, "<u4")))
        out = {}
        for dname, (field, dt) in DATASETS.items():
            d = H.L.H5Dopen2(f, dname.encode(), 0)
            if d < 0:
                raise KeyError(dname)
            sp = H.L.H5Dget_space(d)
            dims = (C.c_ulonglong * 2)()
            H.L.H5Sget_simple_extent_dims(sp, dims, None)
            H.L.H5Sclose(sp)
            arr = np.zeros((dims[0], dims[1]), dtype=dt)
            htype = {"t": H.double, "failure": H.int}.get(field, H.ulong if field.startswith("n_") else H.coords)
            rc = H.L.H5Dread(d, htype, 0, 0, 0, arr.ctypes.data)
            nt = int(_attr(H, d, "nt", H.int, "<i4"))
            H.L.H5Dclose(d)
            if rc < 0:
                raise RuntimeError(f"reading {dname}")
            out[field] = arr[:nt]
        # a spectral run's final Bloch vector (writeSingle, src/logger.cpp:68-70)
        if H.L.H5Aexists(f, b"b_end") > 0:
            v = _attr(H, f, "b_end", H.coords, c3)
            out["b_end"] = np.array([float(v["x"]), float(v["y"]), float(v["z"])])
        return opt, out
    finally:
        H.L.H5Fclose(f)
