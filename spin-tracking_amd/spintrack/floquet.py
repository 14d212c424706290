"""Host-side Floquet routines of the spectral path (integratorType 6), over
the library's C ABI (include/spintrack.h):

    diagonalize(opt)   floquet_diagonalize     src/integrator.cpp:824-867
    rates(...)         floquet_master_equation_rates   src/floquet.cpp:192-279
    markov(...)        integrateFloquetMarkov  src/integrator.cpp:791-822
    integrate(...)     floquet_integrate       src/integrator.cpp:869-889

A spectral ``Particle`` holds its own decomposition (``floquetInfo``) and
spectrum (``spectrum``) and returns b_end from ``floquetResults``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import Ext, Options, check


def floquet_dict(f: L.Floquet) -> dict:
    n = int(f.n_prop)
    props = np.ctypeslib.as_array(f.propagators).reshape(L.FLOQUET_NPROP, 4)[:n].copy()
    return dict(f_modes_0=np.array(f.f_modes_0[:]), f_energies=np.array(f.f_energies[:]),
                frequencies=np.array(f.frequencies[:]), dt=f.dt, n_prop=n, propagators=props)


def _to_struct(fd: dict) -> L.Floquet:
    f = L.Floquet()
    f.f_modes_0[:] = [float(v) for v in fd["f_modes_0"]]
    f.f_energies[:] = [float(v) for v in fd["f_energies"]]
    f.frequencies[:] = [float(v) for v in fd["frequencies"]]
    f.dt = float(fd["dt"])
    f.n_prop = int(fd["n_prop"])
    p = np.ctypeslib.as_array(f.propagators).reshape(L.FLOQUET_NPROP, 4)
    p[:f.n_prop] = np.asarray(fd["propagators"], dtype=np.float64)
    return f


def diagonalize(opt: Options, pulse_axis: int = 2) -> dict:
    """floquet_diagonalize(opt): propagators over one drive period, Floquet
    modes and quasi-energies, and the NW sampling frequencies."""
    f = L.Floquet()
    ext = Ext(int(pulse_axis))
    check(L.load().st_floquet_diagonalize(C.byref(opt), C.byref(ext), C.byref(f)))
    return floquet_dict(f)


def _c(a, shape) -> np.ndarray:
    z = np.ascontiguousarray(np.asarray(a, dtype=complex).reshape(shape))
    return np.ascontiguousarray(np.stack([z.real, z.imag], axis=-1))


def rates(f_modes_0, f_energies, c_op, propagators, period: float, frequencies, power,
          Zeta: np.ndarray, Omicron: np.ndarray):
    """floquet_master_equation_rates: returns (Delta, X, Gamma) and adds to
    Zeta and Omicron (2x2 complex arrays, updated in place)."""
    props = np.ascontiguousarray(propagators, dtype=np.float64).reshape(-1, 4)
    q = lambda v: np.ascontiguousarray(v, dtype=np.float64)  # noqa: E731
    m0, en, co, fr = q(f_modes_0), q(f_energies), q(c_op), q(frequencies)
    pw = _c(power, (L.NW,))
    Z, O = _c(Zeta, (2, 2)), _c(Omicron, (2, 2))
    D = np.zeros((2, 2, L.NK))
    X = np.zeros((2, 2, L.NK, 2))
    G = np.zeros((2, 2, L.NK, 2))
    p = lambda a: a.ctypes.data  # noqa: E731
    check(L.load().st_floquet_rates(p(m0), p(en), p(co), p(props), len(props), float(period), p(fr), p(pw),
                                    p(D), p(X), p(G), p(Z), p(O)))
    Zeta[...] = Z[..., 0] + 1j * Z[..., 1]
    Omicron[...] = O[..., 0] + 1j * O[..., 1]
    return D, X[..., 0] + 1j * X[..., 1], G[..., 0] + 1j * G[..., 1]


def markov(t0: float, tf: float, rho, Zeta, Omicron) -> np.ndarray:
    """integrateFloquetMarkov: the density matrix at tf."""
    r = _c(rho, (2, 2))
    Z, O = _c(Zeta, (2, 2)), _c(Omicron, (2, 2))
    check(L.load().st_floquet_markov(float(t0), float(tf), r.ctypes.data, Z.ctypes.data, O.ctypes.data))
    return r[..., 0] + 1j * r[..., 1]


def integrate(fd: dict, variance, n_samples: int, dt: float, opt: Options) -> np.ndarray:
    """floquet_integrate of a summed covariance spectrum variance[NW,3,3]
    (complex, un-normalised) over n_samples samples of spacing dt."""
    sp = L.Spectrum()
    sp.frequencies[:] = [float(v) for v in fd["frequencies"]]
    sp.dt = float(dt)
    sp.n_samples = int(n_samples)
    v = np.ctypeslib.as_array(sp.variance).reshape(L.NW, 3, 3, 2)
    z = np.asarray(variance, dtype=complex).reshape(L.NW, 3, 3)
    v[..., 0] = z.real
    v[..., 1] = z.imag
    f = _to_struct(fd)
    b = np.zeros(3)
    check(L.load().st_floquet_integrate(C.byref(f), C.byref(sp), C.byref(opt), b.ctypes.data))
    return b
