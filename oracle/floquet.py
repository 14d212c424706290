"""Floquet host math of the spectral path (integratorType 6), in numpy.

TEST INFRASTRUCTURE ONLY, like the rest of oracle/: tests/ use it as the
checker of the library's host-side Floquet routines and of the spectra the
GPU accumulates.

Restated from the reference (paths relative to its repository):
  src/integrator.cpp:791-822   integrateFloquetMarkov  (matrix exponentials
                               by scipy.linalg.expm, an independent routine;
                               the reference uses Eigen's MatrixFunctions)
  src/integrator.cpp:824-867   floquet_diagonalize     (the propagators come
                               from the C oracle's or_integrate_hamiltonian)
  src/integrator.cpp:869-889   floquet_integrate
  src/floquet.cpp:51-64        Spectrum::lookup
  src/floquet.cpp:81-103       CovarianceSpectrum::normalize / extract
  src/floquet.cpp:192-279      floquet_master_equation_rates
  src/floquet.cpp:281-301      bloch_to_density / density_to_bloch
  src/quaternion.cpp:128-135, 169-201   qMult, qEigenval, qEigenvec, toSU2, pow

Quaternions are length-4 float arrays (w, x, y, z); complex matrices are
numpy complex128.  Spectrum::lookup past the last tabulated frequency reads
one element beyond the reference's array (src/floquet.cpp:63); here it
returns the last tabulated power (the nearest frequency).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import scipy.linalg

import oracle as O

NW = 17   # include/floquet.h:34
NK = 11   # include/floquet.h:33


# ------------------------------------------------------------ quaternions
def qmul(a, b) -> np.ndarray:
    """qMult, src/quaternion.cpp:128-135."""
    aw, ax, ay, az = (float(v) for v in a)
    bw, bx, by, bz = (float(v) for v in b)
    return np.array([aw * bw - ax * bx - ay * by - az * bz,
                     aw * bx + ax * bw + ay * bz - az * by,
                     aw * by + ay * bw + az * bx - ax * bz,
                     aw * bz + az * bw + ax * by - ay * bx])


def qconj(q) -> np.ndarray:
    return np.array([q[0], -q[1], -q[2], -q[3]], dtype=np.float64)


def qscale(q, a: float) -> np.ndarray:
    return np.array([q[0] * a, q[1] * a, q[2] * a, q[3] * a], dtype=np.float64)


def q_eigenval(q) -> np.ndarray:
    """qEigenval, src/quaternion.cpp:169-172."""
    return np.array([q[0], 0.0, 0.0, -math.sqrt(1.0 - q[0] * q[0])])


def q_eigenvec(q) -> np.ndarray:
    """qEigenvec, src/quaternion.cpp:174-177: q1 / |q1| as q1 * (1/|q1|)."""
    q1 = np.array([math.sqrt(1.0 - q[0] * q[0]) - q[3], q[2], -q[1], 0.0])
    n = math.sqrt(q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2] + q1[3] * q1[3])
    return qscale(q1, 1.0 / n)


def qpow(q, n: int) -> np.ndarray:
    """pow(quaternion, int), src/quaternion.cpp:190-201 (halving recursion)."""
    if n == 0:
        return np.array([1.0, 0.0, 0.0, 0.0])
    if n == 1:
        return np.asarray(q, dtype=np.float64)
    half = n // 2
    q2 = qpow(q, half)
    qr = qpow(q, n - half - half)
    return qmul(qmul(q2, q2), qr)


def to_su2(q) -> np.ndarray:
    """toSU2, src/quaternion.cpp:179-188."""
    w, x, y, z = (float(v) for v in q)
    return np.array([[complex(w, z), complex(y, x)], [complex(-y, x), complex(w, -z)]])


def bloch_to_density(b, basis=(1.0, 0.0, 0.0, 0.0)) -> np.ndarray:
    """src/floquet.cpp:281-295."""
    B = to_su2(basis)
    x, y, z = (float(v) for v in b)
    rho = np.array([[1 + z, complex(x, -y)], [complex(x, y), 1 - z]])
    return B @ rho @ B.conj().T / 2.0


def density_to_bloch(rho, basis=(1.0, 0.0, 0.0, 0.0)) -> np.ndarray:
    """src/floquet.cpp:297-301 (traces against the Pauli matrices)."""
    B = to_su2(basis)
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    sy = np.array([[0, -1j], [1j, 0]], dtype=complex)
    sz = np.array([[1, 0], [0, -1]], dtype=complex)
    lab = B.conj().T @ rho @ B
    return np.array([np.trace(lab @ sx).real, np.trace(lab @ sy).real, np.trace(lab @ sz).real])


# ----------------------------------------------------- diagonalisation
def integrate_hamiltonian(opt, t0: float, tf: float, q, h: float, pulse_axis: int = 2):
    """integrateHamiltonian (src/integrator.cpp:745-789) through the C oracle;
    returns (status, q_out)."""
    qq = np.array(q, dtype=np.float64)
    e = O._ext(pulse_axis)
    st = O.lib().or_integrate_hamiltonian(t0, tf, O._ptr(qq), O._ptr(opt), C.addressof(e), h)
    return st, qq


def diagonalize(opt, pulse_axis: int = 2) -> dict:
    """floquet_diagonalize, src/integrator.cpp:824-867: 100 propagators over
    one drive period 2 pi / w, the Floquet modes and quasi-energies of the
    last, and the NW sampling frequencies deltaE*i + k*w."""
    t0 = 0.0
    tf = (2 * math.pi) / float(opt["w"])
    n_prop = 100
    y = np.array([1.0, 0.0, 0.0, 0.0])
    h = 1e-6
    props = []
    for i in range(n_prop):
        t1 = t0 + (tf - t0) * i / n_prop
        t2 = t0 + (tf - t0) * (i + 1) / n_prop
        _, y = integrate_hamiltonian(opt, t1, t2, y, h, pulse_axis)
        props.append(y.copy())
    ev = q_eigenval(props[-1])
    evec = q_eigenvec(props[-1])
    ea = abs(math.atan2(ev[3], ev[0])) / (tf - t0)
    eb = -ea
    dE = ea - eb
    freqs = np.zeros(NW)
    for k in range(NK // 2 + 1):
        for i in (-1, 0, 1):
            idx = k * 3 + i
            if idx >= 0:
                freqs[idx] = dE * i + k * float(opt["w"])
    return dict(propagators=np.array(props), f_modes_0=evec, f_energies=ev, frequencies=freqs,
                dt=(tf - t0) / n_prop, n_prop=n_prop)


# ----------------------------------------------------------------- rates
def lookup(freqs, power, f: float) -> complex:
    """Spectrum::lookup, src/floquet.cpp:51-64: scan until the distance
    stops decreasing; negative frequencies give the conjugate."""
    if f < 0:
        return lookup(freqs, power, -f).conjugate()
    best = math.inf
    for i in range(NW):
        d = abs(freqs[i] - f)
        if d < best:
            best = d
        else:
            return complex(power[i - 1])
    return complex(power[NW - 1])


def heaviside(x: float) -> float:
    """src/floquet.cpp:170-176."""
    if x == 0.0:
        return 0.5
    return ((-1.0 if x < 0.0 else 1.0) + 1) / 2


def rates(f_modes_0, f_energies, c_op, propagators, n_prop: int, period: float, freqs, power,
          Zeta: np.ndarray, Omicron: np.ndarray):
    """floquet_master_equation_rates, src/floquet.cpp:192-279.  Returns
    (Delta[2,2,NK], X[2,2,NK], Gamma[2,2,NK]) and adds to Zeta and Omicron
    (2x2 complex) in place."""
    Xq = np.zeros((NK, 4))
    omega = 2 * math.pi / period
    ea = math.atan2(f_energies[3], f_energies[0]) / period
    eb = -math.atan2(f_energies[3], f_energies[0]) / period
    es = (ea, eb)
    kmax = NK // 2
    for i in range(n_prop):
        t = (i + 1) * period / n_prop
        weight = 1.0 / n_prop
        phase = np.array([math.cos(ea * t), 0.0, 0.0, -math.sin(ea * t)])
        fm = qmul(qmul(propagators[i], f_modes_0), phase)
        q = qmul(qmul(qconj(fm), c_op), fm)
        for k in range(NK):
            arg = (k - kmax) * omega * t
            kp = np.array([math.cos(arg), 0.0, 0.0, math.sin(arg)])
            Xq[k] = Xq[k] + qscale(qmul(q, kp), weight)
    X = np.zeros((2, 2, NK), dtype=complex)
    for k in range(NK):
        X[0, 0, k] = complex(Xq[k][3], -Xq[k][0])
        X[1, 1, k] = X[0, 0, k].conjugate()
        X[0, 1, k] = complex(Xq[k][1], -Xq[k][2])
        X[1, 0, k] = complex(Xq[2 * kmax - k][1], Xq[2 * kmax - k][2])
    Xsq = X.real * X.real + X.imag * X.imag
    Delta = np.zeros((2, 2, NK))
    Gamma = np.zeros((2, 2, NK), dtype=complex)
    for k in range(NK):
        for i in range(2):
            for j in range(2):
                f = (es[j] - es[i]) + (k - NK // 2) * omega
                Delta[i, j, k] = f
                Gamma[i, j, k] = Xsq[i, j, k] * lookup(freqs, power, f) * 2.0 * math.pi * heaviside(f)
    for k in range(NK):
        for i in range(2):
            for j in range(2):
                Zeta[i, j] = Zeta[i, j] + Xsq[i, j, k] * lookup(freqs, power, Delta[i, j, k]) / 2.0
    for k in range(NK):
        for i in range(2):
            for j in range(2):
                f = (k - NK // 2) * omega
                Omicron[i, j] = Omicron[i, j] + lookup(freqs, power, f) * X[i, i, k] * X[j, j, k].conjugate()
    return Delta, X, Gamma


def markov(t0: float, tf: float, rho: np.ndarray, Zeta: np.ndarray, Omicron: np.ndarray) -> np.ndarray:
    """integrateFloquetMarkov, src/integrator.cpp:791-822: populations and
    coherences each under a 2x2 linear system, exp(A dt)."""
    dt = tf - t0
    Z, Om = Zeta, Omicron
    rho = np.array(rho, dtype=complex)
    A_diag = np.array([[-Z[1, 0] + Z[0, 1], Z[0, 1] + Z[1, 0]],
                       [Z[1, 0] + Z[0, 1], -Z[1, 0] + Z[0, 1]]])
    p1 = scipy.linalg.expm(A_diag * dt) @ np.array([rho[0, 0], rho[1, 1]])
    rho[0, 0], rho[1, 1] = p1
    d01 = -(Z[0, 0] + Z[0, 1] + Z[0, 1] + Z[1, 1])
    d10 = -(Z[1, 0] + Z[1, 1] + Z[0, 0] + Z[1, 0])
    A_off = np.array([[Om[0, 1] + d01, Om[1, 0]], [Om[0, 1], Om[1, 0] + d10]])
    q1 = scipy.linalg.expm(A_off * dt) @ np.array([rho[0, 1], rho[1, 0]])
    rho[0, 1], rho[1, 0] = q1
    return rho


def integrate(fd: dict, variance: np.ndarray, n_samples: int, dt: float, opt) -> np.ndarray:
    """floquet_integrate, src/integrator.cpp:869-889: normalise the summed
    covariance spectrum (variance[NW,3,3] complex), take its diagonal as the
    noise spectra of sigma_x, sigma_y, sigma_z, accumulate the rates, evolve
    the density matrix over (t0, tf] and return the Bloch vector b_end in the
    lab frame after the whole drive periods."""
    var = np.asarray(variance, dtype=complex) * (dt / n_samples)
    Zeta = np.zeros((2, 2), dtype=complex)
    Om = np.zeros((2, 2), dtype=complex)
    rho = bloch_to_density(np.asarray(opt["yi"], dtype=np.float64), fd["f_modes_0"])
    period = fd["dt"] * fd["n_prop"]
    for i in range(3):
        c_op = np.zeros(4)
        c_op[i + 1] = 1.0
        rates(fd["f_modes_0"], fd["f_energies"], c_op, fd["propagators"], fd["n_prop"], period,
              fd["frequencies"], var[:, i, i], Zeta, Om)
    rho = markov(float(opt["t0"]), float(opt["tf"]), rho, Zeta, Om)
    x = (float(opt["tf"]) - float(opt["t0"])) * float(opt["w"]) / (2 * math.pi)
    n_period = int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)   # C round()
    return density_to_bloch(rho, qmul(fd["f_modes_0"], qpow(fd["f_energies"], n_period)))
