"""The library's host-side Floquet routines (spintrack.floquet over the C
ABI; no GPU work) against the reference's golden vectors and the oracle.

* floquet_diagonalize equals the oracle's restatement (which is pinned to
  integrateMagnusCFET in test_oracle_floquet.py);
* floquet_master_equation_rates + integrateFloquetMarkov reproduce the QuTiP
  tensors and Bloch vector the reference keeps (tests/golden/floquet_qutip.npz,
  the check of tests/floquet_validation.cpp:97-153);
* the closed-form 2x2 exponentials agree with scipy's expm through the
  oracle's integrateFloquetMarkov on random rate matrices;
* floquet_integrate reproduces the free-precession shift S/4 of
  floquet_validation.cpp:234-289 and the oracle's b_end.
"""
import math
import os

import numpy as np
import pytest

import spintrack as st
from spintrack import floquet as SF
from conftest import ROOT, to_oracle_opt

GOLDEN = os.path.join(ROOT, "tests", "golden", "floquet_qutip.npz")


@pytest.fixture(scope="module")
def OF(oracle):
    import floquet
    return floquet


def dressed(tf=1e-3):
    o = st.default_options()
    o.gravity = 0
    o.T = 0.4
    o.B0 = st.Coords(3e-6, 0.0, 0.0)
    o.a = 3e-5
    o.w = 1 / tf * 2 * math.pi
    o.atol = o.rtol = 1e-12
    return o


def test_diagonalize_matches_oracle(OF):
    o = dressed()
    got = SF.diagonalize(o)
    ref = OF.diagonalize(to_oracle_opt(o))
    assert got["n_prop"] == ref["n_prop"] == 100
    assert np.array_equal(got["propagators"], ref["propagators"])
    for k in ("f_modes_0", "f_energies", "frequencies"):
        assert np.array_equal(got[k], ref[k]), k
    assert got["dt"] == ref["dt"]


def test_rates_and_markov_match_qutip():
    g = np.load(GOLDEN)
    tf = 1e-3
    fd = SF.diagonalize(dressed(tf))
    props = fd["propagators"]
    ev, evec = fd["f_energies"], fd["f_modes_0"]
    dE = 2 * (-math.atan2(ev[3], ev[0]) / tf)
    freqs = np.zeros(17)
    power = np.zeros(17, dtype=complex)
    for k in range(6):
        for i in (-1, 0, 1):
            w = dE * i + k * (2 * math.pi / tf)
            if w < 0:
                continue
            freqs[k * 3 + i] = w
            power[k * 3 + i] = 1 / ((w * tf) ** 2 + 1)
    Z = np.zeros((2, 2), dtype=complex)
    Om = np.zeros((2, 2), dtype=complex)
    D, X, G = SF.rates(evec, ev, [0.0, 0.0, 1.0, 0.0], props, tf, freqs, power, Z, Om)
    A = g["A"]
    assert np.allclose(np.abs(4 * math.pi * Z), A + A.T, rtol=1e-8, atol=0)
    assert np.allclose(D, g["Delta"], rtol=1e-10, atol=1e-9)
    assert np.abs(X.real - g["X_re"]).max() < 1e-8
    assert np.abs(X.imag - g["X_im"]).max() < 1e-8
    assert np.abs(np.abs(G) - g["Gamma"]).max() < 1e-8
    rho0 = np.array([[1, -1j], [1j, 1]]) / 2      # Bloch vector (0, 1, 0)
    rho = SF.markov(0.0, 10 * tf, rho0, Z, Om)
    sx = 2 * rho[0, 1].real
    sy = -2 * rho[0, 1].imag
    sz = (rho[0, 0] - rho[1, 1]).real
    assert np.abs(np.array([sx, sy, sz]) - g["b"]).max() < 1e-8


def test_markov_exponentials_match_expm(OF):
    rng = np.random.default_rng(5)
    for _ in range(20):
        Z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        Om = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        rho = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        tf = float(rng.uniform(0.01, 2.0))
        got = SF.markov(0.0, tf, rho, Z, Om)
        ref = OF.markov(0.0, tf, rho, Z, Om)
        assert np.allclose(got, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    # the small-|s| series branch: a nearly scalar rate matrix
    Z = np.array([[0.3, 1e-7], [1e-7, 0.3]], dtype=complex)
    Om = np.array([[0.2, 1e-6], [2e-6, 0.2]], dtype=complex)
    rho = np.array([[0.6, 0.1 - 0.2j], [0.1 + 0.2j, 0.4]])
    assert np.allclose(SF.markov(0.0, 0.5, rho, Z, Om), OF.markov(0.0, 0.5, rho, Z, Om), rtol=1e-13, atol=1e-14)


def free_options():
    o = st.default_options()
    o.gravity = 0
    o.a = 0.0
    o.w = 1e3 * 2 * math.pi
    o.B0 = st.Coords(3e-6, 0.0, 0.0)
    o.E = st.Coords(0.0, 0.0, 0.0)
    o.t0, o.tf = 0.0, 1.2
    o.rtol = o.atol = 1e-12
    o.yi = st.Coords(0.0, 1.0, 0.0)
    return o


def test_free_precession_shift_and_oracle(OF):
    o = free_options()
    fd = SF.diagonalize(o)
    dt, Szz = 1e-5, 5e-3
    nsamp = int(1.2 / dt)
    var = np.zeros((17, 3, 3), dtype=complex)
    i = int(np.argmin(np.abs(fd["frequencies"] - abs(o.gamma * 3e-6))))
    var[i, 2, 2] = Szz / 4 * 1j * (nsamp / dt)
    b_end = SF.integrate(fd, var, nsamp, dt, o)
    w0 = 3e-6 * o.gamma
    w2 = w0 + Szz / 4
    b_ref = np.array([0.0, math.cos(1.2 * w2), -math.sin(1.2 * w2)])
    b0 = np.array([0.0, math.cos(1.2 * w0), -math.sin(1.2 * w0)])
    shift_ref = math.asin(np.linalg.norm(np.cross(b_ref, b0)))
    shift = math.asin(np.linalg.norm(np.cross(b_end, b0)))
    assert abs(shift - shift_ref) <= 1e-6 * abs(shift_ref)
    # the same spectrum through the oracle (scipy expm, numpy complex)
    ofd = OF.diagonalize(to_oracle_opt(o))
    ref = OF.integrate(ofd, var, nsamp, dt, to_oracle_opt(o))
    assert np.abs(b_end - ref).max() < 1e-10


def test_floquet_argument_checks():
    o = dressed()
    o.w = 0.0
    with pytest.raises(st.SpinTrackError):
        SF.diagonalize(o)
