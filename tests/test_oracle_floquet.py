"""Pins of the oracle's spectral path (integratorType 6): the Goertzel
aggregator, integrateHamiltonian and the Floquet host math.

Each test restates a check of the reference's own suite
(tests/floquet_validation.cpp) against an independent answer:
* the quaternion "diagonalisation" reproduces its input (:52-61);
* integrateHamiltonian's propagators rotate a spin exactly as
  integrateMagnusCFET does (:63-95);
* floquet_master_equation_rates and integrateFloquetMarkov reproduce the
  QuTiP golden tensors and Bloch vector the reference keeps
  (tests/golden/floquet_qutip.npz, from scripts/floquet_matrices and
  :97-153);
* the Goertzel recurrence equals the direct DFT up to its known phase
  (:155-182; the reference's version assigns instead of comparing);
* the aggregator's covariance spectrum of a sine equals the brute-force
  double sum of sine_spectrum (:23-47, :184-232);
* a z-noise spectrum at the Larmor frequency shifts the free precession by
  S/4 through floquet_integrate (:234-289).
"""
import ctypes
import math
import os

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden", "floquet_qutip.npz")


@pytest.fixture(scope="module")
def F(oracle):
    import floquet
    return floquet


def dressed_options(oracle, tf=1e-3):
    """The options of floquet_validation.cpp:65-74 (defaults otherwise)."""
    o = oracle.default_options()
    o["gravity"] = False
    o["T"] = 0.4
    o["B0"] = (3e-6, 0.0, 0.0)
    o["a"] = 3e-5
    o["w"] = 1 / tf * 2 * math.pi
    o["atol"] = 1e-12
    o["rtol"] = 1e-12
    return o


def test_quaternion_diagonalisation(F):
    q = np.array([2.0, 3.0, -5.0, 7.0])
    q = q / math.sqrt((q * q).sum())
    ev, evec = F.q_eigenval(q), F.q_eigenvec(q)
    assert ev[1] == 0 and ev[2] == 0
    a = F.qmul(F.qmul(evec, ev), F.qconj(evec))
    assert np.abs(a - q).max() < 1e-12


def propagators(oracle, F, o, tf):
    """The reference test's 100 propagators over one period, beside the spin
    rotated by integrateMagnusCFET over the same slices."""
    y = np.array([1.0, 0.0, 0.0, 0.0])
    s0 = np.array([0.0, 1.0, 0.0])
    s = s0.copy()
    h = 1e-6
    props, worst = [], 0.0
    for i in range(100):
        t1, t2 = tf * i / 100, tf * (i + 1) / 100
        st, y = F.integrate_hamiltonian(o, t1, t2, y, 1e-6)
        assert st == 0
        st, s, h, _, _ = oracle.integrate(o, t1, t2, s, np.zeros(3), np.zeros(3), h, integrator=2)
        assert st == 0
        ys = F.qmul(F.qmul(y, np.r_[0.0, s0]), F.qconj(y))[1:]
        worst = max(worst, np.abs(ys - s).max())
        props.append(y.copy())
    return np.array(props), worst


def test_hamiltonian_matches_magnus(oracle, F):
    o = dressed_options(oracle)
    _, worst = propagators(oracle, F, o, 1e-3)
    assert worst < 1e-8


def test_rates_match_qutip(oracle, F):
    g = np.load(GOLDEN)
    tf = 1e-3
    o = dressed_options(oracle, tf)
    props, _ = propagators(oracle, F, o, tf)
    ev, evec = F.q_eigenval(props[-1]), F.q_eigenvec(props[-1])
    ea = -math.atan2(ev[3], ev[0]) / tf
    dE = ea - (-ea)
    freqs = np.zeros(F.NW)
    power = np.zeros(F.NW, dtype=complex)
    for k in range(F.NK // 2 + 1):
        for i in (-1, 0, 1):
            w = dE * i + k * float(o["w"])
            if w < 0:
                continue
            freqs[k * 3 + i] = w
            power[k * 3 + i] = 1 / ((w * tf) ** 2 + 1)
    Z = np.zeros((2, 2), dtype=complex)
    Om = np.zeros((2, 2), dtype=complex)
    D, X, G = F.rates(evec, ev, np.array([0.0, 0.0, 1.0, 0.0]), props, 100, tf, freqs, power, Z, Om)
    A = g["A"]
    # QuTiP's A is for zero temperature; the reference's Zeta is A + A^T / (4 pi)
    assert np.allclose(np.abs(4 * math.pi * Z), A + A.T, rtol=1e-8, atol=0)
    assert np.allclose(D, g["Delta"], rtol=1e-10, atol=1e-9)
    assert np.abs(X.real - g["X_re"]).max() < 1e-8
    assert np.abs(X.imag - g["X_im"]).max() < 1e-8
    assert np.allclose(np.abs(X) ** 2, g["Xsq"], rtol=1e-7, atol=1e-12)
    assert np.abs(np.abs(G) - g["Gamma"]).max() < 1e-8
    rho = F.markov(0.0, 10 * tf, F.bloch_to_density([0.0, 1.0, 0.0]), Z, Om)
    assert np.abs(F.density_to_bloch(rho) - g["b"]).max() < 1e-8


def test_goertzel_is_the_dft(oracle):
    """One frequency of the aggregator against sum_n x_n e^{-i w n dt}: the
    stage-2 output a + i b is that sum times e^{i w (N-1) dt}."""
    o = oracle.default_options()
    w = 1.3e2
    dt = 1e-4
    freqs = np.zeros(oracle.NW)
    freqs[0] = w
    cfg = oracle.spec_cfg(freqs, dt)
    agg = np.zeros((), dtype=oracle.SPECAGG_DTYPE)
    w0 = np.array([1.1e2, 5.0e2, -2.3e2])
    amp = np.array([1.0, 2.0, 3.0])
    xs = []
    t = (math.floor(0.3 / dt) + 1) * dt
    end = (math.floor(1.2 / dt) + 1) * dt
    while t < end - dt / 2:
        x = amp * np.sin(w0 * t)
        oracle.lib().or_spectrum_update(agg.ctypes.data, ctypes.addressof(cfg), oracle._c3(x))
        xs.append(x)
        t += dt
    xs = np.array(xs)
    N = len(xs)
    assert int(agg["n_samples"]) == N
    dft = (xs * np.exp(-1j * w * dt * np.arange(N))[:, None]).sum(axis=0) * np.exp(1j * w * (N - 1) * dt)
    a = agg["s1"][0] - math.cos(w * dt) * agg["s2"][0]
    b = math.sin(w * dt) * agg["s2"][0]
    assert np.allclose(a + 1j * b, dft, rtol=1e-9, atol=1e-9 * np.abs(dft).max())


def sine_spectrum(a, w1, w2, t0, tf, dt):
    """floquet_validation.cpp:23-47 by numpy: 2 dt/n sum_n [P_n conj(x_n
    e^{i w2 t_n}) - x_n^2/2], P_n the running sum of x_m e^{i w2 t_m}."""
    t = (math.floor(t0 / dt) + 1) * dt
    end = (math.floor(tf / dt) + 1) * dt
    ts = []
    while t < end - dt / 2:
        ts.append(t)
        t += dt
    ts = np.array(ts)
    x = a * np.sin(w1 * ts)
    x2 = x * np.exp(1j * w2 * ts)
    P = np.cumsum(x2)
    total = (P * np.conj(x2) - x * x / 2).sum()
    return total * 2.0 * dt / len(ts)


def test_spectrum_of_a_sine(oracle):
    """floquet_validation.cpp:184-232: the normalised covariance spectrum of
    an injected sine at its own frequency, per axis."""
    o = oracle.default_options()
    o["gravity"] = False
    o["a"] = 1.0
    o["w"] = 1e3 * 2 * math.pi
    o["noiseAmplitudes"] = (1.0, 2.0, 3.0)
    o["noiseFrequencies"] = (1.0e3, 3.5e3, -3.5e3)
    o["B0"] = (3e-6, 0.0, 0.0)
    o["E"] = (0.0, 0.0, 0.0)
    t0, tf, dt = 0.1, 1.2, 1e-5
    freqs = np.zeros(oracle.NW)
    freqs[:3] = o["noiseFrequencies"]
    cfg = oracle.spec_cfg(freqs, dt)
    agg = np.zeros((), dtype=oracle.SPECAGG_DTYPE)
    n = oracle.integrate_spectrum(o, t0, tf, agg, cfg, np.zeros(3), np.zeros(3), dt)
    assert n == int(agg["n_samples"]) and abs(n - 110000) <= 1
    cov = oracle.covariances(agg.reshape(1), cfg)[0]
    var = (cov[..., 0] + 1j * cov[..., 1]) * (dt / n)
    g = float(o["gamma"])
    for i in range(3):
        ref = sine_spectrum(g / 2 * o["noiseAmplitudes"][i], o["noiseFrequencies"][i],
                            o["noiseFrequencies"][i], t0, tf, dt)
        got = var[i, i, i]
        assert abs(got.real - ref.real) <= 1e-6 * abs(ref.real), i
        assert abs(got.imag - ref.imag) <= 1e-6 * abs(ref), i


def test_free_precession_shift(oracle, F):
    """floquet_validation.cpp:234-289: noise S_zz at the Larmor frequency
    moves the free precession to w0 + S_zz/4."""
    o = oracle.default_options()
    o["gravity"] = False
    o["a"] = 0.0
    o["w"] = 1e3 * 2 * math.pi
    o["B0"] = (3e-6, 0.0, 0.0)
    o["E"] = (0.0, 0.0, 0.0)
    o["t0"], o["tf"] = 0.0, 1.2
    o["rtol"] = o["atol"] = 1e-12
    o["yi"] = (0.0, 1.0, 0.0)
    dt = 1e-5
    fd = F.diagonalize(o)
    Szz = 5e-3
    nsamp = int((1.2 - 0.0) / dt)
    var = np.zeros((F.NW, 3, 3), dtype=complex)
    hit = 0
    for i in range(F.NW):
        if abs(fd["frequencies"][i] - abs(o["gamma"] * 3e-6)) < 1e-3:
            var[i, 2, 2] = Szz / 4 * 1j * (nsamp / dt)
            hit += 1
    assert hit == 1
    b_end = F.integrate(fd, var, nsamp, dt, o)
    dur = 1.2
    w0 = 3e-6 * float(o["gamma"])
    w2 = w0 + Szz / 4
    b_ref = np.array([0.0, math.cos(dur * w2), -math.sin(dur * w2)])
    b0 = np.array([0.0, math.cos(dur * w0), -math.sin(dur * w0)])
    shift_ref = math.asin(np.linalg.norm(np.cross(b_ref, b0)))
    shift_fm = math.asin(np.linalg.norm(np.cross(b_end, b0)))
    assert abs(shift_fm - shift_ref) <= 1e-6 * abs(shift_ref)


def test_partial_tree_indices(oracle):
    """src/particle.cpp:826-834 replayed literally on integer tags: slot 0
    ends up with particles 0 and 2^k < n."""
    for n in (1, 2, 3, 5, 8, 13, 64, 100):
        slots = [{j} for j in range(n)]
        for ip in range(n):
            s = 1
            while s < n:
                if ip % (2 * s) == 0 and ip + s < n:
                    slots[ip] = slots[ip] | slots[ip + s]
                s *= 2
        assert sorted(slots[0]) == oracle.partial_tree_indices(n), n
