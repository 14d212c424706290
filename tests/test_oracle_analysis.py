"""Pins of the oracle's restatement of the reference's phase analysis
(oracle/analysis.py) on constructed ensembles whose answers are known, and
the product's frequency-shift formula against it."""
import numpy as np

import analysis as A
from spintrack import analysis as SA


def test_alignment_across_the_branch_cut():
    # phases clustered at pi: half of them come out of atan2 near -pi
    rng = np.random.default_rng(1)
    true = np.pi - 0.01 + 0.02 * rng.standard_normal(20000)
    phis = np.angle(np.exp(1j * true))           # wrapped into (-pi, pi]
    assert (phis < 0).mean() > 0.2
    out, ref = A.shift_datapoints(phis)
    assert np.nanmax(out) - np.nanmin(out) < 0.3    # one cluster again
    assert abs(np.nanmean(out) - np.mean(true)) < 1e-3 or abs(np.nanmean(out) - np.mean(true) + 2 * np.pi) < 1e-3
    means, stds, kept, refs = A.analyze(phis[None, :])
    assert -np.pi <= means[0] <= np.pi
    assert abs(abs(means[0]) - abs(np.angle(np.exp(1j * np.mean(true))))) < 1e-3
    assert abs(stds[0] - 0.02) < 1e-3


def test_five_sigma_cut():
    rng = np.random.default_rng(2)
    phis = 0.3 + 1e-3 * rng.standard_normal(50000)
    phis[123] = 0.3 + 0.5            # far outside
    out, _ = A.shift_datapoints(phis)
    assert np.isnan(out[123]) and np.sum(np.isnan(out)) < 5
    means, stds, kept, _ = A.analyze(phis[None, :])
    assert kept[0] == np.sum(~np.isnan(out))
    assert abs(means[0] - 0.3) < 2e-5


def test_selection_by_last_snapshot():
    S = np.zeros((3, 4, 3))
    S[..., 1] = 1.0
    S[2, 1, :] = np.nan                # NaN at the last snapshot -> dropped everywhere
    fail = np.array([0, 0, 2, 0])      # failure at the last snapshot -> dropped
    phis = A.select(S, fail)
    assert phis.shape == (3, 2)


def test_frequency_shift_recovers_a_known_offset():
    gamma, B0 = -1.832e8, np.array([3e-6, 0.0, 0.0])
    w = gamma * 3e-6
    dw = 2 * np.pi * 5e-9 * 1e3              # 5 uHz = 5000 nHz
    t = np.linspace(0.5, 5.0, 10)
    phase = np.angle(np.exp(1j * np.arctan2(np.cos((w + dw) * t), np.sin((w + dw) * t))))
    # atan2(cos, sin) = pi/2 - angle, so a faster precession lowers the phase
    s_or, e_or = A.frequency_shift(t, phase, np.full(10, 0.1), 1000, gamma, B0)
    assert np.allclose(s_or, -5000.0, rtol=1e-6)
    got = SA.frequency_shift(t, phase, np.full(10, 0.1), np.full(10, 1000.0), gamma, B0)
    assert np.allclose(got["shift"], s_or, rtol=1e-13, atol=0)
    assert np.allclose(got["error"], e_or, rtol=1e-13)
