"""CPU restatement of the reference's ensemble phase analysis -- test
infrastructure only (the checker for st_phase_stats; never part of the
product path).

* ``shift_datapoints``  -- scripts/nedmSim.py:151-178 (histogram-mode
  alignment of angles, 5-sigma cut to NaN)
* ``analyze``           -- benchmarks/CheckingFrequencyShiftFromEField.ipynb,
  analyzeDataset (selection by the last snapshot's failure code and phase,
  per-snapshot shift_datapoints, nanmean wrapped into [-pi, pi], nanstd)
* ``frequency_shift``   -- the same notebook's runTests: phase minus the free
  precession phase atan2(cos(w t), sin(w t)), wrapped, per unit time in nHz,
  and its error std/sqrt(count)/t.
"""
from __future__ import annotations

import numpy as np


def shift_datapoints(phis: np.ndarray) -> np.ndarray:
    """scripts/nedmSim.py:151-178."""
    shifted = np.copy(phis)
    hist, bins = np.histogram(phis, bins=1000)
    location = bins[np.argmax(hist)]
    low = shifted <= (location - np.pi)
    high = shifted >= (location + np.pi)
    shifted[low] += 2.0 * np.pi
    shifted[high] -= 2.0 * np.pi
    mean = np.nanmean(shifted)
    std = np.nanstd(shifted)
    shifted[shifted < (mean - std * 5)] = np.nan
    shifted[shifted > (mean + std * 5)] = np.nan
    return shifted, location


def select(S: np.ndarray, failure_last: np.ndarray) -> np.ndarray:
    """phis (K, M) of the particles the notebook keeps: failure == 0 and a
    finite phase at the last snapshot.  S is (K, N, 3)."""
    valid = np.argwhere(failure_last == 0)[:, 0]
    phis = np.arctan2(S[:, valid, 2], S[:, valid, 1])
    ok = np.argwhere(np.logical_not(np.isnan(phis[-1, :])))[:, 0]
    return phis[:, ok]


def analyze(phis: np.ndarray):
    """analyzeDataset on selected phases (K, M): per snapshot the aligned and
    cut phases; returns means (wrapped), stds, kept counts, reference angles."""
    out = np.empty_like(phis)
    refs = np.empty(phis.shape[0])
    for i in range(phis.shape[0]):
        out[i], refs[i] = shift_datapoints(phis[i])
    means = np.nanmean(out, axis=1)
    means[means > np.pi] -= np.pi * 2.0
    means[means < -np.pi] += np.pi * 2.0
    stds = np.nanstd(out, axis=1)
    kept = np.sum(~np.isnan(out), axis=1)
    return means, stds, kept, refs


def frequency_shift(times, means, stds, count, gamma, B0):
    """runTests: shift = mean - atan2(cos(w t), sin(w t)), w = gamma*|B0|*sign(B0x),
    wrapped into [-pi, pi]; shifts/t/(2 pi)*1e9 and errors
    std/sqrt(count)/t/(2 pi)*1e9."""
    times = np.asarray(times, dtype=np.float64)
    w = gamma * np.sqrt(np.sum(np.square(B0))) * np.sign(B0[0])
    calc = np.arctan2(np.cos(w * times), np.sin(w * times))
    shift = np.asarray(means) - calc
    shift[shift > np.pi] -= 2.0 * np.pi
    shift[shift < -np.pi] += 2.0 * np.pi
    return shift / times / (2.0 * np.pi) * 1e9, np.asarray(stds) / np.sqrt(count) / times / (2.0 * np.pi) * 1e9
