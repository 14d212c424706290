"""Ensemble precession analysis on top of the device reductions.

``phase_series`` runs st_phase_stats over snapshot blocks in HBM: the
reference's per-particle phase analysis (histogram-mode alignment, 5-sigma
cut; benchmarks/CheckingFrequencyShiftFromEField.ipynb:5-27,
scripts/nedmSim.py:151-178) per snapshot and configuration.
``frequency_shift`` turns the mean phases into the frequency shift against
free precession at gamma*|B0| and its error, in nHz, as the notebook's
runTests does.  The notebook divides the phase spread by the square root of
the number of snapshots (phis.shape[0] of a (snapshots, particles) array);
this returns the standard error over the particles kept, and the notebook's
figure as ``errors_notebook``.
"""
from __future__ import annotations

import numpy as np


def phase_series(handle, blocks_dev_ptr: int, n_blocks: int) -> dict:
    """Per snapshot block k and configuration c: mean phase (wrapped), std,
    kept and selected counts, reference angle; arrays shaped (n_blocks, n_cfg)."""
    r = handle.phase_stats(blocks_dev_ptr, n_blocks)
    return dict(mean=r[..., 0], std=r[..., 1], kept=r[..., 2], selected=r[..., 3], ref=r[..., 4])


def frequency_shift(times, mean, std, kept, gamma: float, B0) -> dict:
    """Shift of the mean phase from free precession, per unit time, in nHz."""
    t = np.asarray(times, dtype=np.float64)
    b = np.asarray(B0, dtype=np.float64)
    w = gamma * float(np.sqrt(np.sum(b * b))) * float(np.sign(b[0]))
    free = np.arctan2(np.cos(w * t), np.sin(w * t))
    d = np.asarray(mean, dtype=np.float64) - free
    d = np.where(d > np.pi, d - 2.0 * np.pi, d)
    d = np.where(d < -np.pi, d + 2.0 * np.pi, d)
    scale = 1e9 / (2.0 * np.pi) / t
    std = np.asarray(std, dtype=np.float64)
    return dict(shift=d * scale, error=std / np.sqrt(np.asarray(kept, dtype=np.float64)) * scale,
                errors_notebook=std / np.sqrt(len(t)) * scale)
