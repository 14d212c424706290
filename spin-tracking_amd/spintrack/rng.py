"""Host-side draws from the reference's random stream, vectorised over streams.

Workload generators (field configurations, per-particle durations) draw from
the same generator the particles use, so a sweep is fixed by its seed:
stream (seed, index) is the one initParticlesCPU gives particle ``index``
(src/particle.cpp:646-663: splitmix64 of seed + index, the state words
taken from two outputs' 32-bit halves, three discarded draws), advanced by
the reference's xoshiro variant (src/particle.cpp:59-70) and mapped to
[0, 1) as uniform() does (src/particle.cpp:72-87).
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix(s: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    s = s + np.uint64(0x9E3779B97F4A7C15)
    z = s
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return s, z ^ (z >> np.uint64(31))


class Streams:
    """len(index) independent streams of seed ``seed``."""

    def __init__(self, seed: int, index):
        idx = np.atleast_1d(np.asarray(index, dtype=np.uint64))
        with np.errstate(over="ignore"):
            s = (np.uint64(seed) + idx) & np.uint64(0xFFFFFFFF)   # 32-bit sum, as the reference
            s, r1 = _splitmix(s)
            _, r2 = _splitmix(s)
        lo = np.uint64(0xFFFFFFFF)
        self.x = r1 & lo
        self.y = r1 >> np.uint64(32)
        self.z = r2 & lo
        self.w = r2 >> np.uint64(32)
        for _ in range(3):
            self.next_u64()

    def next_u64(self) -> np.ndarray:
        with np.errstate(over="ignore"):
            r = self.x + self.z
        t = self.y << np.uint64(17)
        self.z = self.z ^ self.x
        self.w = self.w ^ self.y
        self.y = self.y ^ self.z
        self.x = self.x ^ self.w
        self.z = self.z ^ t
        self.w = (self.z << np.uint64(45)) | (self.z >> np.uint64(19))
        return r

    def uniform(self) -> np.ndarray:
        return (self.next_u64() >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniforms(seed: int, index: int, count: int) -> np.ndarray:
    """``count`` consecutive uniforms of one stream."""
    s = Streams(seed, [index])
    return np.array([s.uniform()[0] for _ in range(count)])
