"""The reference's binary output format (src/logger.cpp:14-38).

File = the raw 440-byte ``options`` block (BinaryLogger::writeOptions,
src/logger.cpp:27-29), then one block per snapshot holding, for all N
particles, t[N] f64, pos[N] 3xf64, v[N] 3xf64, S[N] 3xf64, failureState[N]
i32, n_coll[N] u64, n_bounce[N] u64, n_steps[N] u64 (writeSnapshot,
src/logger.cpp:14-25) -- 108*N bytes per block.  There are
floor((tf-t0)/ioutInt)+1 blocks in a complete run; a spectral run
(integratorType 6) appends its final Bloch vector b_end as 3 f64
(BinaryLogger::writeSingle, src/logger.cpp:31-33; particle::postProcess,
src/particle.cpp:845-850).
"""
from __future__ import annotations

import numpy as np

from ._lib import Options
from .options import from_bytes, to_bytes
from .particle import SNAPSHOT_FIELDS, snapshot_block_dtype


class BinaryLogger:
    """src/logger.cpp:14-43 (without the double delete of the FILE*)."""

    def __init__(self, opt: Options, output_name: str):
        self.opt = opt
        self.n = int(opt.numParticles)
        self.f = open(output_name, "wb")
        self.writeOptions()

    def writeOptions(self) -> None:
        self.f.write(to_bytes(self.opt))

    def writeSnapshot(self, t, pos, v, S, failure, n_coll, n_bounce, n_steps) -> None:
        for arr, dt in ((t, "<f8"), (pos, None), (v, None), (S, None), (failure, "<i4"),
                        (n_coll, "<u8"), (n_bounce, "<u8"), (n_steps, "<u8")):
            a = np.ascontiguousarray(arr)
            if dt is not None:
                a = a.astype(dt, copy=False)
            assert a.shape[0] == self.n
            self.f.write(a.tobytes())

    def writeBlock(self, block: bytes | np.ndarray) -> None:
        """A whole snapshot block already in file layout (from the GPU)."""
        b = block if isinstance(block, (bytes, bytearray)) else np.ascontiguousarray(block).tobytes()
        assert len(b) == 108 * self.n
        self.f.write(b)

    def writeSingle(self, name: str, value) -> None:
        """src/logger.cpp:31-33: the raw coords, no name."""
        self.f.write(np.asarray(value, dtype="<f8").reshape(3).tobytes())

    def close(self) -> None:
        if self.f:
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_output(path: str):
    """Reads a binary output file: returns (options, dict of arrays shaped
    (n_snapshots, N)).  The counterpart of scripts/nedmSim.py:6-107 for the
    current options layout; a trailing b_end (spectral runs) is returned
    under "b_end"."""
    with open(path, "rb") as f:
        raw = f.read()
    opt = from_bytes(raw[:440])
    n = int(opt.numParticles)
    bsz = 108 * n
    nblk = (len(raw) - 440) // bsz
    blocks = np.frombuffer(raw, dtype=snapshot_block_dtype(n), count=nblk, offset=440)
    out = {k: np.array(blocks[k]) for k in SNAPSHOT_FIELDS}
    tail = len(raw) - 440 - nblk * bsz
    if tail == 24:
        out["b_end"] = np.frombuffer(raw, dtype="<f8", count=3, offset=440 + nblk * bsz).copy()
    return opt, out
