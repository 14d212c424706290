"""Launch-shape autotuner for the pipeline engine (plain and batched handles).

Knobs (all in st_ext):
* ``chunk_records`` -- trajectory segments staged per particle per pipeline
  chunk: the depth of the segment-record buffer in HBM (records x particles
  x 64 B).  Deeper chunks mean fewer kernel pairs and host synchronisations;
  shallower ones keep the buffer small and shorten the spin kernel's tail.
  0 is the library default: the depth from the staging budget plus overflow
  pages for the particles near their end.
* ``block_size`` -- lanes per workgroup of the kinematics kernel.

``tune`` times a short probe (the first output interval(s) of the real
workload, HIP events on the library's stream) for every candidate shape and
returns the fastest.  Results can be cached in a JSON file keyed by the
workload's signature.
"""
from __future__ import annotations

import itertools
import json
import os

from . import _lib as L
from .options import output_times
from .particle import Particle

# 0: the library's default staging (budget-derived depth plus overflow pages)
DEFAULT_CHUNKS = (0, 256, 512, 1024)
DEFAULT_BLOCKS = (128, 256)


def _signature(cfgs, variant, probe_targets) -> str:
    c0 = cfgs[0]
    return (f"n={sum(int(c.numParticles) for c in cfgs)};cfg={len(cfgs)};integ={int(c0.integratorType)};"
            f"var={int(variant)};gas={int(c0.gas_coll)};t={probe_targets[-1]:.6g}")


def tune(cfgs, variant: int = L.VARIANT_FAST, stop_at=None, pulse_axis: int = 2, probe_intervals: int = 1,
         chunks=DEFAULT_CHUNKS, blocks=DEFAULT_BLOCKS, max_buffer_gb: float = 64.0, cache: str | None = None) -> dict:
    """Fastest (chunk_records, block_size) for a plain (one options block) or
    batched (a list) workload; returns dict(chunk_records, block_size,
    timings={(chunk, block): ms})."""
    cfgs = list(cfgs) if isinstance(cfgs, (list, tuple)) else [cfgs]
    targets = output_times(cfgs[0])[:max(1, probe_intervals)]
    key = _signature(cfgs, variant, targets)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            hit = json.load(f).get(key)
        if hit:
            return dict(chunk_records=hit["chunk_records"], block_size=hit["block_size"], timings={}, cached=True)
    n = sum(int(c.numParticles) for c in cfgs)
    timings = {}
    for chunk, block in itertools.product(chunks, blocks):
        if chunk * n * 64 > max_buffer_gb * 2 ** 30:
            continue
        kw = dict(variant=variant, pulse_axis=pulse_axis, block_size=block, chunk_records=chunk)
        p = Particle.batch(cfgs, **kw) if len(cfgs) > 1 else Particle(cfgs[0], **kw)
        try:
            p.initParticles()
            if stop_at is not None:
                p.set_stop_times(stop_at)
            p.synchronize()
            p.timer_start()
            p.run(targets)
            timings[(chunk, block)] = p.timer_stop()
        finally:
            p.close()
    best = min(timings, key=timings.get)
    out = dict(chunk_records=best[0], block_size=best[1], timings=timings, cached=False)
    if cache:
        data = {}
        if os.path.exists(cache):
            with open(cache) as f:
                data = json.load(f)
        data[key] = dict(chunk_records=best[0], block_size=best[1])
        with open(cache, "w") as f:
            json.dump(data, f, indent=1)
    return out
