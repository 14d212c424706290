"""Equivalent of the reference's run loop ``mainAnalysis`` (src/simulation.cpp:8-97).

    derive tc and sqrtKT_m; create the logger (writes the options header);
    particle p(opt); p.initParticles(); p.outputData(log);
    for i in 0 .. floor((tf-t0)/ioutInt)-1:
        p.runSimulation((i+1)*ioutInt); p.outputData(log)
        print "iter i, duration nextTime, ms"
    p.postProcess(log)          # a spectral run appends b_end

``main_analysis`` keeps that per-interval structure (the drop-in path).
``main_analysis_persistent`` produces the same file from one persistent launch
(st_run) that writes every snapshot block on the GPU.
"""
from __future__ import annotations

import sys
import time

from . import _lib as L
from .logger import BinaryLogger
from .options import derive, output_times
from .particle import Particle


def main_analysis(opt, output_name: str, pulse_axis: int = 2, variant: int = L.VARIANT_EXACT,
                  tc_override=None, log=sys.stdout) -> None:
    opt = derive(type(opt).from_buffer_copy(bytes(opt)), tc_override)
    with BinaryLogger(opt, output_name) as lg, Particle(opt, pulse_axis, variant) as p:
        p.initParticles()
        p.outputData(lg)
        for i, next_t in enumerate(output_times(opt)):
            t0 = time.perf_counter()
            p.runSimulation(next_t)
            p.outputData(lg)
            ms = int((time.perf_counter() - t0) * 1000)
            print(f"iter {i}, duration {next_t:g}, {ms}", file=log)
        p.postProcess(lg)


def main_analysis_persistent(opt, output_name: str, pulse_axis: int = 2,
                             variant: int = L.VARIANT_FAST, tc_override=None) -> None:
    opt = derive(type(opt).from_buffer_copy(bytes(opt)), tc_override)
    targets = output_times(opt)
    with BinaryLogger(opt, output_name) as lg, Particle(opt, pulse_axis, variant) as p:
        p.initParticles()
        bb = p.block_bytes()
        blocks = L.DeviceBuffer((len(targets) + 1) * bb)
        p.snapshot_block_device(blocks.ptr)
        p.run(targets, blocks.ptr + bb)
        host = p.copy_to_host(blocks.ptr, blocks.nbytes)
        blocks.free()
        for k in range(len(targets) + 1):
            lg.writeBlock(host[k * bb:(k + 1) * bb])
        p.postProcess(lg)
