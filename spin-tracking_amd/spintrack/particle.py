"""Host-side mirror of the reference's ``particle`` class (include/particle.h:87-314).

    p = Particle(opt)            # particle p(opt)           -> st_create
    p.initParticles()            # p.initParticles()         -> st_init
    p.runSimulation(next_t)      # p.runSimulation(nextTOut) -> st_advance_to
    p.outputData(logger)         # p.outputData(log)         -> st_snapshot + writeSnapshot

All particle state stays in HBM; only the snapshot / state copy-outs move
data to the host.  ``run(targets, blocks)`` is the persistent multi-interval
form (st_run), which produces the same per-particle results as calling
``runSimulation`` for every target in turn.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import Ext, Options, ParticleState, check

COORDS = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8")])
SNAPSHOT_FIELDS = ("t", "pos", "v", "S", "failure", "n_coll", "n_bounce", "n_steps")

STATE_DTYPE = np.dtype({
    "names": ["S", "v", "v_old", "pos", "pos_old", "t", "t_old", "tf", "next_gas", "h",
              "rng_x", "rng_y", "rng_z", "rng_w", "spare", "hasSpare",
              "n_bounce", "n_coll", "n_steps", "failure", "part_id", "stop", "coll_type",
              "wall_hit"],
    "formats": [COORDS] * 5 + ["<f8"] * 5 + ["<u8"] * 4 + ["<f8", "u1"] + ["<u8"] * 3
               + ["<i4", "<u4", "u1", "u1", "u1"],
    "offsets": [0, 24, 48, 72, 96, 120, 128, 136, 144, 152, 160, 168, 176, 184, 192, 200,
                208, 216, 224, 232, 236, 240, 241, 242],
    "itemsize": C.sizeof(ParticleState),
})


def snapshot_block_dtype(n: int) -> np.dtype:
    """One snapshot block as BinaryLogger::writeSnapshot lays it out
    (src/logger.cpp:14-25): eight consecutive arrays of n entries."""
    return np.dtype({
        "names": list(SNAPSHOT_FIELDS),
        "formats": [("<f8", (n,)), (COORDS, (n,)), (COORDS, (n,)), (COORDS, (n,)),
                    ("<i4", (n,)), ("<u8", (n,)), ("<u8", (n,)), ("<u8", (n,))],
        "offsets": [0, 8 * n, 32 * n, 56 * n, 80 * n, 84 * n, 92 * n, 100 * n],
        "itemsize": 108 * n,
    })


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Particle:
    """GPU-resident ensemble of ``opt.numParticles`` particles."""

    def __init__(self, opt: Options, pulse_axis: int = 2, variant: int = L.VARIANT_EXACT,
                 block_size: int = 0, engine: int = L.ENGINE_PIPELINE, seg_buffer_mb: int = 0,
                 chunk_records: int = 0, graph_chunks: int = 0, spectrum_sum: int = 0, max_events: int = 0,
                 library: str = None, adjoint_window: int = 0):
        self.lib = L.open_library(library) if library else L.load()
        self.opt = Options.from_buffer_copy(bytes(opt))
        self.n = int(opt.numParticles)
        self.ext = Ext(int(pulse_axis), int(variant), int(block_size), int(engine), int(seg_buffer_mb),
                       int(chunk_records), int(graph_chunks), int(spectrum_sum), int(max_events),
                       int(adjoint_window))
        h = C.c_void_p()
        check(self.lib.st_create(C.byref(self.opt), C.byref(self.ext), C.byref(h)), None, self.lib)
        self.h = h

    @classmethod
    def batch(cls, cfgs, pulse_axis: int = 2, variant: int = L.VARIANT_EXACT, block_size: int = 0,
              seg_buffer_mb: int = 0, chunk_records: int = 0, graph_chunks: int = 0) -> "Particle":
        """Batched handle (st_create_batch): the configurations ``cfgs`` (each
        an Options block with its own numParticles) back to back in one
        ensemble.  ``layout()`` gives the particle offsets."""
        self = cls.__new__(cls)
        self.lib = L.load()
        cfgs = list(cfgs)
        arr = (Options * len(cfgs))(*[Options.from_buffer_copy(bytes(c)) for c in cfgs])
        self.opt = Options.from_buffer_copy(bytes(cfgs[0]))
        self.n = int(sum(int(c.numParticles) for c in cfgs))
        self.opt.numParticles = self.n
        self.ext = Ext(int(pulse_axis), int(variant), int(block_size), L.ENGINE_PIPELINE, int(seg_buffer_mb),
                       int(chunk_records), int(graph_chunks))
        h = C.c_void_p()
        check(self.lib.st_create_batch(arr, len(cfgs), C.byref(self.ext), C.byref(h)), None, self.lib)
        self.h = h
        return self

    def layout(self) -> np.ndarray:
        """Particle offsets of the configurations (n_cfg + 1 entries)."""
        n = C.c_int32()
        check(self.lib.st_batch_layout(self.h, C.byref(n), None), self.h, self.lib)
        off = np.zeros(n.value + 1, np.uint32)
        check(self.lib.st_batch_layout(self.h, None, _ptr(off)), self.h, self.lib)
        return off

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.st_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- the reference's interface -----------------------------------------
    def initParticles(self, first: int = 0, stride: int = 1) -> None:
        """include/particle.h:241-252 / src/particle.cpp:638-710."""
        check(self.lib.st_init_subset(self.h, first, stride), self.h, self.lib)

    def runSimulation(self, next_t: float) -> None:
        """src/particle.cpp:865-876: advance every particle to next_t."""
        check(self.lib.st_advance_to(self.h, float(next_t)), self.h, self.lib)

    def snapshot(self) -> dict:
        """The eight arrays of one snapshot, in the reference's types."""
        n = self.n
        out = dict(t=np.empty(n), pos=np.empty(n, COORDS), v=np.empty(n, COORDS),
                   S=np.empty(n, COORDS), failure=np.empty(n, np.int32),
                   n_coll=np.empty(n, np.uint64), n_bounce=np.empty(n, np.uint64),
                   n_steps=np.empty(n, np.uint64))
        check(self.lib.st_snapshot(self.h, *[_ptr(out[k]) for k in SNAPSHOT_FIELDS]), self.h, self.lib)
        return out

    def outputData(self, logger) -> None:
        """src/particle.cpp:852-862."""
        s = self.snapshot()
        logger.writeSnapshot(*[s[k] for k in SNAPSHOT_FIELDS])

    # -- extensions ---------------------------------------------------------
    def run(self, targets, blocks_dev_ptr: int = 0) -> None:
        """Persistent multi-interval run (st_run)."""
        t = np.ascontiguousarray(targets, dtype=np.float64)
        check(self.lib.st_run(self.h, len(t), _ptr(t), blocks_dev_ptr or None), self.h, self.lib)

    def run_noalloc(self, targets, blocks_dev_ptr: int = 0) -> None:
        """st_run without device allocation (static scratch; at most
        NOALLOC_MAX_TARGETS output times)."""
        t = np.ascontiguousarray(targets, dtype=np.float64)
        check(self.lib.st_run_noalloc(self.h, len(t), _ptr(t), blocks_dev_ptr or None), self.h, self.lib)

    def set_stop_times(self, stop_at) -> None:
        if stop_at is None:
            check(self.lib.st_set_stop_times(self.h, None), self.h, self.lib)
            return
        s = np.ascontiguousarray(stop_at, dtype=np.float64)
        assert s.shape == (self.n,)
        check(self.lib.st_set_stop_times(self.h, _ptr(s)), self.h, self.lib)

    def block_bytes(self) -> int:
        return int(self.lib.st_snapshot_block_bytes(self.h))

    def snapshot_block_device(self, dst_ptr: int) -> None:
        check(self.lib.st_snapshot_block_device(self.h, dst_ptr), self.h, self.lib)

    def get_state(self) -> np.ndarray:
        out = np.zeros(self.n, STATE_DTYPE)
        check(self.lib.st_get_state(self.h, _ptr(out)), self.h, self.lib)
        return out

    def set_state(self, state: np.ndarray) -> None:
        s = np.ascontiguousarray(state, dtype=STATE_DTYPE)
        assert s.shape == (self.n,)
        check(self.lib.st_set_state(self.h, _ptr(s)), self.h, self.lib)

    def work_counters(self) -> dict:
        w = np.zeros(6, np.uint64)
        check(self.lib.st_work_counters(self.h, _ptr(w)), self.h, self.lib)
        return dict(accepted=int(w[0]), rejected=int(w[1]), segments=int(w[2]), trips=int(w[3]),
                    events=int(w[4]), kin_records=int(w[5]))

    def particle_counters(self) -> dict:
        a, r, s = (np.zeros(self.n, np.uint64) for _ in range(3))
        check(self.lib.st_particle_counters(self.h, _ptr(a), _ptr(r), _ptr(s)), self.h, self.lib)
        return dict(accepted=a, rejected=r, segments=s)

    def set_trace(self, particle: int, capacity: int = 100000) -> None:
        """Record every integrated segment of one particle (debug tracing)."""
        check(self.lib.st_set_trace(self.h, int(particle), int(capacity)), self.h, self.lib)
        self._trace_cap = int(capacity)

    def get_trace(self) -> np.ndarray:
        """(records, 20) array; columns are spintrack._lib.TRACE_FIELDS."""
        cap = getattr(self, "_trace_cap", 0)
        out = np.zeros((max(cap, 1), L.TRACE_DOUBLES))
        n = C.c_uint32()
        check(self.lib.st_get_trace(self.h, _ptr(out), cap, C.byref(n)), self.h, self.lib)
        return out[: n.value]

    def copy_to_host(self, dev_ptr: int, nbytes: int) -> np.ndarray:
        out = np.empty(int(nbytes), np.uint8)
        check(self.lib.st_memcpy_d2h(self.h, _ptr(out), C.c_void_p(int(dev_ptr)), int(nbytes)), self.h, self.lib)
        return out

    def set_checkpointing(self, enable: bool = True) -> None:
        check(self.lib.st_set_checkpointing(self.h, 1 if enable else 0), self.h, self.lib)

    def adjoint(self, blocks_dev_ptr: int, seeds) -> np.ndarray:
        """Gradient (n_cfg, 15) of a loss of the ensemble-mean spins, given
        seeds[c, k, :] = dL/d<S>(t_k) of the last checkpointed run."""
        n_cfg = len(self.layout()) - 1
        s = np.ascontiguousarray(seeds, dtype=np.float64).reshape(n_cfg, -1, 3)
        out = np.zeros((n_cfg, 15))
        check(self.lib.st_adjoint(self.h, C.c_void_p(int(blocks_dev_ptr)), _ptr(s), _ptr(out)), self.h, self.lib)
        return out

    def block_sums(self, blocks_dev_ptr: int, n_blocks: int) -> np.ndarray:
        """(n_blocks, n_cfg, 4): valid count and sums of Sx, Sy, Sz per
        snapshot block and configuration."""
        n_cfg = len(self.layout()) - 1
        out = np.zeros((int(n_blocks), n_cfg, 4))
        check(self.lib.st_block_sums(self.h, C.c_void_p(int(blocks_dev_ptr)), int(n_blocks), _ptr(out)), self.h, self.lib)
        return out

    def phase_stats(self, blocks_dev_ptr: int, n_blocks: int, phi_dev_ptr: int = 0) -> np.ndarray:
        """(n_blocks, n_cfg, 6): the reference's per-particle phase analysis
        per snapshot block and configuration -- mean phase (wrapped), std,
        phases kept after the 5-sigma cut, phases selected, reference angle
        of the histogram alignment, mean before wrapping."""
        n_cfg = len(self.layout()) - 1
        out = np.zeros((int(n_blocks), n_cfg, 6))
        check(self.lib.st_phase_stats(self.h, C.c_void_p(int(blocks_dev_ptr)), int(n_blocks), _ptr(out),
                                      C.c_void_p(int(phi_dev_ptr)) if phi_dev_ptr else None), self.h, self.lib)
        return out

    def copy_device(self, dst_ptr: int, src_ptr: int, nbytes: int) -> None:
        check(self.lib.st_memcpy_d2d(self.h, C.c_void_p(int(dst_ptr)), C.c_void_p(int(src_ptr)),
                                     int(nbytes)), self.h, self.lib)

    def kernel_timing(self, enable: bool = True) -> None:
        check(self.lib.st_kernel_timing(self.h, 1 if enable else 0), self.h, self.lib)

    def kernel_times(self) -> dict:
        ms = np.zeros(2)
        n = np.zeros(2, np.uint64)
        check(self.lib.st_kernel_times(self.h, _ptr(ms), _ptr(n)), self.h, self.lib)
        return dict(kin_ms=float(ms[0]), spin_ms=float(ms[1]), kin_launches=int(n[0]),
                    spin_launches=int(n[1]))

    def timer_start(self) -> None:
        check(self.lib.st_timer_start(self.h), self.h, self.lib)

    def timer_stop(self) -> float:
        """Milliseconds since timer_start, measured with events on this stream."""
        ms = C.c_float()
        check(self.lib.st_timer_stop(self.h, C.byref(ms)), self.h, self.lib)
        return float(ms.value)

    def set_stream(self, stream_ptr: int) -> None:
        check(self.lib.st_set_stream(self.h, stream_ptr or None), self.h, self.lib)

    def synchronize(self) -> None:
        check(self.lib.st_synchronize(self.h), self.h, self.lib)

    def spinMean(self) -> np.ndarray:
        """particle::spinMean (src/particle.cpp:878-884): the mean spin over
        all particles (a stopped particle's NaN spin makes it NaN)."""
        S = np.zeros(self.n, dtype=COORDS)
        check(self.lib.st_snapshot(self.h, None, None, None, _ptr(S), None, None, None, None), self.h, self.lib)
        # the reference's order, S_sum = S_sum + S[i] (cumsum adds in sequence)
        return np.cumsum(xyz(S), axis=0)[-1] / self.n

    # ---- spectral path (integratorType 6) ----
    def spectrum(self) -> dict:
        """particle::cspec: the summed, un-normalised covariance spectrum."""
        sp = L.Spectrum()
        check(self.lib.st_spectrum_get(self.h, C.byref(sp)), self.h, self.lib)
        v = np.ctypeslib.as_array(sp.variance).reshape(L.NW, 3, 3, 2)
        return dict(frequencies=np.array(sp.frequencies[:]), dt=sp.dt, n_samples=int(sp.n_samples),
                    variance=v[..., 0] + 1j * v[..., 1], raw=v.copy())

    def spectrumAggregators(self) -> dict:
        """particle::getSpectrumAggregators: the last interval's filter state
        s1, s2, imag_diag (n, NW, 3) and sample counts."""
        agg = np.zeros((self.n, L.AGG_FIELDS))
        ns = np.zeros(self.n, dtype=np.uint32)
        check(self.lib.st_spectrum_aggregators(self.h, _ptr(agg), _ptr(ns)), self.h, self.lib)
        a = agg.reshape(self.n, 3, L.NW, 3)
        return dict(s1=a[:, 0].copy(), s2=a[:, 1].copy(), im=a[:, 2].copy(), n_samples=ns)

    def floquetInfo(self) -> dict:
        """particle::fd (floquet_diagonalize of the run's drive)."""
        f = L.Floquet()
        check(self.lib.st_floquet_info(self.h, C.byref(f)), self.h, self.lib)
        from .floquet import floquet_dict
        return floquet_dict(f)

    def floquetResults(self) -> np.ndarray:
        """particle::floquetResults (src/particle.cpp:841-843): b_end."""
        b = np.zeros(3)
        check(self.lib.st_floquet_result(self.h, _ptr(b)), self.h, self.lib)
        return b

    def postProcess(self, logger) -> None:
        """particle::postProcess (src/particle.cpp:845-850): a spectral run
        appends b_end to the output."""
        if int(self.opt.integratorType) == 6:
            logger.writeSingle("b_end", self.floquetResults())


def xyz(a: np.ndarray) -> np.ndarray:
    """COORDS structured array -> (n, 3) float64."""
    return np.stack([a["x"], a["y"], a["z"]], axis=-1)
