"""Gradients of ensemble spin observables with respect to the field
parameters, by the adjoint kernel (st_adjoint).

theta = (Gx, Gy, Gz, E, B0): the nine gradient coefficients dB_i/dx_j (rows
Gx, Gy, Gz of the reference's grad hook, src/integrator.cpp:44-46), the
electric field and the static field -- 15 numbers, the order of
``GRAD_PARAMS``.

    run = ForwardRun(opt)                 # checkpointed forward, blocks kept in HBM
    means = run.run()                     # (K, 3) ensemble-mean spin at the output times
    L, dL_dmeans = phase_loss(means, run.times, opt.t0)
    grad = run.gradient(dL_dmeans)        # (15,) dL/dtheta
    run.close()
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from ._lib import GRAD_PARAMS, Options
from .options import output_times
from .particle import Particle

__all__ = ["GRAD_PARAMS", "ForwardRun", "params_of", "with_params", "phase_loss", "unwrap"]


def params_of(opt: Options) -> np.ndarray:
    return np.array([*opt.Gx, *opt.Gy, *opt.Gz, *opt.E, *opt.B0], dtype=np.float64)


def with_params(opt: Options, theta) -> Options:
    o = Options.from_buffer_copy(bytes(opt))
    t = np.asarray(theta, dtype=np.float64).reshape(15)
    o.Gx = L.Coords(*t[0:3])
    o.Gy = L.Coords(*t[3:6])
    o.Gz = L.Coords(*t[6:9])
    o.E = L.Coords(*t[9:12])
    o.B0 = L.Coords(*t[12:15])
    return o


class ForwardRun:
    """One checkpointed forward run of a plain handle: all output intervals
    in one st_run, snapshot blocks kept in device memory for the adjoint.
    Checkpoints are taken every ``chunk_records`` trajectory segments."""

    def __init__(self, opt: Options, variant: int = L.VARIANT_FAST, chunk_records: int = 256,
                 pulse_axis: int = 2, adjoint_window: int = 0):
        self.opt = Options.from_buffer_copy(bytes(opt))
        self.times = np.asarray(output_times(self.opt), dtype=np.float64)
        self.p = Particle(self.opt, pulse_axis=pulse_axis, variant=variant, chunk_records=chunk_records,
                          adjoint_window=adjoint_window)
        self.blocks = None

    def run(self) -> np.ndarray:
        p = self.p
        p.set_checkpointing(True)
        p.initParticles()
        bb = p.block_bytes()
        self.blocks = L.DeviceBuffer(bb * len(self.times))
        p.run(self.times, self.blocks.ptr)
        sums = p.block_sums(self.blocks.ptr, len(self.times))[:, 0, :]
        if np.any(sums[:, 0] != p.n):
            raise L.SpinTrackError("particles failed in the forward run; the adjoint needs a clean run")
        return sums[:, 1:] / sums[:, :1]

    def gradient(self, dL_dmeans) -> np.ndarray:
        return self.p.adjoint(self.blocks.ptr, np.asarray(dL_dmeans, dtype=np.float64))[0]

    def close(self) -> None:
        if self.blocks is not None:
            self.blocks.free()
            self.blocks = None
        if self.p is not None:
            self.p.close()
            self.p = None


def unwrap(phi: np.ndarray) -> np.ndarray:
    """np.unwrap: consecutive differences brought into [-pi, pi)."""
    return np.unwrap(phi)


def phase_loss(means: np.ndarray, times: np.ndarray, t0: float):
    """L = <Sx>(t_K) + Phi, with Phi the accumulated precession phase of the
    ensemble-mean transverse spin: phi_k = atan2(<Sz>, <Sy>)(t_k),
    unwrapped, fitted by least squares phi_k = a + w*t_k, Phi = w*(t_K - t0).
    Returns (L, dL/dmeans) with dL/dmeans of shape (K, 3)."""
    m = np.asarray(means, dtype=np.float64)
    t = np.asarray(times, dtype=np.float64)
    phi = unwrap(np.arctan2(m[:, 2], m[:, 1]))
    dt = t - t.mean()
    den = float(np.sum(dt * dt))
    w = float(np.sum(dt * (phi - phi.mean()))) / den
    span = float(t[-1] - t0)
    Lval = float(m[-1, 0] + w * span)
    g = np.zeros_like(m)
    c = span * dt / den                      # dPhi/dphi_k
    rho2 = m[:, 1] ** 2 + m[:, 2] ** 2
    g[:, 1] = -c * m[:, 2] / rho2
    g[:, 2] = c * m[:, 1] / rho2
    g[-1, 0] += 1.0
    return Lval, g
