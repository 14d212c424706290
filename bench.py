#!/usr/bin/env python3
"""Throughput benchmark of the spin-tracking hot path on one MI355X.

Workload (the plain-form throughput case): 1,048,576 He-3 atoms, Maxwell-
Boltzmann velocities at 0.45 K, gas collisions with mean free time 2e-3 s,
5% diffuse walls, gravity, B0 = (3e-6,0,0) T, E = (0,0,7.5e6) V/m,
dBx/dz = 1e-8 T/m, DOP853 at rtol = atol = 1e-12, FP64, output every 0.5 s.

One "step" = one output interval: every particle advanced by 0.5 s and its
snapshot block (the reference's file layout) written to HBM.  --steps K
times K consecutive intervals (K = 20 is the full 10 s run); --warmup W runs
W intervals on a separate, freshly seeded ensemble first.  Inputs (the
seeded ensemble) are resident in HBM before the timed region starts.

Prints ONE JSON line with the metric, a roofline object for the dominant
kernel (the persistent lane engine) and the CPU oracle timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spin-tracking_amd"))

# Algorithmic FP64 operations per unit of work (each add, sub, mul, div,
# sqrt or libm call counts 1), for the reference formulation of the path:
# per accepted DOP853 step (12 right-hand-side evaluations plus stage
# combinations, error norm and step control), per rejected step, and per
# event (collision time, move/reflect, sanity check).
FLOPS_ACC_GRAV = 1410.0
FLOPS_ACC_NOGRAV = 1200.0
FLOPS_REJ = 1340.0
FLOPS_EVENT = 75.0
FLOPS_RK45_TRIAL = 650.0
# per trial of integratorTypes 2-5, counted the same way from the reference
# formulation (gamma*B along a segment with gravity: 67 ops; rotation
# quaternion from a vector: 18; quaternion product: 28; Rodrigues rotation of
# a vector: 42; one CFET exponent: 30): MagnusCFET = 5 field samples + 16
# exponents + 16 rotations + controller; RK45/RKF45 quaternion = 6 samples +
# 11 rotation quaternions + 9 products + 2 vector rotations + controller;
# RKF45 = RK45's count.
FLOPS_TRIAL = {1: 650.0, 2: 1520.0, 3: 1000.0, 4: 650.0, 5: 1000.0}
# integratorType 6, per noise sample (the reference's operation order, test
# noise off): the noise field along a gravity segment 55 FP64 operations, 17
# Goertzel filters x 3 axes x 11, the sample clock 1
FLOPS_SAMPLE = 617.0
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector, dense
HBM_PEAK_GBS = 8000.0
# HBM bytes per segment record moved by k_spin (reads + writes), from the
# FETCH_SIZE / WRITE_SIZE counters of the pipeline on this workload, scaled by
# a 1 GiB device copy (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000); the
# algorithmic figure is 64 B per record plus the per-particle spin state.
SPIN_HBM_BYTES_PER_RECORD = 74.5
# The same kernel's executed FP64 rate from the hardware counters
# (SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64, an FMA counted as 2 operations) over
# its rocprofv3 duration on this workload; the fast variant executes fewer
# FP64 operations per step than the reference formulation above counts.
HW_COUNTED = {"frac_fp64_peak": 0.471, "TFLOPs": 37.0,
              "source": "profiles/k_spin_o3_k_kin_fast_dop853_he3_1M_2intervals.md"}


def build_workload(n, integrator):
    from spintrack import configs
    return configs.he3_thermal(n=n, tf=10.0, ioutInt=0.5, seed=1, integrator=integrator)


def algorithmic_flops(opt, w):
    if opt.integratorType in FLOPS_TRIAL:
        return FLOPS_TRIAL[opt.integratorType] * (w["accepted"] + w["rejected"]) + FLOPS_EVENT * w["segments"]
    fa = FLOPS_ACC_GRAV if opt.gravity else FLOPS_ACC_NOGRAV
    return fa * w["accepted"] + FLOPS_REJ * w["rejected"] + FLOPS_EVENT * w["segments"]


def run_gpu(opt, args, targets, persistent, variant, engine):
    """Seeds an ensemble in HBM, then times the K output intervals with HIP
    events recorded on the stream the library launches on."""
    import spintrack as st

    kw = dict(variant=variant, engine=engine, seg_buffer_mb=args.seg_buffer_mb)
    if engine == st._lib.ENGINE_PIPELINE and not args.no_autotune:
        # launch shape from the autotuner (probe: the first output interval)
        from spintrack import autotune
        shape = autotune.tune(opt, variant=variant, probe_intervals=2)
        kw.update(chunk_records=shape["chunk_records"], block_size=shape["block_size"])
        args.tuned_shape = {"chunk_records": shape["chunk_records"], "kin_block": shape["block_size"],
                            "probe_ms": {f"{c}x{b}": round(v, 2) for (c, b), v in shape["timings"].items()}}
    # warm-up on its own ensemble (first-launch costs, clocks)
    if args.warmup > 0:
        wopt = type(opt).from_buffer_copy(bytes(opt))
        wopt.seed = opt.seed + 7919
        with st.Particle(wopt, **kw) as w:
            w.initParticles()
            wt = targets[: args.warmup]
            if persistent:
                w.run(wt)
            else:
                for t in wt:
                    w.runSimulation(t)
            w.synchronize()
    p = st.Particle(opt, **kw)
    p.initParticles()
    if args.copy_calibration_mb and persistent:
        # one device-to-device copy of known size: the reference dispatch
        # that scales the FETCH_SIZE / WRITE_SIZE counters to bytes
        nb = args.copy_calibration_mb << 20
        a, b = st._lib.DeviceBuffer(nb), st._lib.DeviceBuffer(nb)
        p.copy_device(b.ptr, a.ptr, nb)
        p.synchronize()
        a.free()
        b.free()
    bb = p.block_bytes()
    blocks = st._lib.DeviceBuffer(len(targets) * bb)
    p.synchronize()
    if engine == st._lib.ENGINE_PIPELINE:
        p.kernel_timing(True)
    t0 = time.perf_counter()
    p.timer_start()
    if persistent:
        p.run(targets, blocks.ptr)
    else:
        for k, t in enumerate(targets):
            p.runSimulation(t)
            p.snapshot_block_device(blocks.ptr + k * bb)
    kernel_ms = p.timer_stop()
    wall = time.perf_counter() - t0
    kt = p.kernel_times() if engine == st._lib.ENGINE_PIPELINE else None
    w = p.work_counters()
    fails = int((p.snapshot()["failure"] != 0).sum())
    blocks.free()
    p.close()
    return dict(kernel_ms=kernel_ms, wall_s=wall, work=w, failures=fails, ktimes=kt)


def bench_sweep(args, variant):
    """Field-configuration sweep: 64 configurations x 16,384 neutrons with
    their own gradients, Ez polarity and wall model, and per-particle
    durations uniform in [1, 20] s.  Batched = the whole sweep in one
    handle and one st_run call (persistent kernels over all configurations);
    plain = the drop-in path as the reference's loop drives it, one
    configuration after the other, st_advance_to per output interval.  Both
    are timed with HIP events on the library's stream; particle-seconds =
    the sum of the per-particle durations."""
    import numpy as np
    import spintrack as st

    cfgs, stop_at = st.configs.batched_sweep(n_cfg=args.sweep_configs, n_per=args.sweep_per)
    targets = st.output_times(cfgs[0])
    psec = float(np.sum(stop_at - cfgs[0].t0))
    # warm-up: load both kernel families once
    wc, ws = st.configs.batched_sweep(n_cfg=2, n_per=1024, tf=1.0, ioutInt=0.5, dur=(0.5, 1.0))
    with st.Particle.batch(wc, variant=variant) as w:
        w.initParticles()
        w.set_stop_times(ws)
        w.run(st.output_times(wc[0]))
        w.synchronize()
    with st.Particle(wc[0], variant=variant) as w:
        w.initParticles()
        w.runSimulation(0.5)
        w.synchronize()

    # launch shape (chunk depth, kinematics lanes per block) from the autotuner,
    # probing the first output interval of this workload
    from spintrack import autotune
    shape = autotune.tune(cfgs, variant=variant, stop_at=stop_at)
    b = st.Particle.batch(cfgs, variant=variant, chunk_records=shape["chunk_records"],
                          block_size=shape["block_size"])
    b.initParticles()
    b.set_stop_times(stop_at)
    b.synchronize()
    b.kernel_timing(True)
    b.timer_start()
    b.run(targets)
    batched_ms = b.timer_stop()
    kt = b.kernel_times()
    wk = b.work_counters()
    fails = int((b.snapshot()["failure"] != 0).sum())
    b.close()
    spin_flops = FLOPS_ACC_GRAV * wk["accepted"] + FLOPS_REJ * wk["rejected"]
    spin_tf = spin_flops / (kt["spin_ms"] / 1e3) / 1e12

    plain_ms = 0.0
    off = np.cumsum([0] + [c.numParticles for c in cfgs])
    for c, cfg in enumerate(cfgs):
        with st.Particle(cfg, variant=variant) as p:
            p.initParticles()
            p.set_stop_times(stop_at[off[c]:off[c + 1]])
            p.synchronize()
            p.timer_start()
            for t in targets:
                p.runSimulation(t)
            plain_ms += p.timer_stop()
    batched = psec / (batched_ms / 1e3)
    plain = psec / (plain_ms / 1e3)
    return {"workload": f"{len(cfgs)} field configurations x {args.sweep_per} neutrons, durations U[1,20] s, "
                        f"random gradients in [-5e-8,5e-8] T/m, Ez +-7.5e6 V/m, half diffuse 0.1",
            "particle_seconds": psec, "batched_value": batched, "batched_ms": batched_ms,
            "plain_value": plain, "plain_ms": plain_ms, "batched_over_plain": batched / plain,
            "batched_spin_TFLOPs": spin_tf, "batched_spin_frac_fp64_peak": spin_tf / FP64_PEAK_TFLOPS,
            "batched_whole_run_TFLOPs": (spin_flops + FLOPS_EVENT * wk["segments"]) / (batched_ms / 1e3) / 1e12,
            "launch_shape": {"chunk_records": shape["chunk_records"], "kin_block": shape["block_size"],
                             "probe_ms": {f"{c}x{b_}": round(v, 3) for (c, b_), v in shape["timings"].items()}},
            "kernels": kt, "work": wk, "failures": fails}


def bench_rk45(args, variant):
    """The He-3 workload with integratorType 1 (the reference's RK45 with its
    stale-k1 first stage) at the same tolerance, over the first output
    interval(s): RK45 needs far more steps than DOP853 at rtol = 1e-12."""
    import spintrack as st

    opt = build_workload(args.rk45_particles, 1)
    targets = st.output_times(opt)[: args.rk45_steps]
    with st.Particle(opt, variant=variant) as w:   # warm-up
        w.initParticles()
        w.runSimulation(targets[0] * 0.01)
        w.synchronize()
    with st.Particle(opt, variant=variant) as p:
        p.initParticles()
        p.synchronize()
        p.kernel_timing(True)
        p.timer_start()
        p.run(targets)
        ms = p.timer_stop()
        kt = p.kernel_times()
        wk = p.work_counters()
    psec = args.rk45_particles * (targets[-1] - opt.t0)
    tf = FLOPS_RK45_TRIAL * (wk["accepted"] + wk["rejected"]) / (kt["spin_ms"] / 1e3) / 1e12
    return {"workload": f"{args.rk45_particles} He-3 thermal, RK45 rtol=atol=1e-12, {len(targets)} x {opt.ioutInt} s",
            "value": psec / (ms / 1e3), "ms": ms, "kernels": kt, "work": wk,
            "spin_TFLOPs": tf, "spin_frac_fp64_peak": tf / FP64_PEAK_TFLOPS}


# The reference's own published runs (notebook stdout of its GPU build on an
# unrecorded NVIDIA GPU; end-to-end totals in ms, init and output included):
# name -> (params beyond the shared block, total ms, citation)
REFERENCE_RUNS = {
    "magnus_neutrons_no_gravity": (
        {"numParticles": 65536, "diffuse": "0", "gravity": "false", "E": "75.0e5, 0, 0", "tf": "2000.0",
         "ioutInt": "20", "integratorType": "2"},
        27310, "scripts/ComparingVCrossENeutrons.ipynb:285 (params :93-121)"),
    "magnus_neutrons_gravity": (
        {"numParticles": 65536, "diffuse": "0", "gravity": "true", "E": "75.0e5, 0, 0", "tf": "2000.0",
         "ioutInt": "20", "integratorType": "2"},
        293964, "scripts/ComparingVCrossENeutrons.ipynb:1024"),
    "trajectory_1M_specular": (
        {"numParticles": 1000000, "diffuse": "0", "gravity": "true", "tf": "500.0", "ioutInt": "10.0",
         "integratorType": "-1"},
        11871, "scripts/PhysicsSanityChecks/EnergyConservationCheck.ipynb:192 (params :35-71)"),
    "trajectory_1M_diffuse": (
        {"numParticles": 1000000, "diffuse": "1.0", "gravity": "true", "tf": "500.0", "ioutInt": "10.0",
         "integratorType": "-1"},
        15130, "scripts/PhysicsSanityChecks/EnergyConservationCheck.ipynb:140"),
}
REFERENCE_COMMON = {"dist": "C", "V": "5.0", "T": "0.4", "gas_coll": "false", "B0": "3.0e-6, 0, 0",
                    "numPerGPUBlock": "256", "t0": "0.0", "a": "0", "w": "0", "h": "1e-5",
                    "keepStepSize": "false", "fixedStepSize": "false", "rtol": "1.0e-12", "atol": "1.0e-12",
                    "hmax": "1.0", "hmin": "1.0e-9", "seed": "0", "m": "1.674927e-27",
                    "gamma": "-1.832471850e8"}


def bench_reference_runs(args):
    """The reference's published GPU runs, repeated with the same params.txt
    through the shipped C++ driver (spintrack_main: parse, seed, advance per
    interval, write every snapshot to a binary file), timed end to end like
    the reference's own total.  The file goes to a temporary directory in
    memory (/dev/shm where there is one): on a disk, a run that follows
    another run's 5.4 GB file waits on that file's write-back (the diffuse
    trajectory run measured 2.5-5.7 s on disk against 0.93 s of kernels)."""
    import subprocess
    import tempfile

    exe = os.path.join(ROOT, "spin-tracking_amd", "spintrack_main")
    out = {}
    shm = "/dev/shm" if os.path.isdir("/dev/shm") and os.access("/dev/shm", os.W_OK) else None
    with tempfile.TemporaryDirectory(dir=shm) as d:
        for name, (extra, ref_ms, cite) in REFERENCE_RUNS.items():
            kv = dict(REFERENCE_COMMON, **{k: str(v) for k, v in extra.items()})
            params = os.path.join(d, name + ".txt")
            with open(params, "w") as f:
                f.write("".join(f"{k}, {v}\n" for k, v in kv.items()))
            def run(*flags):
                t0 = time.perf_counter()
                r = subprocess.run([exe, params, os.path.join(d, name + ".bin"), *flags], capture_output=True,
                                   text=True, timeout=900)
                wall = time.perf_counter() - t0
                if r.returncode != 0:
                    raise RuntimeError(f"{name}: driver failed: {r.stderr[-500:]}")
                os.remove(os.path.join(d, name + ".bin"))
                return float(r.stdout.strip().splitlines()[-1]), wall

            # the exact variant (the driver's default) is the like-for-like figure
            total_ms, wall = run()
            psec = int(kv["numParticles"]) * float(kv["tf"])
            out[name] = {"total_ms": total_ms, "wall_s": wall, "value": psec / (total_ms / 1e3),
                         "reference_total_ms": ref_ms, "reference_value": psec / (ref_ms / 1e3),
                         "speedup": ref_ms / total_ms, "reference_source": cite}
            if kv["integratorType"] != "-1":
                fast_ms, _ = run("--fast")
                out[name].update(fast_total_ms=fast_ms, fast_speedup=ref_ms / fast_ms)
            print(f"[bench] reference run {name}: {out[name]}", file=sys.stderr, flush=True)
    return {"note": "same params.txt as the reference's published runs (its GPU build on an unrecorded "
                    "NVIDIA GPU); end-to-end totals, output file included", "runs": out}


def bench_snapshot(args):
    """The snapshot stream of the per-interval path: k_snapshot turns the
    SoA particle state into one block of the reference's file layout (112 B
    of state read and 108 B written per particle), coordinate fields staged
    through LDS.  HIP events around 50 calls on the library's stream."""
    import spintrack as st

    n = args.particles
    opt = build_workload(n, 0)
    with st.Particle(opt, variant=st.VARIANT_FAST) as p:
        p.initParticles()
        bb = p.block_bytes()
        buf = st._lib.DeviceBuffer(bb)
        p.snapshot_block_device(buf.ptr)
        p.synchronize()
        reps = 50
        p.timer_start()
        for _ in range(reps):
            p.snapshot_block_device(buf.ptr)
        ms = p.timer_stop() / reps
        buf.free()
    nbytes = (112 + 108) * n
    gbs = nbytes / (ms / 1e3) / 1e9
    return {"kernel": "k_snapshot", "particles": n, "us_per_block": ms * 1e3, "algorithmic_bytes": nbytes,
            "GBs": gbs, "frac_hbm_peak": gbs / HBM_PEAK_GBS}


# integratorType -> (name, rtol, simulated seconds): tolerances each method
# can meet above hmin = 1e-8 (RKF45's "error" is the step's whole increment,
# RKF45Q's the sixth rotation factor)
INTEGRATOR_LEG = {2: ("MagnusCFET", 1e-12, 20.0), 3: ("RK45 quaternion", 1e-12, 20.0),
                  4: ("RKF45", 1e-4, 0.01), 5: ("RKF45 quaternion", 1e-4, 0.05)}


def bench_integrators(args, variant):
    """integratorTypes 2-5 on neutrons of the reference cell (5 m/s,
    specular walls, gravity, the same fields) with spins transverse to B0,
    one output interval of the simulated time listed per type; the CPU
    oracle on a 2048-particle sample of the same ensemble beside it."""
    import spintrack as st
    from spintrack import configs

    out = {}
    n = args.integrator_particles
    for integ, (name, rtol, tf) in INTEGRATOR_LEG.items():
        opt = configs.neutron_cell(n=n, tf=tf, ioutInt=tf)
        opt.integratorType = integ
        opt.rtol = opt.atol = rtol
        opt.yi.x, opt.yi.y, opt.yi.z = 0.0, 0.6, 0.8
        targets = st.output_times(opt)
        with st.Particle(opt, variant=variant) as w:   # warm-up
            w.initParticles()
            w.runSimulation(targets[0] * 0.01)
            w.synchronize()
        with st.Particle(opt, variant=variant) as p:
            p.initParticles()
            p.synchronize()
            p.kernel_timing(True)
            p.timer_start()
            p.run(targets)
            ms = p.timer_stop()
            kt = p.kernel_times()
            wk = p.work_counters()
        psec = n * (targets[-1] - opt.t0)
        trials = wk["accepted"] + wk["rejected"]
        tfl = FLOPS_TRIAL[integ] * trials / (kt["spin_ms"] / 1e3) / 1e12
        cpu = cpu_baseline(opt, n, 2048, 1)
        out[name] = {"integratorType": integ, "rtol": rtol, "simulated_s": tf, "value": psec / (ms / 1e3),
                     "ms": ms, "kernels": kt, "trials_per_particle_s": trials / psec,
                     "spin_TFLOPs": tfl, "spin_frac_fp64_peak": tfl / FP64_PEAK_TFLOPS,
                     "cpu_oracle": cpu["value"], "cpu_sample": cpu["sample"],
                     "gpu_over_cpu": psec / (ms / 1e3) / cpu["value"]}
        print(f"[bench] integrator {name}: {out[name]}", file=sys.stderr, flush=True)
    return {"workload": f"{n} neutrons, spins transverse to B0, one interval per type", "types": out}


def bench_spectral(args, variant):
    """integratorType 6 (noise spectrum + Floquet-Markov end point): He-3 of
    the headline workload (Maxwell-Boltzmann at 0.45 K, gas collisions with
    tc = 2e-3 s, 5% diffuse walls, gravity), drive 3e-5 T at 1 kHz along z,
    the noise field sampled every 1e-4 s into 17 Goertzel filters, 1 s in two
    output intervals, every particle's spectrum summed; the CPU oracle on a
    2048-particle sample of the first interval beside it."""
    import math

    import spintrack as st
    from spintrack import configs

    n = args.spectral_particles

    def options(m):
        o = configs.he3_thermal(n=m, tf=1.0, ioutInt=0.5, integrator=6)
        o.a, o.w, o.h = 3e-5, 2 * math.pi * 1000.0, 1e-4
        return o

    opt = options(n)
    targets = st.output_times(opt)
    with st.Particle(options(min(n, 65536)), spectrum_sum=1, variant=variant) as w:   # warm-up
        w.initParticles()
        w.runSimulation(0.01)
        w.synchronize()
    with st.Particle(opt, spectrum_sum=1, variant=variant) as p:
        p.initParticles()
        p.synchronize()
        p.kernel_timing(True)
        p.timer_start()
        p.run(targets)
        ms = p.timer_stop()
        kt = p.kernel_times()
        wk = p.work_counters()
        b_end = p.floquetResults()
        spec = p.spectrum()
    psec = n * (targets[-1] - opt.t0)
    samples = wk["accepted"]
    tfl = FLOPS_SAMPLE * samples / (kt["spin_ms"] / 1e3) / 1e12
    cpu = cpu_baseline(opt, n, 2048, 1)
    out = {"workload": f"{n} He-3, integratorType 6, h = 1e-4 s, 17 filters, 1 s in 2 intervals",
           "variant": args.variant,
           "value": psec / (ms / 1e3), "unit": "particle-seconds/s", "ms": ms, "kernels": kt,
           "samples": samples, "samples_per_s": samples / (ms / 1e3),
           "sampling_kernel_TFLOPs": tfl, "sampling_kernel_frac_fp64_peak": tfl / FP64_PEAK_TFLOPS,
           "flops_per_sample": FLOPS_SAMPLE, "spectrum_samples": spec["n_samples"],
           "b_end": [float(v) for v in b_end],
           "cpu_oracle": cpu["value"], "cpu_sample": cpu["sample"], "gpu_over_cpu": psec / (ms / 1e3) / cpu["value"]}
    print(f"[bench] spectral: {out}", file=sys.stderr, flush=True)
    return out


def bench_dressing(args, variant):
    """Spin dressing: 262,144 neutrons + 262,144 He-3 co-tracked, B0 =
    (1e-6,0,0) T plus 1e-5 T * cos(2 pi 3000 t) along y, gas collisions,
    2% diffuse walls, gravity, DOP853 rtol = atol = 1e-12, 2 s with an output
    every 0.1 s.  Batched = both species in one handle and one st_run call;
    plain = each species on its own handle, st_advance_to per interval."""
    import numpy as np
    import spintrack as st

    cfgs = st.configs.dressing_cotracked(n=args.dressing_per)
    targets = st.output_times(cfgs[0])
    psec = float(sum(c.numParticles for c in cfgs)) * (targets[-1] - cfgs[0].t0)
    b = st.Particle.batch(cfgs, variant=variant, pulse_axis=1)
    b.initParticles()
    bb = b.block_bytes()
    blocks = st._lib.DeviceBuffer(bb * len(targets))
    b.synchronize()
    b.kernel_timing(True)
    b.timer_start()
    b.run(targets, blocks.ptr)
    batched_ms = b.timer_stop()
    kt = b.kernel_times()
    wk = b.work_counters()
    sums = b.block_sums(blocks.ptr, len(targets))
    blocks.free()
    b.close()
    plain_ms = 0.0
    for c in cfgs:
        with st.Particle(c, variant=variant, pulse_axis=1) as p:
            p.initParticles()
            p.synchronize()
            p.timer_start()
            for t in targets:
                p.runSimulation(t)
            plain_ms += p.timer_stop()
    pol = np.hypot(sums[-1, :, 2], sums[-1, :, 3]) / np.maximum(sums[-1, :, 0], 1)
    tf = (FLOPS_ACC_GRAV * wk["accepted"] + FLOPS_REJ * wk["rejected"]) / (kt["spin_ms"] / 1e3) / 1e12
    return {"workload": f"{len(cfgs)} x {args.dressing_per} (neutrons, He-3) co-tracked, B0 1e-6 T, "
                        "1e-5 T cos(2 pi 3000 t) along y, gas, 2% diffuse, 2 s",
            "particle_seconds": psec, "batched_value": psec / (batched_ms / 1e3), "batched_ms": batched_ms,
            "plain_value": psec / (plain_ms / 1e3), "plain_ms": plain_ms,
            "batched_over_plain": plain_ms / batched_ms, "kernels": kt, "work": wk,
            "steps_per_particle_second": wk["accepted"] / psec,
            "batched_spin_TFLOPs": tf, "batched_spin_frac_fp64_peak": tf / FP64_PEAK_TFLOPS,
            "final_transverse_polarisation": {"neutron": float(pol[0]), "He-3": float(pol[1])}}


ADJ_PARAMS = list(range(9)) + [11, 12]   # the nine gradient coefficients, E.z, B0.x


def bench_adjoint(args, variant):
    """Differentiation workload: 131,072 neutrons, 50 s, outputs every 0.5 s,
    loss = <Sx>(50 s) + accumulated precession phase (least-squares fit over
    the 100 snapshots).  Times the plain forward run, the checkpointed
    forward run (kinematic checkpoints every 256 segments) and the adjoint
    pass (HIP events / host clock around st_adjoint); checks the gradient
    against central finite differences of the forward run (step 1e-3 of
    each parameter, and Richardson-extrapolated)."""
    import numpy as np
    import spintrack as st
    from spintrack import adjoint as A

    opt = st.configs.adjoint_case(n=args.adjoint_particles)
    times = np.asarray(st.output_times(opt))
    kw = dict(variant=variant, chunk_records=256)

    def forward(o, checkpoint):
        p = st.Particle(o, **kw)
        p.set_checkpointing(checkpoint)
        p.initParticles()
        buf = st._lib.DeviceBuffer(p.block_bytes() * len(times))
        p.synchronize()
        p.timer_start()
        p.run(times, buf.ptr)
        ms = p.timer_stop()
        sums = p.block_sums(buf.ptr, len(times))[:, 0, :]
        return p, buf, ms, sums

    def loss(o):
        p, buf, _, sums = forward(o, False)
        buf.free()
        p.close()
        return A.phase_loss(sums[:, 1:] / sums[:, :1], times, o.t0)[0]

    # warm-up (kernel loading)
    wo = st.configs.adjoint_case(n=2048, tf=1.0)
    r = A.ForwardRun(wo, variant=variant)
    r.gradient(A.phase_loss(r.run(), r.times, wo.t0)[1])
    r.close()

    p, buf, fwd_ms, _ = forward(opt, False)
    buf.free()
    p.close()
    p, buf, fwd_ck_ms, sums = forward(opt, True)
    if np.any(sums[:, 0] != opt.numParticles):
        raise RuntimeError("particle failures in the differentiation workload")
    Lval, seeds = A.phase_loss(sums[:, 1:] / sums[:, :1], times, opt.t0)
    t0 = time.perf_counter()
    grad = p.adjoint(buf.ptr, seeds)[0]
    adj_ms = (time.perf_counter() - t0) * 1e3
    buf.free()
    p.close()
    out = {"workload": f"{opt.numParticles} neutrons, 50 s, diffuse 0.01, gradients U[-1e-8,1e-8] T/m, "
                       "Ez 7.5e6 V/m, B0x 3e-6 T, loss <Sx>(T) + fitted precession phase over 100 snapshots",
           "loss": Lval, "forward_ms": fwd_ms, "forward_checkpointed_ms": fwd_ck_ms, "adjoint_ms": adj_ms,
           "adjoint_over_forward": adj_ms / fwd_ms,
           "gradient": {st._lib.GRAD_PARAMS[i]: float(grad[i]) for i in ADJ_PARAMS},
           "particle_seconds": opt.numParticles * (times[-1] - opt.t0)}
    if not args.no_adjoint_fd:
        # central differences of the adaptive-step forward run at two steps:
        # 1e-3 of each parameter (the stated check) and 1e-5; the loss carries
        # a ~1e-12 noise floor from accept/reject decisions that move with
        # theta, so weakly coupled parameters resolve only to that floor
        theta = A.params_of(opt)
        rel = {}
        for step in (1e-3,) + ((1e-5,) if args.adjoint_fd_ladder else ()):
            r = {}
            for i in ADJ_PARAMS:
                d = step * abs(theta[i])
                tp, tm = theta.copy(), theta.copy()
                tp[i] += d
                tm[i] -= d
                fd = (loss(A.with_params(opt, tp)) - loss(A.with_params(opt, tm))) / (2 * d)
                r[st._lib.GRAD_PARAMS[i]] = abs(grad[i] - fd) / abs(fd)
            rel[step] = r
        best = {k: min(r[k] for r in rel.values()) for k in rel[1e-3]}
        # the accuracy check: the differentiation workload's own ensemble
        # (131,072 neutrons, 50 s) and loss with fixed steps (h = hmin), where
        # the forward map is smooth in theta, differenced at the stated 1e-3
        # of each parameter; and a smaller case with a step ladder
        # (plus 1e-2 and 1e-4: B0.x's difference carries the loss's curvature
        # at 1e-3, and the weakest couplings change the loss by a few ulp
        # there, so each parameter is also reported at its best step)
        c5 = adjoint_fd_fixed_step(variant, n=args.adjoint_particles, tf=50.0, h=args.adjoint_fd_h,
                                   steps=(1e-3, 1e-2, 1e-4))
        rel3 = c5["rel_err_step_1e-3"]
        out["fd_check"] = {"case": c5["workload"],
                           "params_within_1e-6_step_1e-3": sum(v < 1e-6 for v in rel3.values()),
                           "params_within_1e-6": c5["params_within_1e-6_best_step"],
                           "params_within_1e-6_or_loss_resolution": c5["params_within_1e-6_or_resolution"],
                           "params": c5["params"], "rel_err_step_1e-3": rel3,
                           "rel_err_best_step": c5["rel_err_best_step"],
                           "loss_resolution_step_1e-3": c5["loss_resolution_step_1e-3"]}
        out["fd_fixed_step"] = adjoint_fd_fixed_step(variant)
        # the adaptive 50 s run: its loss carries a ~1e-12 noise floor from
        # accept/reject decisions that move with theta, so weakly coupled
        # parameters resolve only to that floor (informational)
        ad = {"note": "adaptive steps: differences limited by the controller's accept/reject noise",
              "rel_err_step_1e-3": rel[1e-3],
              "loss_change_step_1e-3": {st._lib.GRAD_PARAMS[i]: 2e-3 * abs(theta[i] * grad[i])
                                        for i in ADJ_PARAMS},
              "params_within_1e-6": sum(v < 1e-6 for v in best.values())}
        if args.adjoint_fd_ladder:
            ad["rel_err_step_1e-5"] = rel[1e-5]
            ad["rel_err_best"] = best
        out["fd_adaptive"] = ad
    return out


def adjoint_fd_fixed_step(variant, n=4096, tf=10.0, h=1e-4, steps=(1e-3, 1e-4, 1e-5)):
    """The same physics and loss with fixed integration steps (hmin = h, so
    the steps stay at h), which make the forward map smooth in theta: central
    differences at 1e-3, 1e-4 and 1e-5 of each parameter against the adjoint
    gradient.  Large steps see the loss's non-linearity (the phase fit over
    an aliased precession), small ones its rounding; the best of the three
    and the 1e-3 figure are reported, with the loss's own resolution (4 ulp
    of the loss over the 1e-3 difference's change) per parameter."""
    import numpy as np
    import spintrack as st
    from spintrack import adjoint as A

    opt = st.configs.adjoint_case(n=n, tf=tf)
    st.set_fields(opt, fixedStepSize=True, h=h, hmin=h)
    r = A.ForwardRun(opt, variant=variant)
    try:
        m = r.run()
        Lval, seeds = A.phase_loss(m, r.times, opt.t0)
        grad = r.gradient(seeds)
    finally:
        r.close()

    def loss(th):
        rr = A.ForwardRun(A.with_params(opt, th), variant=variant)
        try:
            return A.phase_loss(rr.run(), rr.times, opt.t0)[0]
        finally:
            rr.close()

    theta = A.params_of(opt)
    rel3, best, res = {}, {}, {}
    for i in ADJ_PARAMS:
        name = st._lib.GRAD_PARAMS[i]
        errs = []
        for step in steps:
            d = step * abs(theta[i])
            tp, tm = theta.copy(), theta.copy()
            tp[i] += d
            tm[i] -= d
            fd = (loss(tp) - loss(tm)) / (2 * d)
            errs.append(abs(grad[i] - fd) / abs(fd))
            if step == 1e-3:
                res[name] = 4 * float(np.spacing(abs(Lval))) / (2 * d * abs(fd))
        rel3[name] = errs[0]
        best[name] = min(errs)
    return {"workload": f"{n} neutrons, {tf:g} s, fixed steps of {h:g} s, same loss, differences at "
                        f"{', '.join(f'{x:g}' for x in steps)} of each parameter", "loss": Lval,
            "rel_err_step_1e-3": rel3, "rel_err_best_step": best, "loss_resolution_step_1e-3": res,
            "params_within_1e-6_best_step": sum(v < 1e-6 for v in best.values()),
            "params_within_1e-6_or_resolution": sum(best[k] < max(1e-6, res[k]) for k in best),
            "params": len(best)}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(opt, n_total, sample_particles, sample_intervals, threads=None):
    """The CPU oracle (the reference CPU path restated in C, OpenMP over
    particles) on a deterministic subset: every k-th particle of the same
    ensemble (unbiased: each particle depends only on seed + index) for the
    first sample_intervals output intervals, on `threads` host threads (all
    the process may use by default)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    import spintrack as st

    stride = max(1, n_total // sample_particles)
    m = n_total // stride
    o = np.frombuffer(bytes(opt), dtype=O.OPTIONS_DTYPE)[0].copy()
    o["numParticles"] = m
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = threads or avail
    E = O.Ensemble(o, m, first=0, stride=stride)
    targets = st.output_times(opt)[:sample_intervals]
    t0 = time.perf_counter()
    prev = opt.t0
    for t in targets:
        E.advance_to(t, prev_t=prev, nthreads=cores)
        prev = t
    dt = time.perf_counter() - t0
    psec = m * (targets[-1] - opt.t0)
    return dict(value=psec / dt, unit="particle-seconds/s", cores=cores, kind="port",
                sample=f"every {stride}th particle ({m} particles) for {len(targets)} output "
                       f"intervals ({targets[-1] - opt.t0:g} s); {dt:.2f} s wall on {cores} threads",
                seconds=dt)


def cpu_baselines(opt, n_total, steps, sample_all, sample_one):
    """The headline's CPU baseline: the oracle over the same simulated span
    as the GPU run (every output interval, so the per-particle work has the
    run's mix of early and late intervals), on all host threads and on one
    thread, with the host's CPU model."""
    allc = cpu_baseline(opt, n_total, sample_all, steps)
    one = cpu_baseline(opt, n_total, sample_one, steps, threads=1)
    allc.update(cpu_model=cpu_model(), threads_available=allc["cores"],
                single_thread={"value": one["value"], "sample": one["sample"], "cores": 1},
                scaling_all_over_one=allc["value"] / one["value"])
    return allc


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="output intervals timed (20 = full 10 s run)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=1048576)
    ap.add_argument("--variant", choices=["fast", "exact"], default="fast")
    ap.add_argument("--integrator", type=int, default=0, choices=[0, 1])
    ap.add_argument("--no-plain", action="store_true", help="skip the per-interval (plain) run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-sample", type=int, default=8192, help="particles in the all-threads CPU sample")
    ap.add_argument("--cpu-sample-1t", type=int, default=512, help="particles in the one-thread CPU sample")
    ap.add_argument("--engine", choices=["pipeline", "fused"], default="pipeline")
    ap.add_argument("--seg-buffer-mb", type=int, default=0, help="pipeline staging budget (0 = library default)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched field-sweep leg")
    ap.add_argument("--no-autotune", action="store_true", help="library default launch shape for the main leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiation leg")
    ap.add_argument("--no-rk45", action="store_true", help="skip the RK45 leg")
    ap.add_argument("--rk45-steps", type=int, default=1, help="output intervals of the RK45 leg")
    ap.add_argument("--rk45-particles", type=int, default=1048576,
                    help="He-3 atoms of the RK45 leg (the first of the seeded ensemble)")
    ap.add_argument("--no-integrators", action="store_true", help="skip the integratorType 2-5 leg")
    ap.add_argument("--no-reference-runs", action="store_true",
                    help="skip repeating the reference's published runs through the C++ driver")
    ap.add_argument("--integrator-particles", type=int, default=1048576)
    ap.add_argument("--no-spectral", action="store_true", help="skip the integratorType 6 leg")
    ap.add_argument("--spectral-particles", type=int, default=1048576)
    ap.add_argument("--no-dressing", action="store_true", help="skip the spin-dressing leg")
    ap.add_argument("--dressing-per", type=int, default=262144, help="particles per species")
    ap.add_argument("--no-adjoint-fd", action="store_true", help="skip its finite-difference check")
    ap.add_argument("--adjoint-particles", type=int, default=131072)
    ap.add_argument("--adjoint-fd-h", type=float, default=5e-4,
                    help="fixed step of the finite-difference check on the differentiation ensemble")
    ap.add_argument("--adjoint-fd-ladder", action="store_true",
                    help="also difference at 1e-5 of each parameter")
    ap.add_argument("--sweep-configs", type=int, default=64)
    ap.add_argument("--sweep-per", type=int, default=16384, help="neutrons per sweep configuration")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="issue one device-to-device copy of this many MiB before the timed run "
                         "(reference dispatch for hardware byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by particle index, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import spintrack as st
    opt = build_workload(args.particles, args.integrator)
    all_targets = st.output_times(opt)
    steps = max(1, min(args.steps, len(all_targets)))
    targets = all_targets[:steps]
    variant = st.VARIANT_FAST if args.variant == "fast" else st.VARIANT_EXACT
    engine = st._lib.ENGINE_PIPELINE if args.engine == "pipeline" else st._lib.ENGINE_FUSED
    print(f"[bench] {args.particles} He-3, {steps} intervals of {opt.ioutInt} s, variant={args.variant}, "
          f"engine={args.engine}", file=sys.stderr, flush=True)

    res = run_gpu(opt, args, targets, True, variant, engine)
    psec = args.particles * (targets[-1] - opt.t0)
    value = psec / (res["kernel_ms"] / 1e3)
    w = res["work"]
    flops = algorithmic_flops(opt, w)
    kt = res["ktimes"]
    if kt:
        # dominant kernel: the spin kernel; its algorithmic work is the
        # integrator trial steps (the per-event work runs in k_kin)
        fa = FLOPS_ACC_GRAV if opt.gravity else FLOPS_ACC_NOGRAV
        spin_flops = (FLOPS_RK45_TRIAL * (w["accepted"] + w["rejected"]) if opt.integratorType == 1
                      else fa * w["accepted"] + FLOPS_REJ * w["rejected"])
        per_launch_ms = kt["spin_ms"] / max(1, kt["spin_launches"])
        traffic = SPIN_HBM_BYTES_PER_RECORD * w["kin_records"] / max(1, kt["spin_launches"])
        achieved_tf = spin_flops / (kt["spin_ms"] / 1e3) / 1e12
        kernel = f"k_spin<{args.variant}, {'RK45' if opt.integratorType == 1 else 'DOP853'}>"
        roof_extra = {"launches": kt["spin_launches"], "avg_launch_ms": per_launch_ms,
                      "kernel_ms": kt["spin_ms"], "kin_kernel_ms": kt["kin_ms"],
                      "kin_launches": kt["kin_launches"], "whole_run_TFLOPs": flops / (res["kernel_ms"] / 1e3) / 1e12,
                      "traffic_unit": "HBM bytes per k_spin launch (measured bytes per record x records per launch)",
                      "traffic_GBs": traffic / (per_launch_ms / 1e3) / 1e9}
    else:
        achieved_tf = flops / (res["kernel_ms"] / 1e3) / 1e12
        kernel = f"k_engine<{args.variant}, persistent>"
        roof_extra = {"launches": 1, "avg_launch_ms": res["kernel_ms"]}
        traffic = None
    print(f"[bench] persistent: {res['kernel_ms']:.1f} ms, {value:.4g} particle-s/s, kernels={kt}, "
          f"work={w}", file=sys.stderr, flush=True)

    plain = None
    if not args.no_plain:
        r2 = run_gpu(opt, args, targets, False, variant, engine)
        plain = psec / (r2["kernel_ms"] / 1e3)
        print(f"[bench] plain: {r2['kernel_ms']:.1f} ms, {plain:.4g} particle-s/s", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baselines(opt, args.particles, steps, args.cpu_sample, args.cpu_sample_1t)
        print(f"[bench] cpu: {cpu}", file=sys.stderr, flush=True)

    sweep = None
    if not args.no_sweep:
        sweep = bench_sweep(args, variant)
        print(f"[bench] sweep: {sweep}", file=sys.stderr, flush=True)

    rk45 = None
    if not args.no_rk45:
        rk45 = bench_rk45(args, variant)
        print(f"[bench] rk45: {rk45}", file=sys.stderr, flush=True)

    refruns = None
    if not args.no_reference_runs:
        refruns = bench_reference_runs(args)

    snap = bench_snapshot(args)
    print(f"[bench] snapshot stream: {snap}", file=sys.stderr, flush=True)

    integ = None
    if not args.no_integrators:
        integ = bench_integrators(args, variant)

    spectral = None
    if not args.no_spectral:
        spectral = bench_spectral(args, variant)

    dressing = None
    if not args.no_dressing:
        dressing = bench_dressing(args, variant)
        print(f"[bench] dressing: {dressing}", file=sys.stderr, flush=True)

    adj = None
    if not args.no_adjoint:
        adj = bench_adjoint(args, variant)
        print(f"[bench] adjoint: {adj}", file=sys.stderr, flush=True)

    line = {
        "metric": "particle-seconds/s, 1 MI355X FP64 rtol=1e-12: batched vs plain; adjoint/fwd cost",
        "value": value,
        "batched_over_plain": sweep["batched_over_plain"] if sweep else None,
        "adjoint_over_forward": adj["adjoint_over_forward"] if adj else None,
        "adjoint_fd_check": adj.get("fd_check") if adj else None,
        "unit": "particle-seconds/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": res["kernel_ms"] / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded ensemble, reference RNG)",
        "config": {"workload": "He-3 thermal 0.45 K, gas tc=2e-3 s, 5% diffuse, gravity, "
                               "DOP853 rtol=atol=1e-12, output every 0.5 s",
                   "particles": args.particles, "sim_seconds": targets[-1] - opt.t0,
                   "variant": args.variant, "integrator": "DOP853" if opt.integratorType == 0 else "RK45",
                   "path": f"st_run, {args.engine} engine",
                   "launch_shape": getattr(args, "tuned_shape", None)},
        "plain_value": plain,
        "roofline": dict({"bound": "fp64_valu", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": achieved_tf / FP64_PEAK_TFLOPS, "traffic": traffic,
                          "kernel": kernel,
                          "compute": "FP64 vector ALU (no MFMA); peak = the FP64 vector rate; achieved = the "
                                     "reference formulation's FP64 operations per trial step (algorithmic)",
                          "hardware_counted": HW_COUNTED,
                          "work": w}, **roof_extra),
        "cpu_baseline": cpu,
        "failures": res["failures"],
        "batched_sweep": sweep,
        "rk45": rk45,
        "integrators": integ,
        "spectral": spectral,
        "snapshot_stream": snap,
        "reference_runs": refruns,
        "dressing": dressing,
        "adjoint": adj,
    }
    # the figures a reader looks for first, last on the line (where a
    # truncated tail still shows them)
    line["summary"] = {
        "value": value, "plain_value": plain, "ms_per_step": res["kernel_ms"] / steps,
        "k_spin_frac_fp64_peak": achieved_tf / FP64_PEAK_TFLOPS,
        "k_spin_ms": kt["spin_ms"] if kt else None, "k_kin_ms": kt["kin_ms"] if kt else None,
        "sweep_batched_over_plain": sweep["batched_over_plain"] if sweep else None,
        "sweep_batched_spin_frac": sweep["batched_spin_frac_fp64_peak"] if sweep else None,
        "dressing_batched_over_plain": dressing["batched_over_plain"] if dressing else None,
        "dressing_batched_spin_frac": dressing["batched_spin_frac_fp64_peak"] if dressing else None,
        "adjoint_over_forward": adj["adjoint_over_forward"] if adj else None,
        "adjoint_fd_within_1e-6": (adj["fd_check"]["params_within_1e-6"] if adj and adj.get("fd_check") else None),
        "adjoint_fd_within_1e-6_or_resolution": (adj["fd_check"]["params_within_1e-6_or_loss_resolution"]
                                                 if adj and adj.get("fd_check") else None),
        "adjoint_fd_case": adj["fd_check"]["case"] if adj and adj.get("fd_check") else None,
        "cpu_all_threads": cpu["value"] if cpu else None,
        "cpu_one_thread": cpu["single_thread"]["value"] if cpu else None,
        "gpu_over_cpu_all_threads": value / cpu["value"] if cpu else None,
    }
    # numpy scalars (counts, differences) print as plain JSON numbers
    print(json.dumps(line, default=lambda o: o.item() if hasattr(o, "item") else str(o)))


if __name__ == "__main__":
    main()
