"""Writes the golden fixtures of tests/test_golden.py into tests/golden/golden.npz.

No golden output exists in the reference for this path (its own tests pin
free precession, the Maxwell-Boltzmann initial state and interpolation, all
restated as known-answer tests elsewhere), so the fixtures are:
* RNG: the seeded state words and first outputs of the reference's xoshiro
  variant (src/particle.cpp:59-70, 646-663), from a big-int restatement in
  this script that shares no code with the C oracle;
* trajectories: snapshots of small ensembles from the C oracle, the CPU
  restatement pinned by tests/test_oracle_*.py -- a specular neutron run,
  a He-3 run with gas collisions and diffuse walls, and one short run per
  integratorType 1-6 (type 6 also with its last interval's filter states,
  the summed spectrum and b_end);
* integrator calls: in/out tuples of single integrator calls of types 0-5.
They freeze the checked oracle so that later changes to it, to the device
code or to the toolchain show up as differences, and give the GPU tests
reference outputs that need no CPU run.

Run from the repository root:  python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "spin-tracking_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

M64 = (1 << 64) - 1
RNG_SEEDS = (0, 1, 2, 3, 4, 4294967295)
RNG_INDICES = (0, 1, (1 << 20) - 1, 12345678)
RNG_OUTPUTS = 64


def bigint_stream(seed, index, n):
    """splitmix64 twice from the 32-bit sum seed + index, 32-bit halves as
    the four state words, three discarded draws, then n outputs x + z."""
    s = (seed + index) & 0xFFFFFFFF
    words = []
    for _ in range(2):
        s = (s + 0x9E3779B97F4A7C15) & M64
        r = s
        r = ((r ^ (r >> 30)) * 0xBF58476D1CE4E5B9) & M64
        r = ((r ^ (r >> 27)) * 0x94D049BB133111EB) & M64
        r ^= r >> 31
        words += [r & 0xFFFFFFFF, r >> 32]
    x, y, z, w = words

    def nxt():
        nonlocal x, y, z, w
        res = (x + z) & M64
        t = (y << 17) & M64
        z ^= x
        w ^= y
        y ^= z
        x ^= w
        z ^= t
        w = ((z << 45) | (z >> 19)) & M64
        return res

    for _ in range(3):
        nxt()
    state = [x, y, z, w]
    return state, [nxt() for _ in range(n)]


def trajectory_cases():
    """name -> (options, output times); small enough for a CPU run in seconds."""
    import spintrack as st
    from spintrack import configs

    cases = {}
    o = configs.neutron_cell(n=64, tf=0.2, ioutInt=0.05, seed=11)
    o.yi.x, o.yi.y, o.yi.z = 0.0, 0.6, 0.8
    cases["neutrons"] = o
    cases["he3"] = configs.he3_thermal(n=64, tf=0.1, ioutInt=0.05, seed=12)
    for integ, (rtol, tf) in {1: (1e-10, 0.01), 2: (1e-12, 0.2), 3: (1e-12, 0.2), 4: (1e-4, 0.004),
                              5: (1e-4, 0.02)}.items():
        o = configs.neutron_cell(n=64, tf=tf, ioutInt=tf / 2, seed=20 + integ)
        o.integratorType = integ
        o.rtol = o.atol = rtol
        o.yi.x, o.yi.y, o.yi.z = 0.0, 0.6, 0.8
        cases[f"type{integ}"] = o
    o = configs.neutron_cell(n=64, tf=0.2, ioutInt=0.1, seed=26)
    o.integratorType = 6
    o.a, o.w, o.h = 3e-5, 2 * math.pi * 1000.0, 1e-4
    o.Gy = st.Coords(2e-9, 0.0, 5e-9)
    cases["type6"] = o
    return {k: (v, st.output_times(v)) for k, v in cases.items()}


def run_oracle(O, opt, targets):
    from conftest import to_oracle_opt

    E = O.Ensemble(to_oracle_opt(opt), opt.numParticles)
    snaps = [E.snapshot()]
    prev = opt.t0
    for t in targets:
        E.advance_to(t, prev_t=prev)
        prev = t
        snaps.append(E.snapshot())
    return snaps, E


def integrator_cases(O):
    from test_oracle_integrators import field_opt

    rng = np.random.default_rng(2024)
    rows = []
    for integ in range(6):
        o = field_opt(O)
        o["rtol"] = o["atol"] = {4: 1e-5, 5: 1e-5}.get(integ, 1e-11)
        o["hmin"] = 1e-12
        for _ in range(4):
            y = rng.normal(size=3)
            y /= np.linalg.norm(y)
            p0, v0 = rng.uniform(-0.03, 0.03, 3), rng.normal(0, 5, 3)
            t0 = rng.uniform(0, 1)
            tf = t0 + 10 ** rng.uniform(-4, -2.5)
            h0 = 10 ** rng.uniform(-6, -4)
            stt, yo, ho, na, nr = O.integrate(o, t0, tf, y, p0, v0, h0, integrator=integ)
            rows.append((integ, t0, tf, *y, *p0, *v0, h0, stt, *yo, ho, na, nr))
    return np.array(rows, dtype=np.float64), bytes(o.tobytes())


def main():
    import oracle as O
    from spintrack.options import to_bytes

    out = {}
    st_words, outs = [], []
    for seed in RNG_SEEDS:
        for index in RNG_INDICES:
            s, r = bigint_stream(seed, index, RNG_OUTPUTS)
            st_words.append(s)
            outs.append(r)
    out["rng_seeds"] = np.array([s for s in RNG_SEEDS for _ in RNG_INDICES], dtype=np.uint32)
    out["rng_indices"] = np.array([i for _ in RNG_SEEDS for i in RNG_INDICES], dtype=np.uint32)
    out["rng_state"] = np.array(st_words, dtype=np.uint64)
    out["rng_outputs"] = np.array(outs, dtype=np.uint64)
    names = []
    for name, (opt, targets) in trajectory_cases().items():
        snaps, E = run_oracle(O, opt, targets)
        names.append(name)
        if E.spectral:
            for k in ("s1", "s2", "im", "n_samples"):
                out[f"{name}_agg_{k}"] = np.ascontiguousarray(E.agg[k])
            out[f"{name}_cspec"] = E.cspec
            out[f"{name}_cspec_n"] = np.array(E.cspec_n, dtype=np.int64)
            out[f"{name}_b_end"] = E.floquet_result()
        out[f"{name}_options"] = np.frombuffer(to_bytes(opt), dtype=np.uint8)
        out[f"{name}_targets"] = np.asarray(targets, dtype=np.float64)
        for f in ("t", "pos", "v", "S", "failure", "n_coll", "n_bounce", "n_steps"):
            out[f"{name}_{f}"] = np.stack([s[f] for s in snaps])
        print(f"{name}: {opt.numParticles} particles, {len(targets)} intervals", flush=True)
    out["cases"] = np.array(names)
    rows, optb = integrator_cases(O)
    out["integrator_calls"] = rows
    out["integrator_options"] = np.frombuffer(optb, dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"))


if __name__ == "__main__":
    main()
