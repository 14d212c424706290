"""The device kinematics and RNG (spin-tracking_amd/csrc/st_device.h), built
for the host by hipcc, against the oracle's C restatement, bit for bit, on
random and edge-case states (tests/host_device_math_main.cpp): seeding,
uniform / exponential / normal draws (normal3 with and without a cached
spare), and one full event -- next collision time (src/particle.cpp:121-257),
move and collision (:263-388), sanity check (:412-444) -- with gravity on
and off, specular and diffuse walls, Maxwell-Boltzmann and isotropic gas
collisions, the interval cap, +-0 / +-inf / NaN / tiny velocities, positions
on the walls and an apex at the ceiling; and the same event in the split
order the kinematics kernel runs it (phase A, the re-emission job with the
deferred gas-time draw, phase B), which must end identical.  Diffuse
re-emission velocities are
compared to 1e-15 relative (sin/cos vs sincos in the last bit); the RNG
stream after them stays exact.  No GPU is used.
"""
import os
import subprocess

import pytest

from conftest import ROOT

HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_device_kinematics_and_rng_match_oracle(oracle, tmp_path):
    exe = str(tmp_path / "host_device_math")
    odir = os.path.join(ROOT, "oracle")
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-ffp-contract=off", "-w",
                    os.path.join(ROOT, "tests", "host_device_math_main.cpp"), "-o", exe,
                    f"-L{odir}", "-loracle", f"-Wl,-rpath,{odir}"], check=True, timeout=300)
    r = subprocess.run([exe, "400000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
    # the diffuse branch was exercised
    n_diffuse = int(r.stdout.split("diffuse events ")[1].split()[0])
    assert n_diffuse > 100000
    # the split order (phase A, re-emission job, phase B) of the kinematics
    # kernel ran on every event, gas jobs included
    assert int(r.stdout.split("gas jobs ")[1].split()[0]) > 100000
