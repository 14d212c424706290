"""The C++ host driver (spin-tracking_amd/host): the params.txt grammar
restated in C++ gives the same 440-byte options block as the Python
restatement (both follow src/optionsParser.cpp:40-197 and
src/simulation.cpp:9-11); on the GPU, the driver's output file equals the
Python mainAnalysis mirror's byte for byte, per interval and persistent."""
import io
import os
import subprocess

import numpy as np
import pytest

import spintrack as st
from conftest import ROOT

EXE = os.path.join(ROOT, "spin-tracking_amd", "spintrack_main")

PARAMS = """numParticles, 192
dist, M
T, 0.45
gas_coll, TRUE
gravity, true
diffuse, 0.05
B0, 3e-6, 0, 0
E, 0, 0, 7.5e6
Gx, 0, 0, 1e-8
Gy, 1e-9, 0, 0
yi, 0, 1, 0
L, 0.07, 0.1, 0.4
tf, 0.2
ioutInt, 0.05
seed, 7
nmax, 1000000
integratorType, 0
fixedStepSize, False
keepStepSize, yes
m, 5.008e-27
gamma, -2.038e8
h, 1e-3
rtol, 1e-11
atol , 1e-11
bogus, 3
justone
"""


def write(tmp_path, text=PARAMS):
    p = tmp_path / "params.txt"
    p.write_text(text)
    return str(p)


def driver(*args, check=True):
    return subprocess.run([EXE, *args], capture_output=True, check=check, timeout=120)


def test_driver_parser_matches_python(tmp_path):
    assert os.path.exists(EXE), "build with make -C spin-tracking_amd"
    path = write(tmp_path)
    r = driver(path, str(tmp_path / "unused.bin"), "--print-options")
    o = st.options.parse_params(PARAMS, log=io.StringIO())
    st.options.derive(o)
    assert r.stdout == bytes(o)
    assert b"Unrecognized Option: bogus" in r.stderr and b"justone" in r.stderr
    # the tc override replaces the derived mean free time
    r2 = driver(path, str(tmp_path / "unused.bin"), "--print-options", "--tc", "2e-3")
    o.tc = 2e-3
    assert r2.stdout == bytes(o)


def test_driver_rejects_malformed_numbers(tmp_path):
    # std::stod throws in the reference; the driver reports and exits non-zero
    path = write(tmp_path, "tf, abc\n")
    r = driver(path, str(tmp_path / "x.bin"), "--print-options", check=False)
    assert r.returncode != 0 and b"params file" in r.stderr
    with pytest.raises(ValueError):
        st.options.parse_params("tf, abc\n", log=io.StringIO())


@pytest.mark.gpu
@pytest.mark.parametrize("persistent", [False, True])
def test_driver_output_matches_python_main_analysis(tmp_path, persistent):
    path = write(tmp_path)
    out_cpp = str(tmp_path / "cpp.bin")
    out_py = str(tmp_path / "py.bin")
    args = [path, out_cpp] + (["--persistent"] if persistent else [])
    print(f"[driver] {' '.join(args)}", flush=True)
    r = driver(*args)
    assert r.stdout.splitlines()[-1].strip().isdigit()   # total milliseconds, as the reference prints
    opt = st.read_params(path, log=io.StringIO())
    if persistent:
        st.simulation.main_analysis_persistent(opt, out_py, variant=st.VARIANT_EXACT)
    else:
        st.simulation.main_analysis(opt, out_py, log=io.StringIO())
    a, b = open(out_cpp, "rb").read(), open(out_py, "rb").read()
    assert len(a) == len(b) == 440 + 5 * 108 * 192
    assert a == b
    hdr, snaps = st.read_output(out_cpp)
    assert hdr.numParticles == 192 and snaps["t"].shape == (5, 192)
    assert np.all(snaps["failure"][-1] == 0)


@pytest.mark.gpu
@pytest.mark.parametrize("persistent", [False, True])
def test_driver_spectral_run_appends_b_end(tmp_path, persistent):
    """integratorType 6 through the driver: the file ends with b_end
    (particle::postProcess, src/particle.cpp:845-850) and equals the Python
    mainAnalysis mirror's byte for byte."""
    text = PARAMS.replace("integratorType, 0", "integratorType, 6").replace("h, 1e-3", "h, 1e-4")
    text += "a, 3e-5\nw, 6283.185307179586\n"
    path = write(tmp_path, text)
    out_cpp = str(tmp_path / "cpp.bin")
    out_py = str(tmp_path / "py.bin")
    args = [path, out_cpp] + (["--persistent"] if persistent else [])
    print(f"[driver] spectral {' '.join(args)}", flush=True)
    driver(*args)
    opt = st.read_params(path, log=io.StringIO())
    assert opt.integratorType == 6 and opt.w > 0
    if persistent:
        st.simulation.main_analysis_persistent(opt, out_py, variant=st.VARIANT_EXACT)
    else:
        st.simulation.main_analysis(opt, out_py, log=io.StringIO())
    a, b = open(out_cpp, "rb").read(), open(out_py, "rb").read()
    assert len(a) == len(b) == 440 + 5 * 108 * 192 + 24
    assert a == b
    _, snaps = st.read_output(out_cpp)
    assert snaps["t"].shape == (5, 192)
    assert np.all(np.isfinite(snaps["b_end"])) and 0 < np.linalg.norm(snaps["b_end"]) <= 1 + 1e-9

