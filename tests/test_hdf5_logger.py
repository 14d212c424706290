"""The host driver's HDF5Logger (src/logger.cpp:40-188 of the reference) on
the CPU: a test program built from the driver's sources writes the same
synthetic snapshots through createLogger to a binary and to an HDF5 file;
the HDF5 file holds every option as a root attribute with the reference's
names and types, and the eight [nsave, N] datasets with their row counters,
with values identical to the binary file's.  On the GPU, the driver's own
HDF5 output of a run equals its binary output of the same run."""
import os
import subprocess

import numpy as np
import pytest

import spintrack as st
from spintrack import h5, logger
from conftest import ROOT
from test_host_driver import PARAMS, driver, write

HOST = os.path.join(ROOT, "spin-tracking_amd", "host")
PKG = os.path.join(ROOT, "spin-tracking_amd")
HDF5_INC = "/opt/conda/include"

needs_hdf5 = pytest.mark.skipif(not (os.path.exists(os.path.join(HDF5_INC, "hdf5.h")) and h5.hdf5_available()),
                                reason="no HDF5 C library on this host")


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("h5") / "host_logger_main")
    srcs = [os.path.join(ROOT, "tests", "host_logger_main.cpp")] + [
        os.path.join(HOST, f) for f in ("params.cpp", "run.cpp", "hdf5_logger.cpp")]
    subprocess.run(["g++", "-O1", "-std=c++17", "-DST_HDF5", f'-DST_HDF5_LIB="{h5.DEFAULT_LIB}"',
                    "-isystem", HDF5_INC, f"-I{HOST}", *srcs, f"-L{PKG}", "-lspintrack", "-ldl",
                    f"-Wl,-rpath,{PKG}", "-o", exe], check=True, timeout=300)
    return exe


def fields_equal(a, b):
    for k in a:
        assert np.asarray(a[k]).tobytes() == np.asarray(b[k]).tobytes(), k


@needs_hdf5
@pytest.mark.parametrize("block", [False, True])
def test_hdf5_matches_binary(tmp_path, writer, block):
    params = write(tmp_path)
    K = 4                                        # nsave = (0.2 - 0)/0.05 + 1 = 5 rows
    extra = ["block"] if block else []
    subprocess.run([writer, params, str(tmp_path / "o.bin"), str(K), *extra], check=True, timeout=60)
    subprocess.run([writer, params, str(tmp_path / "o.H5"), str(K), *extra], check=True, timeout=60)
    ob, b = logger.read_output(str(tmp_path / "o.bin"))
    oh, d = h5.read_hdf5(str(tmp_path / "o.H5"))
    assert bytes(oh) == bytes(ob)                # every option, with the reference's names and types
    assert set(d) == set(b)
    for k in b:
        assert d[k].shape == (K, ob.numParticles), k
    fields_equal(d, b)


@needs_hdf5
def test_hdf5_rows_and_extension_rules(tmp_path, writer):
    params = write(tmp_path)
    # more snapshots than the nsave rows the reference sizes the datasets for
    r = subprocess.run([writer, params, str(tmp_path / "o.hdf5"), "6"], capture_output=True, timeout=60)
    assert r.returncode == 5 and b"more snapshots" in r.stderr
    _, d = h5.read_hdf5(str(tmp_path / "o.hdf5"))
    assert d["t"].shape[0] == 5
    # any other extension is the binary format
    subprocess.run([writer, params, str(tmp_path / "o.dat"), "2"], check=True, timeout=60)
    assert os.path.getsize(tmp_path / "o.dat") == 440 + 2 * 108 * 192


@needs_hdf5
@pytest.mark.gpu
@pytest.mark.parametrize("persistent", [False, True])
def test_driver_hdf5_equals_binary(tmp_path, persistent):
    params = write(tmp_path)
    extra = ["--persistent"] if persistent else []
    driver(params, str(tmp_path / "run.bin"), *extra)
    driver(params, str(tmp_path / "run.h5"), *extra)
    ob, b = logger.read_output(str(tmp_path / "run.bin"))
    oh, d = h5.read_hdf5(str(tmp_path / "run.h5"))
    assert bytes(oh) == bytes(ob)
    assert d["t"].shape[0] == b["t"].shape[0] == 5    # initial state + 4 intervals
    fields_equal(d, b)
    assert np.all(np.isfinite(st.xyz(d["S"][-1])))


@needs_hdf5
@pytest.mark.gpu
def test_driver_hdf5_spectral_b_end(tmp_path):
    """integratorType 6: the HDF5 file carries b_end as a root attribute
    (HDF5Logger::writeSingle, src/logger.cpp:68-70), equal to the binary
    file's trailing b_end."""
    text = open(write(tmp_path)).read().replace("integratorType, 0", "integratorType, 6")
    text = text.replace("h, 1e-3", "h, 1e-4") + "a, 3e-5\nw, 6283.185307179586\n"
    (tmp_path / "p6.txt").write_text(text)
    params = str(tmp_path / "p6.txt")
    print("[hdf5] spectral driver run", flush=True)
    driver(params, str(tmp_path / "run6.bin"))
    driver(params, str(tmp_path / "run6.h5"))
    ob, b = logger.read_output(str(tmp_path / "run6.bin"))
    oh, d = h5.read_hdf5(str(tmp_path / "run6.h5"))
    assert bytes(oh) == bytes(ob) and oh.integratorType == 6
    fields_equal(d, b)
    assert np.array_equal(d["b_end"], b["b_end"]) and np.all(np.isfinite(b["b_end"]))

