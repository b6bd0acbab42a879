"""The C++ facade (include/skt_environment.hpp), compiled and linked against
libskt.so the way a C++ caller of the reference's environmentGPU would use it.

CPU: the facade and tests/cpp/facade_step.cpp compile and link (every symbol
they use resolves in libskt.so). GPU: the program steps one environment
through the facade and one through the raw C ABI; the sky fields must agree
bit for bit, and the facade must reject foreign device arrays. Its readTable
reads the DWD fixture and initialises an environment on the device, which
must equal skt_env_init on the skt_sounding_init_fields arrays.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "skewtcloudsim_amd")
SRC = os.path.join(ROOT, "tests", "cpp", "facade_step.cpp")


def _build(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    if not os.path.exists(os.path.join(LIBDIR, "libskt.so")):
        pytest.fail("libskt.so is not built (run __graft_entry__.build())")
    exe = os.path.join(str(tmp_path), "facade_step")
    cmd = [gxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC, "-L", LIBDIR,
           "-lskt", "-Wl,-rpath," + LIBDIR, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def test_facade_compiles_and_links(tmp_path):
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
def test_facade_steps_like_the_c_abi(tmp_path):
    exe = _build(tmp_path)
    dwd = os.path.join(ROOT, "tests", "golden", "sekundenwerte_aero_01303_fixture.zip")
    r = subprocess.run([exe, dwd], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "facade ok" in r.stdout, r.stdout + r.stderr
