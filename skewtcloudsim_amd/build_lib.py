"""Builds libskt.so (the HIP library) for gfx950 in-tree with hipcc.

Usage: python -m skewtcloudsim_amd.build_lib
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["skt_kernels.hip", "skt_misc.hip", "skt_mg.hip", "skt_api.cpp", "skt_batch.cpp", "skt_serve.cpp", "skt_sounding.cpp", "skt_ingest.cpp"]
LIBS = ["-lz"]  # zip members of the radiosonde archives are deflated
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # two roundings for a*b+c, as the CPU restatement is compiled
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-parameter",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    out = os.path.join(HERE, "libskt.so")
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES] + LIBS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libskt.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
