"""Builds the test-support HIP libraries next to this file (test
infrastructure; the product is skewtcloudsim_amd/libskt.so)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from skewtcloudsim_amd.build_lib import HIPCC_FLAGS, hipcc  # noqa: E402

LIBS = {"libfacecheck.so": ["face_check.hip"], "libpowcheck.so": ["pow_check.hip"]}


def build():
    for out, srcs in LIBS.items():
        cmd = [hipcc()] + HIPCC_FLAGS + ["-o", os.path.join(HERE, out)] + [os.path.join(HERE, s) for s in srcs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed building {out}")


if __name__ == "__main__":
    build()
