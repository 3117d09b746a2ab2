"""Builds the test-only self-test kernels (tests/gpu/_build/libmpqp_selftest.so)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from mpqpcudasolver_amd.build import HIPCC_FLAGS, hipcc  # noqa: E402

OUT = os.path.join(HERE, "_build", "libmpqp_selftest.so")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", OUT, os.path.join(HERE, "selftest.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(r.stderr[-4000:])
    return OUT


if __name__ == "__main__":
    print(build())
