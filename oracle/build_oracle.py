"""Build the C oracle (oracle/libdp_oracle.so) with gcc.

-ffp-contract=off keeps the FP64 arithmetic un-fused, like the reference's
Numba kernels; -march=x86-64-v3 keeps the library runnable on any AVX2 host.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "libdp_oracle.so")


def build(quiet: bool = False) -> str:
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
           "-shared", "-fPIC", os.path.join(HERE, "dp_oracle.c"), "-o", OUT + ".tmp", "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("gcc failed building the C oracle")
    os.replace(OUT + ".tmp", OUT)
    if not quiet:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
