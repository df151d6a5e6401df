"""Build libdompap_hip.so (gfx950) in-tree.

Steps:
  1. embed dp_device.h and dp_pair.hip as C string literals
     (dp_embedded.inc) for run-time compilation with hiprtc;
  2. hipcc the host library and the pipeline kernels into
     dompap_amd/libdompap_hip.so.

Usage: python dompap_amd/csrc/build.py [--quiet]
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "libdompap_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed-FP32 instructions are disabled for all HIP code of this project.
NO_PACKED = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _embed() -> str:
    def lit(name, path):
        text = open(path).read()
        delim = "DPRAW"
        if f"){delim}\"" in text:
            raise RuntimeError(f"raw-string delimiter collides in {path}")
        # split into chunks below compiler string-literal limits
        chunks = [text[i:i + 8000] for i in range(0, len(text), 8000)] or [""]
        body = "\n".join(f'R"{delim}({c}){delim}"' for c in chunks)
        return f"static const char {name}[] =\n{body};\n"

    inc = os.path.join(HERE, "dp_embedded.inc")
    content = ("// generated by build.py -- do not edit\n"
               + lit("dp_src_device_h", os.path.join(HERE, "dp_device.h"))
               + lit("dp_src_pair", os.path.join(HERE, "dp_pair.hip")))
    old = open(inc).read() if os.path.exists(inc) else None
    if old != content:
        with open(inc, "w") as f:
            f.write(content)
    return inc


def build(quiet: bool = False) -> str:
    _embed()
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", *NO_PACKED,
           "-fvisibility=hidden",
           os.path.join(HERE, "dp_lib.hip"),
           "-o", OUT + ".tmp", "-lhiprtc"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libdompap_hip.so")
    os.replace(OUT + ".tmp", OUT)
    if not quiet:
        warn = "\n".join(line for line in res.stderr.splitlines() if "packed-fp32-ops" not in line)
        if warn.strip():
            sys.stderr.write(warn + "\n")
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(quiet="--quiet" in sys.argv)
