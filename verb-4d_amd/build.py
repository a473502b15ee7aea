"""Builds the gfx950 shared library libverb4d_hip.so in-tree.

Each HIP source is compiled separately: the assembly, convection and radial
files without floating-point contraction (their arithmetic must match the
reference's operation order bit for bit), the band-LU file with hipcc's
default contraction. Packed FP32 instructions are disabled for all device
code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libverb4d_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# (source, extra flags)
SOURCES = [
    ("radial.hip", ["-ffp-contract=off"]),
    ("convection.hip", ["-ffp-contract=off"]),
    ("local.hip", ["-ffp-contract=off"]),
    ("band_lu.hip", []),
    ("gmres.hip", []),
    ("driver.hip", []),
    ("api.hip", []),
]

COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
    "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for %s" % cmd[-1])
    return r


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    for src, extra in SOURCES:
        obj = os.path.join(OUT_DIR, src.replace(".hip", ".o"))
        cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
