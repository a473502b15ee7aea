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
    ("driver.hip", ["-ffp-contract=off"]),
    ("remap.hip", ["-ffp-contract=off"]),
    ("da.hip", []),
    ("da_conv.hip", []),
    ("api.hip", []),
    ("noalloc.hip", []),
    ("unet.hip", []),
    ("formats.cpp", []),  # host only: compiled with g++
]

HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]

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


def _stale(obj, src):
    """True when obj is older than its source or any shared header."""
    if not os.path.exists(obj):
        return True
    deps = [src, os.path.join(ROOT, "include", "verb4d_hip.h")]
    deps += [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose=False, incremental=False, jobs=8):
    """Compiles every source (in parallel), then links. incremental=True
    recompiles only objects older than their source or a header."""
    os.makedirs(OUT_DIR, exist_ok=True)
    objs, procs = [], []
    for src, extra in SOURCES:
        obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        path = os.path.join(CSRC, src)
        if incremental and not _stale(obj, path):
            continue
        if src.endswith(".cpp"):
            cmd = ["g++"] + HOST_FLAGS + extra + ["-c", path, "-o", obj]
        else:
            cmd = [HIPCC] + COMMON + extra + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        while len([p for p in procs if p[0].poll() is None]) >= jobs:
            procs[0][0].wait()
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True), cmd))
    failed = []
    for p, cmd in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(" ".join(cmd) + "\n" + out + err)
            failed.append(cmd[-3])
    if failed:
        raise RuntimeError("hipcc failed for %s" % ", ".join(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, incremental="-i" in sys.argv))
