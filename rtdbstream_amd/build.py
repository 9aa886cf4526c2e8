"""Builds the HIP engine (librtdbs.so) for gfx950 in-tree with hipcc.

The library is the product: C++ host engine + hand-written HIP kernels behind
the C ABI of include/rtdbs.h. Python only drives the compiler here.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "librtdbs.so")
RUNNER = os.path.join(PKG, "rtdbs_run")  # command-line runner, links librtdbs.so
BUILD_DIR = os.path.join(PKG, "_build")

SOURCES = ["rtdbs_kernels.hip", "rtdbs_engine.cpp", "rtdbs_server.cpp", "rtdbs_tune.cpp", "rtdbs_validate.hip", "rtdbs_dataset.cpp"]

# -ffp-contract=off: distances must be bit-identical to the FP32 reference
# order (no fused multiply-add). Packed FP32 ops are disabled for all HIP code.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
    "-Wall",
    "-Wno-unused-result",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    return out


def build_library(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    cc = hipcc()
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([cc] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        outs = list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    outs.append(_run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]))
    os.replace(tmp, LIB)
    # the runner: host code only, linked against the library next to it
    robj = os.path.join(BUILD_DIR, "rtdbs_run.o")
    outs.append(_run([cc] + HIP_FLAGS + ["-c", os.path.join(CSRC, "rtdbs_run.cpp"), "-o", robj]))
    rtmp = RUNNER + ".tmp"
    outs.append(_run([cc, "-o", rtmp, robj, "-L" + PKG, "-lrtdbs", "-Wl,-rpath,$ORIGIN"]))
    os.replace(rtmp, RUNNER)
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
