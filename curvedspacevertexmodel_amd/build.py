"""Build the HIP library libsvm.so in-tree (hipcc, gfx950).

    python -m curvedspacevertexmodel_amd.build

Flags: -ffp-contract=off keeps every FP64 expression evaluated exactly as
written (the parity mode's bit-exactness depends on it); packed FP32
instructions are disabled for the whole library.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsvm.so")

SOURCES = ["svm_capi.hip"]
HEADERS = ["svm_fp64.h", "svm_geom.h", "svm_device.h", "svm_kernels.h", "svm_delaunay.h", "svm_exact.h"]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def command(out=OUT, extra=()):
    return [
        hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-fno-fast-math",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-result",
        "-I", os.path.join(HERE, "..", "include"),
        *[os.path.join(CSRC, s) for s in SOURCES],
        "-o", out, *extra,
    ]


def build(verbose=False):
    cmd = command(OUT + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libsvm.so")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
