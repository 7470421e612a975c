"""Build of the HIP library libdmtz.so for gfx950 (hipcc, in-tree output).

Flags: -ffp-contract=off keeps every floating-point operation separately rounded (the edit
arithmetic must match the CPU restatement bit for bit); packed FP32 instructions are disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdmtz.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + CSRC,
    "-I" + os.path.join(ROOT, "include"),
]

SOURCES = [os.path.join(CSRC, "dmtz_hip.hip"), os.path.join(CSRC, "learned.hip")]


def hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=True):
    """Compile libdmtz.so (always recompiles)."""
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("hipcc failed with code %d" % r.returncode)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
