"""Build of the HIP library libdmtz.so for gfx950 (hipcc, in-tree output).

Flags: -ffp-contract=off keeps every floating-point operation separately rounded (the edit
arithmetic must match the CPU restatement bit for bit); packed FP32 instructions are disabled.
The translation units compile in parallel into objects under _build/ (git-ignored), which are
then linked into dmtz_amd/libdmtz.so.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdmtz.so")
OBJDIR = os.path.join(HERE, "_build")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + CSRC,
    "-I" + os.path.join(ROOT, "include"),
]

# one translation unit per field type for the correction path, so they compile in parallel
SOURCES = ["dmtz_core.hip", "dmtz_f32.hip", "dmtz_f64.hip", "dmtz_codec.hip", "dmtz_noalloc.hip", "learned.hip"]


def hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _filter(text):
    return "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)


def build(verbose=True):
    """Compile libdmtz.so (always recompiles every unit)."""
    os.makedirs(OBJDIR, exist_ok=True)
    procs = []
    objs = []
    for s in SOURCES:
        obj = os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o")
        cmd = [hipcc()] + HIPCC_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, s)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        out = _filter(out)
        if p.returncode != 0:
            failed.append(s)
            sys.stderr.write(out + "\n")
        elif out.strip() and verbose:
            sys.stderr.write(out + "\n")
    if failed:
        raise RuntimeError("hipcc failed on %s" % ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(_filter(r.stdout) + "\n")
        raise RuntimeError("hipcc link failed with code %d" % r.returncode)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
