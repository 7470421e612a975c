"""The device headers under AddressSanitizer and UndefinedBehaviorSanitizer, on the host.

tests/native/hostcheck.cpp compiles the headers the HIP kernels use (stencil.h, lower_star.h,
gradient.h, mismatch.h, repair.h, edit.h) for the CPU. Here it is built with ASan + UBSan, every
error fatal, by g++ and by ROCm's clang (the front end hipcc uses), once with normal inlining and
once with every helper kept out of line (-fno-inline), and test_hostcheck.py runs against each
build in a child process with the sanitizer runtime preloaded. Out-of-bounds accesses, signed
overflow, invalid shifts, misaligned or null accesses and similar undefined behaviour in the
classifier, the ID decoders, the pairing lookups and the repair rules then stop the run.
"""
import glob
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "native", "hostcheck.cpp")
FLAGS = ["-g", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-I" + os.path.join(ROOT, "dmtz_amd", "csrc")]


def _gcc_runtime():
    r = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _clang():
    cxx = "/opt/rocm/llvm/bin/clang++"
    rt = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    return (cxx, rt[0]) if os.path.exists(cxx) and rt else (None, None)


def _run(so, preload, libdir=None):
    env = dict(os.environ, DMTZ_HOSTCHECK_SO=so, LD_PRELOAD=preload, ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    if libdir:
        env["LD_LIBRARY_PATH"] = libdir + os.pathsep + env.get("LD_LIBRARY_PATH", "")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(HERE, "test_hostcheck.py")], env=env, capture_output=True, text=True,
                       timeout=900, cwd=ROOT)
    assert r.returncode == 0, (r.stdout[-4000:], r.stderr[-4000:])
    assert " passed" in r.stdout


@pytest.mark.parametrize("inline", [True, False], ids=["inlined", "out_of_line"])
def test_headers_clean_under_gcc_sanitizers(tmp_path, inline):
    rt = _gcc_runtime()
    if rt is None:
        pytest.skip("g++ has no AddressSanitizer runtime")
    so = str(tmp_path / "libhostcheck_gcc.so")
    opt = ["-O2"] if inline else ["-O1", "-fno-inline"]
    subprocess.check_call(["g++"] + opt + FLAGS + ["-o", so, SRC])
    _run(so, rt)


def test_headers_clean_under_clang_sanitizers(tmp_path):
    cxx, rt = _clang()
    if cxx is None:
        pytest.skip("ROCm's clang or its sanitizer runtime is missing")
    so = str(tmp_path / "libhostcheck_clang.so")
    subprocess.check_call([cxx, "-O1", "-fno-inline", "-shared-libsan"] + FLAGS + ["-o", so, SRC])
    _run(so, rt, os.path.dirname(rt))
