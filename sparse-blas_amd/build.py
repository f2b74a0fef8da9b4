"""Build libspx.so (gfx950 HIP kernels + C ABI + drop-in C++ API) and
libspxgen.so (host-only synthetic matrix generators) in-tree.

    python sparse-blas_amd/build.py          # both libraries and the test_spmv driver
    python sparse-blas_amd/build.py --lib    # libspx.so (and the driver)

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions are
disabled for all device code (-packed-fp32-ops).
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPX_OFFLOAD_ARCH", "gfx950")

LIB = os.path.join(HERE, "libspx.so")
GENLIB = os.path.join(HERE, "libspxgen.so")

DEVICE_SOURCES = ["kernels.hip", "build.hip", "noalloc.hip", "seg.hip", "spmm.hip", "select.hip", "plan.cpp", "stage.cpp", "dropin.cpp", "select_api.cpp", "mmio.cpp"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Wno-unused-result", "-ffp-contract=fast",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _run(cmd):
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if err.strip():
        sys.stderr.write(err + "\n")
    return time.time() - t


def build_lib():
    objs = []
    for src in DEVICE_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", path, "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


ASAN_DIR = os.path.join(HERE, "_asan")
ASAN_LIB = os.path.join(ASAN_DIR, "libspx.so")
# Host code instrumented with AddressSanitizer and UndefinedBehaviorSanitizer
# (device code is built as usual): the CPU tests run against it with the
# clang ASan runtime preloaded (tests/test_sanitized_host.py).
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]


def asan_runtime():
    """The clang ASan runtime hipcc links the sanitized host code against."""
    base = "/opt/rocm/lib/llvm/lib/clang"
    for v in sorted(os.listdir(base)) if os.path.isdir(base) else []:
        p = os.path.join(base, v, "lib", "linux", "libclang_rt.asan-x86_64.so")
        if os.path.exists(p):
            return p
    return None


def build_asan():
    os.makedirs(ASAN_DIR, exist_ok=True)
    objs = []
    flags = [f for f in HIP_FLAGS if f != "-O3"] + ["-O1", "-g"] + SAN_HOST
    for src in DEVICE_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(ASAN_DIR, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *flags, *lang, "-c", path, "-o", obj])
        objs.append(obj)
    tmp = ASAN_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-shared-libsan", *SAN_HOST, "-o", tmp, *objs])
    os.replace(tmp, ASAN_LIB)
    for o in objs:
        os.remove(o)
    return ASAN_LIB


def build_gen():
    tmp = GENLIB + ".tmp"
    _run(["g++", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-Wall", "-o", tmp,
          os.path.join(CSRC, "matgen.cpp")])
    os.replace(tmp, GENLIB)
    return GENLIB


HARNESS = os.path.join(HERE, "test_spmv")


def build_harness():
    """The reference's test driver workflow (csrc/test_spmv.cpp) linked
    against libspx.so and libspxgen.so, found next to the binary at run time."""
    tmp = HARNESS + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", tmp, os.path.join(CSRC, "test_spmv.cpp"),
          "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-L", HERE, "-lspx", "-lspxgen",
          "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HARNESS)
    return HARNESS


def build_handle_check():
    """The persistent-handle check program (tests/handle_check.cpp): a C++
    caller of include/anonymouslib_hip.h linked against libspx.so."""
    out = os.path.join(ROOT, "tests", "handle_check")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", tmp, os.path.join(ROOT, "tests", "handle_check.cpp"),
          "-I", os.path.join(ROOT, "include"), "-L", HERE, "-lspx",
          "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,-rpath,$ORIGIN/../sparse-blas_amd"])
    os.replace(tmp, out)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true")
    ap.add_argument("--gen", action="store_true")
    ap.add_argument("--asan", action="store_true", help="only the sanitized host build")
    a = ap.parse_args(argv)
    if a.asan:
        print("built", build_asan())
        return
    both = not (a.lib or a.gen)
    if a.gen or both:
        print("built", build_gen())
    if a.lib or both:
        print("built", build_lib())
    if both or (os.path.exists(LIB) and os.path.exists(GENLIB)):
        print("built", build_harness())
        print("built", build_handle_check())
    if both:
        print("built", build_asan())


if __name__ == "__main__":
    main()
