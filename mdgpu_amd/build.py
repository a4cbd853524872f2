"""Build the in-tree shared library mdgpu_amd/libmdgpu_amd.so for gfx950.

Host sources (*.cpp) are compiled with g++, device sources (*.hip) with
hipcc --offload-arch=gfx950; everything is linked with hipcc into one .so
that exports the C ABI of include/mdgpu_c.h and the C++ drop-in of
include/mdgpu.h.  Run:  python -m mdgpu_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libmdgpu_amd.so")
CLI = os.path.join(HERE, "mdgpu")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
              f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
             "-Wno-unused-parameter", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-ffp-contract=fast"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.stderr


SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]


def build_sanitized():
    """Host code (the analysis, plan, drop-in and C ABI: every *.cpp) built
    with AddressSanitizer + UndefinedBehaviorSanitizer into
    libmdgpu_amd_san.so; the device code is compiled as usual.  Run the CPU
    tests against it with MDGPU_AMD_LIB pointing at it and the gcc
    sanitizer runtimes preloaded."""
    obj = os.path.join(HERE, "_build_san")
    os.makedirs(obj, exist_ok=True)
    jobs = []
    for src in sorted(glob.glob(os.path.join(CSRC, "*.cpp"))):
        base = os.path.basename(src)
        flots = ("float", "double") if base == "dropin.cpp" else (None,)
        for t in flots:
            extra = [f"-DFLOTANTE={t}"] if t else []
            o = os.path.join(obj, base + (f".{t}" if t else "") + ".o")
            jobs.append(["g++", *HOST_FLAGS, *SAN_FLAGS, *extra, "-c", src, "-o", o])
    for src in sorted(glob.glob(os.path.join(CSRC, "*.hip"))):
        jobs.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", os.path.join(obj, os.path.basename(src) + ".o")])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    out = os.path.join(HERE, "libmdgpu_amd_san.so")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", out, *[j[-1] for j in jobs],
          "-lasan", "-lubsan"])
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in sorted(glob.glob(os.path.join(CSRC, "*.cpp"))):
        if os.path.basename(src) == "dropin.cpp":
            # the C++ drop-in once per FLOTANTE (its symbols carry it)
            for t in ("float", "double"):
                obj = os.path.join(OBJ, f"dropin.{t}.o")
                jobs.append(["g++", *HOST_FLAGS, f"-DFLOTANTE={t}", "-c", src, "-o", obj])
            continue
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        jobs.append(["g++", *HOST_FLAGS, "-c", src, "-o", obj])
    for src in sorted(glob.glob(os.path.join(CSRC, "*.hip"))):
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for warn in ex.map(_run, jobs):
            if verbose and warn.strip():
                print(warn, file=sys.stderr)
    objs = [j[-1] for j in jobs]
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs])
    shutil.move(tmp, OUT)
    # command-line driver over the drop-in C++ interface
    # (FP64 drop-in, like the reference driver built with FLOTANTE = double)
    _run(["g++", "-O2", "-std=c++17", "-DFLOTANTE=double", os.path.join(HERE, "cli", "mdgpu_main.cpp"), "-o", CLI,
          f"-L{HERE}", "-lmdgpu_amd", "-Wl,-rpath,$ORIGIN"])
    return OUT


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        print(build_sanitized())
    else:
        print(build(verbose="-v" in sys.argv))
