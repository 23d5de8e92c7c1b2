"""Builds hipace2_amd/lib/libhpq.so (the HIP kernels, the C++ host layer and the C ABI
of include/hpq.h) for gfx950 with hipcc, in-tree.

The FP64 path is compiled with -ffp-contract=off so that every multiply and add rounds
separately, as the reference CPU build computes them; the FP32 kernels of the mixed-precision
path (the *_f32.hip units) are compiled with FMA contraction. FP64 atomics use the hardware
global_atomic_add_f64 (-munsafe-fp-atomics; all buffers are coarse-grained device memory).
Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libhpq.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["poisson.hip", "mg.hip", "particles.hip", "deposit.hip", "bins.hip", "dst_mfma.hip", "adjoint.hip", "insitu.hip", "salame.hip", "fields.hip", "gather.hip", "particles_f32.hip", "deposit_f32.hip", "hipace.cpp"]
# sources a translation unit includes besides the headers
INCLUDES = {"particles_f32.hip": ["particles.hip"], "deposit_f32.hip": ["deposit.hip"]}
HEADERS = ["internal.hpp", "shape.hpp", "push.hpp", "tangent.hpp", os.path.join("..", "..", "include", "hpq.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HPQ_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-command-line-argument"]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _flags(src: str):
    """Compile flags of one translation unit: FMA contraction for the FP32 units only."""
    if src.endswith("_f32.hip"):
        return [("-ffp-contract=fast" if f == "-ffp-contract=off" else f) for f in COMMON]
    return COMMON


def _newer(target: str, deps) -> bool:
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def build(force: bool = False, verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hdrs = [os.path.join(CSRC, h) for h in HEADERS]
    objs, cmds = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OBJDIR, src + ".o")
        objs.append(obj)
        deps = [path] + hdrs + [os.path.join(CSRC, d) for d in INCLUDES.get(src, [])]
        if force or _newer(obj, deps):
            cmds.append([HIPCC] + _flags(src) + DEVICE + ["-c", path, "-o", obj])
    # translation units compile in parallel (hipcc is single-threaded per file)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "8"))))
    for i in range(0, len(cmds), jobs):
        procs = []
        for cmd in cmds[i:i + jobs]:
            if verbose:
                print(" ".join(cmd), file=sys.stderr)
            procs.append((cmd, subprocess.Popen(cmd)))
        failed = [cmd for cmd, pr in procs if pr.wait() != 0]
        if failed:
            raise subprocess.CalledProcessError(1, failed[0])
    if force or _newer(LIB, objs):
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
        if verbose:
            print(" ".join(cmd), file=sys.stderr)
        subprocess.run(cmd, check=True)
        os.replace(tmp, LIB)
    return LIB


# host-side AddressSanitizer + UndefinedBehaviorSanitizer (device code is built as usual);
# used by tests/test_sanitizers.py with clang's ASan runtime preloaded
SANITIZE_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                 "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]


def build_sanitized(lib: str, objdir: str) -> str:
    """Builds a copy of the library with the host code sanitized into `lib`."""
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        cmds.append([HIPCC] + _flags(src) + DEVICE + SANITIZE_HOST + ["-c", os.path.join(CSRC, src), "-o", obj])
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "8"))))
    for i in range(0, len(cmds), jobs):
        procs = [(cmd, subprocess.Popen(cmd, stderr=subprocess.DEVNULL)) for cmd in cmds[i:i + jobs]]
        failed = [cmd for cmd, pr in procs if pr.wait() != 0]
        if failed:
            raise subprocess.CalledProcessError(1, failed[0])
    subprocess.run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", lib] + objs, check=True)
    return lib


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
    print(LIB)
