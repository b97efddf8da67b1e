"""Build libhypre_mi355x.so (all HIP sources, gfx950) in-tree.

    python -m hypre_amd.build            # compile + link
    python -m hypre_amd.build --sanitize # host ASan + UBSan build (_asan/, examples/*_asan)
    python -m hypre_amd.build --clean

Objects go to hypre_amd/_build/, the library to hypre_amd/libhypre_mi355x.so.
Compiled with -ffp-contract=off (the setup must reproduce the reference's
floating-point operation order) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhypre_mi355x.so")
ARCH = os.environ.get("HYPRE_MI355X_ARCH", "gfx950")

SOURCES = ["hx_runtime.hip", "hx_solve.hip", "hx_setup.hip", "hx_amg.hip", "hx_ij.hip",
           "hx_mlp.hip", "hx_hmis.hip", "hx_agg.hip", "hx_noalloc.hip", "hx_stats.hip"]

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_examples(verbose)
    return LIB


EXAMPLES = os.path.join(HERE, "examples")


def build_examples(verbose: bool = True) -> None:
    """Plain-C drivers that use only hypre's public API, linked against the library."""
    for src in sorted(os.listdir(EXAMPLES)) if os.path.isdir(EXAMPLES) else []:
        if not src.endswith(".c"):
            continue
        exe = os.path.join(EXAMPLES, src[:-2])
        cmd = ["gcc", "-O2", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"),
               os.path.join(EXAMPLES, src), "-L", HERE, "-lhypre_mi355x",
               "-Wl,-rpath,$ORIGIN/..", "-lm", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"example build failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"built {exe}")


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer build: every source
# with the host half of the compilation instrumented (-Xarch_host before each
# sanitizer flag; device code is built as usual), linked into
# _asan/libhypre_mi355x_asan.so, and the C example drivers built with clang's
# sanitizer runtime against it (examples/*_asan).  For host-code checks on a
# GPU machine:  ASAN_OPTIONS=detect_leaks=0 examples/ij_amg_pcg_asan 20 7
ASAN_DIR = os.path.join(HERE, "_asan")
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=all"]


def build_sanitized(verbose: bool = True) -> str:
    os.makedirs(ASAN_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def comp(src):
        obj = os.path.join(ASAN_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *CFLAGS, *SAN_FLAGS, "-g", "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stderr}")
        return obj

    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(comp, srcs))
    lib = os.path.join(ASAN_DIR, "libhypre_mi355x_asan.so")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *SAN_FLAGS, "-o", lib, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    for obj in objs:
        os.remove(obj)
    clang = os.path.join(os.path.dirname(os.path.realpath(hipcc())), "..", "lib", "llvm", "bin",
                         "clang")
    if not os.path.exists(clang):
        clang = "/opt/rocm/llvm/bin/clang"
    for src in sorted(os.listdir(EXAMPLES)):
        if not src.endswith(".c"):
            continue
        exe = os.path.join(EXAMPLES, src[:-2] + "_asan")
        cmd = [clang, "-O1", "-g", "-std=c99", "-fsanitize=address,undefined",
               "-fno-omit-frame-pointer", "-fno-sanitize-recover=all", "-I",
               os.path.join(ROOT, "include"), os.path.join(EXAMPLES, src), lib,
               "-Wl,-rpath,$ORIGIN/../_asan", "-lm", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"example build failed: {src}\n{' '.join(cmd)}\n{r.stderr}")
        if verbose:
            print(f"built {exe}")
    return lib


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)
    if os.path.exists(LIB):
        os.remove(LIB)


if __name__ == "__main__":
    if "--clean" in sys.argv:
        clean()
    elif "--sanitize" in sys.argv:
        build_sanitized()
    else:
        build()
