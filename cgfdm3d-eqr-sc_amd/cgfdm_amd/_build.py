"""Compile the HIP library for gfx950 in-tree (libcgfdm_amd.so next to this file)."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(os.path.dirname(ROOT), "include")
LIB = os.path.join(PKG, "libcgfdm_amd.so")
OBJ = os.path.join(ROOT, "_build")

SOURCES = ["setup.hip", "step_fp32.hip", "step_fp64.hip", "step_fp16.hip", "step_bf16.hip", "aweno_fp32.hip",
           "aweno_fp64.hip", "adjoint_fp32.hip", "adjoint_fp64.hip", "capi.cpp"]
HEADERS = ["common.h", "kernels.h", "setup.h", "cgfdm_step_impl.h", "launch_impl.h", "aweno.h", "aweno_impl.h", "adjoint.h", "adjoint_impl.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         # packed FP32 arithmetic disabled: scalar FP32 instructions only
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", CSRC, "-I", INCLUDE]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in r.stderr.splitlines() if "unknown" not in l or "error" in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return out, errs


HOST = os.path.join(ROOT, "host")
HOST_SOURCES = ["json_lite.cpp", "params.cpp", "model_build.cpp", "source_build.cpp", "terrain.cpp", "io_files.cpp"]
RUNNER = os.path.join(PKG, "cgfdm_run")


def build_runner(verbose=True):
    """The params.json driver (C++ host on the C ABI), linked against the library
    next to it; the host sources it relies on are inside the library (and are
    also built on their own by the host tests)."""
    srcs = [os.path.join(HOST, "cgfdm_run.cpp")]
    tmp = RUNNER + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", INCLUDE, "-I", HOST, "-o", tmp] + srcs + \
          ["-L", PKG, "-lcgfdm_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed for cgfdm_run:\n" + r.stderr)
    os.replace(tmp, RUNNER)
    if verbose:
        print("built", RUNNER, file=sys.stderr)
    return RUNNER


# host sources linked into the library: the params.json session entry points
# (cgfdm_create_from_params, cgfdm_write_outputs) and the readers, builders and
# writers they use
LIB_HOST_SOURCES = HOST_SOURCES + ["session.cpp"]


def _compile_host(src):
    out = os.path.join(OBJ, "host_" + os.path.splitext(src)[0] + ".o")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", "-I", INCLUDE, "-I", HOST, "-c", os.path.join(HOST, src),
           "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{r.stderr}")
    return out, r.stderr


def build(verbose=True, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
        host_results = list(ex.map(_compile_host, LIB_HOST_SOURCES))
    objs = [o for o, _ in results] + [o for o, _ in host_results]
    tmp = LIB + ".tmp"
    # -z defs: an unresolved symbol fails the link instead of the first load
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-z,defs", "-o", tmp] + objs
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB, file=sys.stderr)
    build_runner(verbose)
    return LIB


if __name__ == "__main__":
    build()
