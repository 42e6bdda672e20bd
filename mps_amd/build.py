"""Builds libmps_amd.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

    python -m mps_amd.build            # incremental
    python -m mps_amd.build --force    # rebuild everything
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmps_amd.so")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")

SOURCES = ["grid.hip", "stages.hip", "ppe.hip", "mg.hip", "adjoint.hip", "capi.hip"]
HOST = os.path.join(HERE, "host")
HOST_SOURCES = ["particle.cpp", "logger.cpp", "settings.cpp", "exporter.cpp", "async_export.cpp", "bucket.cpp",
                "mps.cpp", "simulation.cpp", "hooks.cpp"]
HOST_LIB = os.path.join(HERE, "libmps_host.so")
CLI = os.path.join(HERE, "mps")
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-ffp-contract=off",
              f"-I{INCLUDE}"]
HEADERS = [os.path.join(CSRC, "mps_internal.hpp"), os.path.join(INCLUDE, "mps_amd.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPS_AMD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 instructions stay off (hipcc's host pass warns it does not
    # know the feature; that is harmless)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src, force):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if not force and not _stale(obj, [path] + HEADERS):
        return obj, None
    cmd = [HIPCC] + FLAGS + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, f"$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    return obj, None


def build(force: bool = False, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        results = list(ex.map(lambda s: _compile(s, force), SOURCES))
    errs = [e for _, e in results if e]
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs))
    objs = [o for o, _ in results]
    if force or _stale(LIB, objs):
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    _build_host(force)
    if verbose:
        print(LIB)
    return LIB


def _build_host(force):
    """The C++ class surface (MPS, Bucket, Particle, Settings, Exporter,
    Simulation) as libmps_host.so over libmps_amd.so, plus the `mps` CLI."""
    srcs = [os.path.join(HOST, f) for f in HOST_SOURCES]
    hdrs = [os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".hpp")] + HEADERS
    link = ["-L" + HERE, "-lmps_amd", "-Wl,-rpath,$ORIGIN"]
    if force or _stale(HOST_LIB, srcs + hdrs + [LIB]):
        cmd = [CXX] + HOST_FLAGS + ["-shared", "-o", HOST_LIB] + srcs + link
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host build failed:\n$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    main = os.path.join(HOST, "main.cpp")
    if force or _stale(CLI, [main, HOST_LIB] + hdrs):
        cmd = [CXX] + HOST_FLAGS + ["-o", CLI, main, "-L" + HERE, "-lmps_host", "-lmps_amd", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"cli build failed:\n$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args()
    try:
        build(force=a.force, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
