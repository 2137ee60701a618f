"""Builds libaurora_hip.so in-tree with hipcc for gfx950.

Usage: ``python -m aurora_amd.build`` (or ``__graft_entry__.build()``).
Objects go to ``aurora_amd/_build/``; the shared library is written next to
this file so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libaurora_hip.so")
HOST = os.path.join(HERE, "host")
HOST_BIN = os.path.join(HOST, "aurora")
# a C++ program of the test suite that drives the host classes
HOST_CHECK_SRC = os.path.join(HERE, "..", "tests", "cpp", "host_api_check.cpp")
HOST_CHECK_BIN = os.path.join(HERE, "..", "tests", "cpp", "host_api_check")
ARCH = os.environ.get("AURORA_OFFLOAD_ARCH", "gfx950")

SOURCES = ["capi.cpp", "host_scene.cpp", "exr_io.cpp", "halton.cpp", "bvh_build.hip", "wavefront.hip", "parity.hip", "subdiv.hip", "adjoint.hip", "batch.hip"]
EXTRA_SOURCES_FILE = os.path.join(CSRC, "sources.txt")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-Wno-unused-but-set-variable"]
# Packed FP32 instructions stay off for every device compile.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _sources():
    srcs = list(SOURCES)
    if os.path.exists(EXTRA_SOURCES_FILE):
        with open(EXTRA_SOURCES_FILE) as fh:
            srcs += [ln.strip() for ln in fh if ln.strip() and not ln.startswith("#")]
    return srcs


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    cmd = [hipcc()] + COMMON + ["-I", CSRC, "-I", os.path.join(HERE, "..", "include")]
    if src.endswith(".hip"):
        cmd += DEVICE + ["-x", "hip"]
    else:
        cmd += ["-D__HIP_PLATFORM_AMD__"]
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> str:
    """Compiles every source and links the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lamdhip64", "-lz"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stdout))
    os.replace(tmp, LIB)
    build_host()
    if verbose:
        print("built", LIB, "and", HOST_BIN)
    return LIB


def _link_host(src: str, out: str) -> str:
    cxx = os.environ.get("CXX") or shutil.which("g++") or "g++"
    tmp = out + ".tmp"
    rpath = os.path.relpath(HERE, os.path.dirname(out))
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(HERE, "..", "include"),
           src, "-o", tmp, "-L", HERE, "-laurora_hip", "-Wl,-rpath,$ORIGIN/" + rpath]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0 or r.stdout.strip():
        raise RuntimeError("host build %s: %s\n%s" % ("failed" if r.returncode else "warned", " ".join(cmd), r.stdout))
    os.replace(tmp, out)
    return out


def build_host() -> str:
    """Links the C++ command line (host/aurora_main.cpp) and the test suite's
    C++ program against the library with the host compiler; the binaries find
    the library through their rpath."""
    _link_host(os.path.join(HOST, "aurora_main.cpp"), HOST_BIN)
    if os.path.exists(HOST_CHECK_SRC):
        _link_host(HOST_CHECK_SRC, HOST_CHECK_BIN)
    return HOST_BIN


if __name__ == "__main__":
    build()
    sys.exit(0)
