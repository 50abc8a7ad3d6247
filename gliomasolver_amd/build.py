"""Build libgliomahip.so in-tree for gfx950 (hipcc), the host executables in
gliomasolver_amd/bin, and the oracle checker.

Usage: python -m gliomasolver_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgliomahip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no FMA contraction anywhere, the reference's float/double
# rounding points are reproduced exactly.  Packed-FP32 ops are disabled for
# every HIP translation unit.
COMMON = ["-O3", "-fPIC", "-pthread", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include")]
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fno-gpu-rdc"]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith(".hip") or name.endswith(".cpp"):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += HIP_FLAGS + ["-x", "hip"]
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build_library(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    return LIB


APPS = os.path.join(PKG, "apps")
BIN = os.path.join(PKG, "bin")


def build_apps() -> list:
    """Host executables over the C ABI (brain, likelihood, rd_pipeline), linked
    against the in-tree library with an $ORIGIN-relative rpath."""
    os.makedirs(BIN, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    out = []
    for name in sorted(os.listdir(APPS)):
        if not name.endswith(".cpp"):
            continue
        exe = os.path.join(BIN, name[:-4])
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + APPS,
               os.path.join(APPS, name), "-o", exe + ".tmp", "-L" + PKG, "-lgliomahip",
               "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("app build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def build_oracle() -> None:
    """Checker only (oracle/): liboracle.so, and oracle/_ref when the reference tree exists."""
    make = shutil.which("make")
    if make is None:
        raise RuntimeError("make not found")
    r = subprocess.run([make, "-s", "-C", os.path.join(ROOT, "oracle"), "all"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s\n%s" % (r.stdout, r.stderr))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args(argv)
    print(build_library(a.jobs))
    for exe in build_apps():
        print(exe)
    if not a.no_oracle:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
