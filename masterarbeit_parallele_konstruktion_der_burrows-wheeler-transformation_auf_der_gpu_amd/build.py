"""Build libmibwt.so (HIP, gfx950) and the CLI binaries in-tree with hipcc.

Usage: python build.py [-j N] [--verbose]
Outputs (git-ignored, shipped with the tree): lib/libmibwt.so, bin/pfp_bwt, bin/lyndon_bwt,
bin/cubwt_compat.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
BIN_DIR = os.path.join(HERE, "bin")
LIB = os.path.join(LIB_DIR, "libmibwt.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for every device compilation.
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I" + CSRC,
          "-I" + os.path.join(HERE, "..", "include")]

LIB_SOURCES = ["abi.hip", "pfp.hip", "lyndon.hip", "selector.hip", "unbwt.hip", "tune.hip"]
CLI_SOURCES = {"pfp_bwt": "cli.cpp", "lyndon_bwt": "cli.cpp", "cubwt_compat": "cli.cpp"}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r


def _compile(src, verbose):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, *DEVICE_FLAGS, *COMMON, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-c", path, "-o", obj]
    _run(cmd, verbose)
    return obj


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(BIN_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), LIB_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", *objs, "-o", tmp, "-ldl"], verbose)
    os.replace(tmp, LIB)
    cli = os.path.join(CSRC, "cli.cpp")
    if os.path.exists(cli):
        for name in CLI_SOURCES:
            out = os.path.join(BIN_DIR, name)
            _run(["g++", "-O2", "-std=c++17", "-DMIBWT_CLI_" + name.upper(),
                  "-I" + os.path.join(HERE, "..", "include"), cli, "-o", out + ".tmp",
                  "-L" + LIB_DIR, "-lmibwt", "-Wl,-rpath,$ORIGIN/../lib"], verbose)
            os.replace(out + ".tmp", out)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.verbose))
