"""Builds libtioga_amd.so (the HIP engine and its C ABI) in-tree for gfx950.

    python -m tioga_amd.build        # or tioga_amd.build.build()

Every translation unit is compiled by hipcc with -ffp-contract=off (the FP64
numerics must not be contracted into FMAs, so results match the reference
bit for bit) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.environ.get("TIOGA_AMD_OUT", os.path.join(HERE, "libtioga_amd.so"))
BUILD = os.environ.get("TIOGA_AMD_BUILD_DIR", os.path.join(HERE, "_build"))

SOURCES = ["tg_engine.hip", "tg_amr.hip", "tg_adt.hip", "tg_dedupe.hip", "tg_obb.hip", "tg_prims.hip", "tg_update.hip", "tg_points.hip", "tg_seq.cpp", "tg_serve.cpp", "tg_packets.cpp", "tg_capi.cpp"]
ARCH = os.environ.get("TIOGA_AMD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
         "-Wno-unused-function", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
FLAGS += os.environ.get("TIOGA_AMD_CFLAGS", "").split()


def hipcc() -> str:
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), f"--offload-arch={ARCH}", *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


ROOT = os.path.dirname(HERE)
CXX_DRIVER_SRC = os.path.join(ROOT, "tests", "cxx", "tioga_class_driver.cpp")
CXX_DRIVER = os.path.join(ROOT, "tests", "cxx", "tioga_class_driver")


def build_cxx_driver(verbose: bool = False) -> str:
    """The C++ caller of TIOGA::tioga that tests/test_cxx_class.py runs: a
    host program (g++) linked against libtioga_amd.so in this tree."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", CXX_DRIVER_SRC, "-o", CXX_DRIVER + ".tmp", "-L" + HERE,
           "-ltioga_amd", "-Wl,-rpath,$ORIGIN/../../tioga_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(CXX_DRIVER + ".tmp", CXX_DRIVER)
    if verbose:
        print("built", CXX_DRIVER)
    return CXX_DRIVER


HO_CB_SRC = os.path.join(ROOT, "tests", "cxx", "ho_callbacks.c")
HO_CB_LIB = os.path.join(ROOT, "tests", "cxx", "libho_callbacks.so")


def build_test_callbacks(verbose: bool = False) -> str:
    """The high-order callbacks of the parity cases (tests/cxx/ho_callbacks.c),
    the same source the reference driver links, as a library the tests hand
    to tioga_set_highorder_callback_."""
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", HO_CB_SRC, "-o", HO_CB_LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(HO_CB_LIB + ".tmp", HO_CB_LIB)
    if verbose:
        print("built", HO_CB_LIB)
    return HO_CB_LIB


if __name__ == "__main__":
    build(verbose=True)
    build_cxx_driver(verbose=True)
    build_test_callbacks(verbose=True)
    sys.exit(0)
