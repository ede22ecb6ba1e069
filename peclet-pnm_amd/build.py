"""Build libpnm_hip.so (gfx950 HIP kernels + C-ABI) and the pybind11 module peclet.pnm._pnm.

Everything is built in-tree: objects under peclet-pnm_amd/_build/, the two shared libraries
into peclet-pnm_amd/peclet/pnm/ (the module finds the library through an $ORIGIN rpath).

    python peclet-pnm_amd/build.py            # incremental
    python peclet-pnm_amd/build.py --force    # rebuild everything
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
PKG = os.path.join(HERE, "peclet", "pnm")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PNM_OFFLOAD_ARCH", "gfx950")

# Contraction off: FP32 centroids and FP64 sums must round as the written expressions do.
# Packed FP32 instructions are disabled for correctness on this hardware.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]

SOURCES = [
    "k_util.hip",
    "k_extract.hip",
    "k_segment.hip",
    "k_topology.hip",
    "k_flow.hip",
    "k_tile.hip",
    "k_stokes.hip",
    "k_surrogate.hip",
    "k_batch.hip",
    "pipeline.hip",
    "capi.cpp",
]
HEADERS = ["pnm_internal.hpp", "k_common.hpp"]

LIB = os.path.join(PKG, "libpnm_hip.so")


def _newer(src_list, target):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(s) > t for s in src_list)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}):\n{' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def build_lib(force=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    hdrs = [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(INCLUDE, "pnm_capi.h")]
    objs, todo = [], []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, s + ".o")
        objs.append(obj)
        if force or _newer([src] + hdrs, obj):
            lang = ["-x", "hip"] if s.endswith(".hip") else ["-x", "hip"]
            todo.append([HIPCC, *HIP_FLAGS, *lang, "-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, todo):
            pass
    if force or todo or _newer(objs, LIB):
        tmp = LIB + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
        os.replace(tmp, LIB)
    return LIB


def build_module(force=False):
    import pybind11  # noqa: PLC0415

    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    target = os.path.join(PKG, "_pnm" + suffix)
    src = os.path.join(HERE, "binding", "pnm_module.cpp")
    deps = [src, os.path.join(INCLUDE, "pnm_capi.h"), LIB]
    if not (force or _newer(deps, target)):
        return target
    tmp = target + ".tmp"
    _run([
        "g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
        f"-I{INCLUDE}", f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
        src, "-o", tmp, f"-L{PKG}", "-lpnm_hip", "-Wl,-rpath,$ORIGIN",
    ])
    os.replace(tmp, target)
    return target


SYNTH = os.path.join(PKG, "libpnm_synth.so")


def build_synth(force=False):
    """Seeded synthetic-geometry generators (host C + OpenMP), used by the bench and tests."""
    src = os.path.join(CSRC, "synth.c")
    if not (force or _newer([src], SYNTH)):
        return SYNTH
    tmp = SYNTH + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared", "-Wall", src, "-o", tmp,
          "-lm"])
    os.replace(tmp, SYNTH)
    return SYNTH


def build(force=False):
    lib = build_lib(force)
    mod = build_module(force)
    syn = build_synth(force)
    return lib, mod, syn


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args()
    for p in build(a.force):
        print(p)
    sys.exit(0)
