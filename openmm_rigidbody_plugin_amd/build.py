"""Build the in-tree gfx950 libraries of the rigid-body step.

    python -m openmm_rigidbody_plugin_amd.build         # librbhip.so (+ host plugin)

Outputs (git-ignored, shipped with the tree):
  openmm_rigidbody_plugin_amd/librbhip.so          HIP kernels + C ABI (include/rbhip.h)
  openmm_rigidbody_plugin_amd/librigidbody_hip.so  C++ plugin layer (RigidBodyIntegrator,
                                                   HipIntegrateRigidBodyStepKernel, ...)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "librbhip.so")
PLUGIN = os.path.join(PKG, "librigidbody_hip.so")

ARCH = os.environ.get("RBHIP_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP library)")


# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", INCLUDE, "-I", CSRC]

UNITS = [
    ("rbhip_kernels.hip", []),
    # FP64 body-state build: unfused operations, in the reference's order
    ("rbhip_setup.hip", ["-ffp-contract=off"]),
    ("rbhip_api.hip", []),
    # device bookkeeping (rocPRIM sort and scan)
    ("rbhip_bookkeeping.hip", []),
    # job server: host code over the C ABI
    ("rbhip_server.cpp", []),
]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature for this target" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()

    def compile_unit(unit):
        src, extra = unit
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        out = _run([cc, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        return obj

    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        objs = list(ex.map(compile_unit, UNITS))
    tmp = LIB + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_plugin(verbose: bool = False) -> str | None:
    """C++ host plugin layer linked against librbhip.so (built if its sources exist)."""
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp")) \
        if os.path.isdir(HOST) else []
    if not srcs:
        return None
    cxx = shutil.which("g++") or "g++"
    tmp = PLUGIN + ".tmp"
    out = _run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", INCLUDE, "-I", HOST,
                *srcs, "-o", tmp, "-L", PKG, "-lrbhip", "-Wl,-rpath,$ORIGIN"])
    if verbose and out.strip():
        print(out)
    os.replace(tmp, PLUGIN)
    return PLUGIN


def build_all(verbose: bool = False):
    lib = build_hip(verbose)
    plugin = build_plugin(verbose)
    return lib, plugin


if __name__ == "__main__":
    print(build_all(verbose="-v" in sys.argv))
