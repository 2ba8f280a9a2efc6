"""Builds the native libraries of axicor_amd in-tree.

* ``axicor_amd/lib/libaxicor_day.so`` — the Day-Phase engine and its C-ABI
  (``include/axicor_day.h``), HIP for gfx950 only.
* ``axicor_amd/lib/libaxicor_host.so`` — host C++: the seeded synthetic brain
  generator (``.state``/``.axons`` blobs, axon paths, input bitmasks) and the
  Night-Phase sprouting pass.
* ``axicor_amd/lib/libaxicor_surrogate.so`` — the parameter-search surrogate
  (``include/axicor_surrogate.h``): MFMA BF16 GEMMs, GELU, Adam; HIP for gfx950.

Run ``python -m axicor_amd.build`` (or ``__graft_entry__.build()``). Every
call compiles; the loaders in :mod:`axicor_amd._ffi` only load what is there.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
LIBDIR = os.path.join(ROOT, "lib")

DAY_LIB = os.path.join(LIBDIR, "libaxicor_day.so")
GEN_LIB = os.path.join(LIBDIR, "libaxicor_host.so")
SUR_LIB = os.path.join(LIBDIR, "libaxicor_surrogate.so")

# gfx950 only. Packed-FP32 instructions are switched off for every HIP file.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the Day-Phase engine needs ROCm's hipcc")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build_day_lib() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = DAY_LIB + ".tmp"
    _run([_hipcc(), *HIPCC_FLAGS, "-o", tmp, os.path.join(CSRC, "day_engine.hip")])
    os.replace(tmp, DAY_LIB)
    return DAY_LIB


def build_gen_lib() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = GEN_LIB + ".tmp"
    # -ffp-contract=off: the sprouting scores are f32 sums evaluated as the
    # reference writes them, without fused multiply-adds.
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-Wall",
          "-o", tmp, os.path.join(CSRC, "brain_gen.cpp"), os.path.join(CSRC, "night_host.cpp")])
    os.replace(tmp, GEN_LIB)
    return GEN_LIB


def build_surrogate_lib() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = SUR_LIB + ".tmp"
    _run([_hipcc(), *HIPCC_FLAGS, "-o", tmp, os.path.join(CSRC, "surrogate.hip")])
    os.replace(tmp, SUR_LIB)
    return SUR_LIB


def build_all() -> list[str]:
    return [build_day_lib(), build_gen_lib(), build_surrogate_lib()]


if __name__ == "__main__":
    for p in build_all():
        print("built", os.path.relpath(p, REPO))
