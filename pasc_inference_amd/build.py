"""Builds libpasc_hip.so (gfx950) in-tree with hipcc.

The fused-pass instantiations are split into one object per (storage type, block size) so that
they compile in parallel; objects go to pasc_inference_amd/_build/ and the shared library to
pasc_inference_amd/libpasc_hip.so.
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
LIB = os.path.join(HERE, "libpasc_hip.so")
ARCH = os.environ.get("PASC_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# packed FP32 instructions are disabled (the compiler must not form v_pk_*_f32)
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-result"]


def _units():
    units = []
    for s in ("float", "double"):
        for nt in (64, 128, 256, 512):
            units.append((os.path.join(CSRC, "pass_inst.hip"), f"pass_{s}_{nt}.o",
                          [f"-DPASS_S={s}", f"-DPASS_NT={nt}"]))
        units.append((os.path.join(CSRC, "grid_inst.hip"), f"grid_{s}.o", [f"-DPASS_S={s}"]))
        units.append((os.path.join(CSRC, "img_inst.hip"), f"img_{s}.o", [f"-DPASS_S={s}"]))
    units.append((os.path.join(CSRC, "engine.hip"), "engine.o", []))
    units.append((os.path.join(CSRC, "adjoint.hip"), "adjoint.o", []))
    units.append((os.path.join(CSRC, "fem.hip"), "fem.o", []))
    units.append((os.path.join(CSRC, "noalloc.hip"), "noalloc.o", []))
    units.append((os.path.join(CSRC, "capi.hip"), "capi.o", []))
    units.append((os.path.join(CSRC, "generate.cpp"), "generate.o", []))
    return units


def _compile(src, obj, extra):
    out = os.path.join(BUILD, obj)
    cmd = [HIPCC] + COMMON + extra + ["-c", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src} {extra}:\n{r.stderr[-6000:]}")
    return out


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    units = _units()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda u: _compile(*u), units))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    build_host(verbose)
    if verbose:
        print(f"built {LIB}")
    return LIB


HOST_DRIVERS = ("test_signal1d", "test_image")
HOST_DRIVER = os.path.join(HERE, "host", "test_signal1d")


def build_host(verbose: bool = False) -> str:
    """The C++ host layer's drivers (host/examples/*.cpp), linked to the library."""
    gxx = shutil.which("g++") or "g++"
    root = os.path.dirname(HERE)
    for name in HOST_DRIVERS:
        cmd = [gxx, "-std=c++17", "-O2", "-Wall", "-Wextra",
               "-I", os.path.join(root, "include"), "-I", os.path.join(HERE, "host", "include"),
               os.path.join(HERE, "host", "examples", name + ".cpp"),
               "-L", HERE, "-lpasc_hip", "-Wl,-rpath,$ORIGIN/..",
               "-o", os.path.join(HERE, "host", name)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host driver {name} build failed:\n{r.stderr[-6000:]}")
        if verbose and r.stderr.strip():
            print(r.stderr)
    return HOST_DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
