"""Build the gfx950 library, the host layer + pybind11 module, and the CPU checkers.

    python ultrasound-engine_amd/build.py            # everything
    python ultrasound-engine_amd/build.py --hip      # only libusx_hip.so

Outputs (in-tree, git-ignored, shipped with the tree):
    ultrasound-engine_amd/usx/libusx_hip.so          C ABI + HIP kernels (hipcc, gfx950)
    ultrasound-engine_amd/usx/_usx.<abi>.so          usx:: host layer + bindings (g++)
    oracle/liboracle_usx.so, oracle/_ref/...         CPU checkers (oracle/Makefile)

Flags. HIP code: -O3, -ffp-contract=off and correctly rounded FP32 divide and
sqrt, so expressions written in the reference's operation order round as the
reference's x86-64 build does; packed FP32 instructions are disabled. Host
code: g++ -O3 without -march and without fast-math, the reference's own
defaults (no FMA contraction), because the host builds the filter taps, the
mixing rotator and the per-transmit constants that must match it bitwise.
"""

from __future__ import annotations

import argparse
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
OUT = os.path.join(PKG, "usx")
BUILD = os.path.join(PKG, "_build")

HIP_SOURCES = ["context.hip", "demod.hip", "beamform.hip", "mv.hip", "apod.hip", "sim.hip", "display.hip",
               "doppler.hip"]
CORE_SOURCES = ["types.cpp", "geometry.cpp", "gpu.cpp", "beamformer.cpp", "pipeline.cpp", "doppler.cpp",
                "simulator.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-fhip-fp32-correctly-rounded-divide-sqrt",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "csrc")]
CXX_FLAGS = ["-std=gnu++17", "-O3", "-DNDEBUG", "-fPIC", "-fopenmp", "-Wall",
             "-Wno-unused-parameter", "-I" + os.path.join(ROOT, "include"),
             "-I" + os.path.join(PKG, "core", "include")]


def _run(cmd, quiet=False):
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(line for line in (r.stdout + r.stderr).splitlines() if noise not in line)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{err}")
    if err.strip() and not quiet:
        print(err)


def _newer(target, deps):
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(d) <= t for d in deps)


def build_hip(force=True):
    os.makedirs(BUILD, exist_ok=True)
    headers = [os.path.join(PKG, "csrc", h) for h in ("kernels.hpp", "device_math.hpp")]
    headers.append(os.path.join(ROOT, "include", "usx_gpu.h"))
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        s = os.path.join(PKG, "csrc", src)
        o = os.path.join(BUILD, src + ".o")
        objs.append(o)
        if force or not _newer(o, [s] + headers):
            jobs.append([HIPCC, *HIP_FLAGS, "-c", s, "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, max(1, len(jobs)))) as ex:
        list(ex.map(_run, jobs))
    target = os.path.join(OUT, "libusx_hip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", target])
    return target


def build_module(force=True):
    import pybind11

    os.makedirs(BUILD, exist_ok=True)
    objs, jobs = [], []
    incs = CXX_FLAGS + ["-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"]]
    headers = [os.path.join(PKG, "core", "include", "usx", h)
               for h in os.listdir(os.path.join(PKG, "core", "include", "usx"))]
    headers += [os.path.join(PKG, "core", "src", "gpu.hpp"), os.path.join(ROOT, "include", "usx_gpu.h")]
    for src in CORE_SOURCES:
        s = os.path.join(PKG, "core", "src", src)
        o = os.path.join(BUILD, src + ".o")
        objs.append(o)
        if force or not _newer(o, [s] + headers):
            jobs.append(["g++", *CXX_FLAGS, "-c", s, "-o", o])
    b = os.path.join(PKG, "bindings", "pymodule.cpp")
    bo = os.path.join(BUILD, "pymodule.o")
    objs.append(bo)
    if force or not _newer(bo, [b] + headers):
        jobs.append(["g++", *incs, "-fvisibility=hidden", "-c", b, "-o", bo])
    with ThreadPoolExecutor(max_workers=min(8, max(1, len(jobs)))) as ex:
        list(ex.map(_run, jobs))
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    target = os.path.join(OUT, "_usx" + suffix)
    _run(["g++", "-shared", "-fopenmp", *objs, "-L" + OUT, "-lusx_hip",
          "-Wl,-rpath,$ORIGIN", "-o", target])
    return target


def build_oracle():
    _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "port"], quiet=True)
    if os.path.isdir("/root/reference/core/src"):
        _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "ref"], quiet=True)


def build_all(force=True):
    build_hip(force)
    build_module(force)
    build_oracle()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--hip", action="store_true", help="only the HIP library")
    ap.add_argument("--module", action="store_true", help="only the host layer + module")
    ap.add_argument("--oracle", action="store_true", help="only the CPU checkers")
    ap.add_argument("--incremental", action="store_true", help="skip up-to-date objects")
    a = ap.parse_args(argv)
    force = not a.incremental
    if a.hip:
        print(build_hip(force))
    elif a.module:
        print(build_module(force))
    elif a.oracle:
        build_oracle()
    else:
        build_all(force)
        print("built", OUT)


if __name__ == "__main__":
    sys.exit(main())
