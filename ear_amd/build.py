"""Builds the native libraries of the package in-tree.

* ``ear_amd/libear_hip.so``  -- HIP kernels + the C ABI of include/ear.h (gfx950)
* ``ear_amd/bin/EAR``        -- host C++ CLI (``EAR render`` / ``EAR calc T60``)
  and ``ear_amd/libear_host.so`` (Scene / Mesh / Material / Recorder over the ABI)

Every float kernel is compiled with -ffp-contract=off and without fast-math so
that each operation rounds exactly once, in source order, on host and device.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(CSRC, "host")
BUILD = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EAR_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["ear_api.hip", "lbvh.hip", "tracer.hip", "convolve.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    # keep hipcc from emitting packed fp32 math (the gfx950 packed forms are avoided)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT}/include",
]

LIB = os.path.join(PKG, "libear_hip.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature: harmless
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def build_hip(force: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    headers = [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    headers.append(os.path.join(ROOT, "include", "ear.h"))
    objs = []
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        for src in HIP_SOURCES:
            s = os.path.join(CSRC, src)
            o = os.path.join(BUILD, src.replace(".hip", ".o"))
            objs.append(o)
            if force or _newer(o, [s] + headers):
                jobs.append(ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", s, "-o", o]))
        for j in jobs:
            j.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-L/opt/rocm/lib", "-lhipfft",
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


def build_host(force: bool = True) -> str | None:
    """Host C++ layer (reference API names) linked against libear_hip.so."""
    if not os.path.isdir(HOST):
        return None
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    if not srcs:
        return None
    os.makedirs(os.path.join(PKG, "bin"), exist_ok=True)
    lib_srcs = [s for s in srcs if not s.endswith("main.cpp")]
    cxx = shutil.which("g++") or "g++"
    flags = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", f"-I{ROOT}/include", f"-I{HOST}", "-Wall"]
    host_lib = os.path.join(PKG, "libear_host.so")
    _run([cxx, *flags, "-shared", "-o", host_lib, *lib_srcs, f"-L{PKG}", "-lear_hip", "-Wl,-rpath,$ORIGIN"])
    main = os.path.join(HOST, "main.cpp")
    if os.path.exists(main):
        exe = os.path.join(PKG, "bin", "EAR")
        _run([cxx, *flags, "-o", exe, main, f"-L{PKG}", "-lear_host", "-lear_hip",
              "-Wl,-rpath,$ORIGIN/.."])
    return host_lib


def build_all(force: bool = True) -> None:
    build_hip(force)
    build_host(force)


if __name__ == "__main__":
    build_all(force="--incremental" not in sys.argv)
    print("built", LIB)
