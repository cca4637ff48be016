"""Builds libsp_hip.so (the gfx950 C-ABI library) in-tree with hipcc.

    python -m stencilparareal_amd.build            # compile if sources changed
    python -m stencilparareal_amd.build --force    # always compile

Each .hip file is compiled to an object under stencilparareal_amd/_build/ and the
objects are linked into stencilparareal_amd/libsp_hip.so.  The reference-order
sweeps (sp_staged.hip, sp_misc.hip) are compiled with -ffp-contract=off.  The C++
drivers stencilparareal_amd/bin/parareal and bin/convection (host/*_main.cpp over the
C++ API in include/sp/) are then built with g++ against that library.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsp_hip.so")
# C++ drivers over include/sp/: executable name -> source
DRIVERS = {"parareal": "parareal_main.cpp", "convection": "convection_main.cpp"}
ARCH = os.environ.get("SP_OFFLOAD_ARCH", "gfx950")

SOURCES = {
    "sp_api.hip": ["-ffp-contract=off"],
    "sp_staged.hip": ["-ffp-contract=off"],
    "sp_misc.hip": ["-ffp-contract=off"],
    "sp_fused.hip": [],
    "sp_batch.hip": [],
    "sp_spectral.hip": [],
}
HEADERS = ["sp_internal.hpp", "sp_handles.hpp", os.path.join("..", "..", "include", "stencilparareal.h")]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def common_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
        # packed FP32 instructions are kept out of every code object
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
    ]


def _stale(target: str, deps: list[str]) -> bool:
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def build(force: bool = False, verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    hdrs = [os.path.join(CSRC, h) for h in HEADERS]
    procs, objs = [], []
    for src, extra in SOURCES.items():
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(o)
        if force or _stale(o, [s] + hdrs):
            cmd = [cc] + common_flags() + extra + ["-c", s, "-o", o]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
            if len(procs) >= jobs:
                _wait(procs)
                procs = []
    _wait(procs)
    if force or _stale(LIB, objs):
        tmp = LIB + ".tmp"
        # hipFFT (the spectral coarse propagator), found through the ROCm tree at run time
        rocm_lib = os.path.join(os.path.dirname(os.path.dirname(cc)), "lib")
        cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + \
              ["-L", rocm_lib, "-lhipfft", f"-Wl,-rpath,{rocm_lib}"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stdout)
        os.replace(tmp, LIB)
    build_driver(force=force, verbose=verbose)
    return LIB


def build_driver(force: bool = False, verbose: bool = False) -> list[str]:
    """g++ build of the driver executables in bin/ (rpath to the library one directory up)."""
    hdr_dir = os.path.join(ROOT, "include")
    hdrs = [os.path.join(hdr_dir, "sp", f) for f in os.listdir(os.path.join(hdr_dir, "sp"))]
    hdrs.append(os.path.join(hdr_dir, "stencilparareal.h"))
    out = []
    for exe, src in DRIVERS.items():
        src = os.path.join(HERE, "host", src)
        target = os.path.join(HERE, "bin", exe)
        out.append(target)
        if not (force or _stale(target, [src, LIB] + hdrs)):
            continue
        os.makedirs(os.path.dirname(target), exist_ok=True)
        cxx = os.environ.get("CXX", "g++")
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", hdr_dir, src, "-L", HERE, "-lsp_hip",
               "-Wl,-rpath,$ORIGIN/..", "-o", target + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{exe} build failed:\n" + r.stdout)
        if r.stdout.strip():
            print(r.stdout, file=sys.stderr)
        os.replace(target + ".tmp", target)
    return out


def _wait(procs):
    errs = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            errs.append(f"{src}:\n{text}")
        elif text.strip():
            # keep compiler warnings visible, minus the known host-pass note on the
            # device-only target feature
            lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
            if lines:
                print(f"[{src}]\n" + "\n".join(lines), file=sys.stderr)
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(force=a.force, verbose=a.verbose))
