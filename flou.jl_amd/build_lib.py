"""Build libflouhip.so (gfx950) from flou.jl_amd/csrc with hipcc.

Usage: python flou.jl_amd/build_lib.py [-j JOBS]

Every translation unit is compiled for gfx950 only.  The FP64 plain kernels and the
reductions are built with -ffp-contract=off so that no a*b+c is fused where the reference
rounds twice; packed FP32 instructions are disabled for all device code.

Also builds libflouvtkhdf.so, the host-side VTKHDF writer, with g++ against the HDF5 C
library found under $FLOU_HDF5_DIR (default /opt/conda); without HDF5 that library is
skipped and only VTKHDF output is unavailable.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libflouhip.so")
VTKHDF_LIB = os.path.join(HERE, "libflouvtkhdf.so")
HDF5_DIR = os.environ.get("FLOU_HDF5_DIR", "/opt/conda")

ARCH = "gfx950"
BASE = ["hipcc", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument"]
EXACT = ["-ffp-contract=off"]


def units():
    u = [("capi.o", "capi.cpp", EXACT), ("reduce.o", "reduce.hip", EXACT), ("limiter.o", "limiter.hip", EXACT),
         ("output.o", "output.hip", []),
         ("rhs_fast.o", "rhs_fast.hip", ["-ffp-contract=fast"])]
    # the fast adjoint, one unit per two-point flux (StdAverage, Chandrasekhar, Kennedy-Gruber)
    for tp, name in ((0, "std"), (1, "ch"), (2, "kg")):
        u.append((f"vjp_fast_{name}.o", "vjp_fast.hip", ["-ffp-contract=fast", f"-DFLOUHIP_FVJP_TP={tp}"]))
    for prec in ("F64", "F32"):
        for nd in (1, 2, 3):
            u.append((f"rhs_plain_{prec.lower()}_{nd}d.o", "rhs_plain.hip",
                      EXACT + [f"-DFLOUHIP_PREC_{prec}", f"-DFLOUHIP_ND={nd}"]))
            u.append((f"vjp_plain_{prec.lower()}_{nd}d.o", "vjp_plain.hip",
                      [f"-DFLOUHIP_PREC_{prec}", f"-DFLOUHIP_ND={nd}"]))
    return u


def _compile(obj, src, flags):
    out = os.path.join(BUILD, obj)
    cmd = BASE + flags + ["-c", os.path.join(CSRC, src), "-o", out]
    t = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(line for line in (p.stdout + p.stderr).splitlines()
                     if "packed-fp32-ops" not in line and line.strip())
    return obj, p.returncode, msgs, time.time() - t


def build(jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    us = units()
    objs = []
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, rc, msgs, dt in ex.map(lambda a: _compile(*a), us):
            objs.append(os.path.join(BUILD, obj))
            if verbose:
                print(f"  [{dt:5.1f}s] {obj}" + ("" if rc == 0 else "  FAILED"), flush=True)
            if rc != 0:
                failed.append((obj, msgs))
            elif msgs and verbose:
                print(msgs)
    if failed:
        for obj, msgs in failed:
            print(f"--- {obj}\n{msgs}", file=sys.stderr)
        raise RuntimeError("libflouhip build failed")
    tmp = LIB + ".tmp"
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC"] + objs + ["-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print(f"  built {LIB}")
    build_vtkhdf(verbose)
    return LIB


def build_vtkhdf(verbose=True):
    """libflouvtkhdf.so: g++, HDF5's C library by run path, libstdc++ linked statically so
    that the HDF5 directory's own C++ runtime is never picked up in its place."""
    inc, libdir = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    if not os.path.isfile(os.path.join(inc, "hdf5.h")):
        if verbose:
            print(f"  HDF5 not found under {HDF5_DIR}: {os.path.basename(VTKHDF_LIB)} not built "
                  "(VTKHDF output unavailable)")
        return None
    tmp = VTKHDF_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{inc}", os.path.join(CSRC, "vtkhdf.cpp"),
           f"-L{libdir}", "-lhdf5", f"-Wl,-rpath,{libdir}", "-static-libstdc++", "-static-libgcc", "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("libflouvtkhdf build failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, VTKHDF_LIB)
    if verbose:
        print(f"  built {VTKHDF_LIB}")
    return VTKHDF_LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    build(a.j)
