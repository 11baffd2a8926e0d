"""Builds libcpptraj_rmsd.so (gfx950) in-tree with hipcc.

    python -m cpptraj_amd.build [--verbose]

Outputs in cpptraj_amd/_lib/:
  libcpptraj_rmsd.so       the kernels and the C ABI (include/cpptraj_rmsd.h);
                           every HIP source is compiled for gfx950 only, with
                           packed-FP32 instruction selection disabled;
  libcpptraj_rmsd_host.so  the C++ host layer (host/*.cpp: rms2d / rms /
                           cluster-cache classes) linked against the former.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libcpptraj_rmsd.so")
HOSTLIB = os.path.join(LIBDIR, "libcpptraj_rmsd_host.so")
HOSTDIR = os.path.join(HERE, "host")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["capi.hip", "capi_grad.hip", "capi_batch.hip", "capi_noalloc.hip", "k_prep.hip",
           "k_pairwise.hip", "k_nofit.hip", "k_fit.hip", "k_grad.hip", "k_matrix.hip", "k_perres.hip", "k_avgcorr.hip",
           "k_symm.hip"]
HOST_SOURCES = ["host_core.cpp", "host_rms.cpp", "host_capi.cpp"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(LIBDIR, "obj", os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-I", CSRC, "-I", INCLUDE, "-c", src, "-o", obj]
    if verbose:
        cmd.append("-Rpass-analysis=kernel-resource-usage")
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [ln for ln in (r.stdout + r.stderr).splitlines()
            if "packed-fp32-ops" not in ln]
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.join(LIBDIR, "obj"), exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose: bool = False) -> str:
    """The host layer: plain C++ (g++), HIP runtime API only, linked to LIB."""
    objs = []
    for s in HOST_SOURCES:
        obj = os.path.join(LIBDIR, "obj", s + ".o")
        cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
               "-I", INCLUDE, "-I", HOSTDIR, "-I", os.path.join(ROCM, "include"), "-c",
               os.path.join(HOSTDIR, s), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {s}:\n" + r.stdout + r.stderr)
        if verbose and (r.stdout or r.stderr):
            print(r.stdout + r.stderr)
        objs.append(obj)
    tmp = HOSTLIB + ".tmp"
    cmd = ["g++", "-shared", "-fPIC", "-o", tmp, *objs, "-L", LIBDIR, "-lcpptraj_rmsd",
           "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, HOSTLIB)
    return HOSTLIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.verbose))
    sys.exit(0)
