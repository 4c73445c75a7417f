"""Builds the HIP shared library libkscr.so (gfx950) in-tree.

    python -m kamisimulator_amd.build

The library holds the step kernels and the C ABI of include/kscr.h.  It is
compiled with hipcc directly (no torch extension): the boundary is a plain C
ABI, and the built .so is what travels with the source tree.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libkscr.so")

SOURCES = ["kscr_api.hip", "kscr_api_adjoint.hip", "kscr_api_batch.hip", "kscr_api_lbvh.hip", "kscr_step.hip",
           "kscr_colored.hip", "kscr_adjoint.hip", "kscr_contact.hip", "kscr_lbvh.hip", "kscr_sagfree.cpp",
           "kscr_meshsdf.hip", "kscr_selftest.hip", "kscr_lbvh_static.hip", "kscr_strain.hip"]
HEADERS = ["kscr_math.h", "kscr_device.h", "kscr_solve.h", "kscr_contact.h", "kscr_lbvh.h", "kscr_api_internal.h", "kscr_sync.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-function",
    # keep hipcc from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # IEEE division and square root, and no FMA contraction: every FP32
    # operation rounds once, in the order the source states, so the device
    # results are bit-identical to the FP32 CPU restatement
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-ffp-contract=off",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


HOST = os.path.join(PKG, "host")
HOST_SOURCES = ["json_lite.cpp", "hair.cpp", "curve_io.cpp", "yarn_scene.cpp", "scr_solver.cpp", "solver_cli.cpp"]
HOST_HEADERS = ["json_lite.h", "hair.h", "curve_io.h", "yarn_scene.h", "scr_solver.h"]
CLI = os.path.join(PKG, "kscr_solver_cli")
ROCM_LIB = "/opt/rocm/lib"


def build(verbose=True):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(REPO, "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose=True):
    """The host-side solver (C++ over the C ABI) and its command-line driver,
    compiled with the host compiler and linked against libkscr.so."""
    cxx = shutil.which("g++") or "g++"
    tmp = CLI + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(REPO, "include"), "-I", HOST, "-o", tmp] + \
          [os.path.join(HOST, s) for s in HOST_SOURCES] + \
          ["-L", PKG, "-lkscr", "-L", ROCM_LIB, "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + ROCM_LIB]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, CLI)
    return CLI


def needs_build():
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    deps = [os.path.join(CSRC, s) for s in SOURCES + HEADERS] + [os.path.join(REPO, "include", "kscr.h")]
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


if __name__ == "__main__":
    build()
    sys.exit(0)
