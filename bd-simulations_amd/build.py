"""Builds libbdk_hip.so (HIP kernels + C-ABI) in-tree for gfx950, and the C++
driver bdk_sim (the main.cpp-style run loop over the C-ABI).

Usage: python bd-simulations_amd/build.py [--param-header path/to/parameter.h]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "libbdk_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Device code: gfx950 only; no FP contraction (position arithmetic must round
# like the reference); packed-FP32 instructions disabled.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib():
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    for src in ("bdk_kernels.hip", "bdk_noalloc.hip", "bdk_api.cpp", "bdk_serve.cpp"):
        obj = os.path.join(HERE, "_build", src + ".o")
        extra = ["-x", "hip"] if src.endswith(".hip") else []
        run([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)


def driver_cmd(main_src, exe, param_header=None, extra=()):
    """g++ line of a driver over the host shim, compiled against the constants
    of a reference-format parameter.h (default: host/small_run/parameter.h).
    host/ itself holds no parameter.h, so the quoted #include "parameter.h" of
    every shim file falls through to the -I directory of the header given."""
    inc = os.path.dirname(os.path.abspath(param_header)) if param_header else os.path.join(HOST, "small_run")
    return (["g++", "-O2", "-std=c++17", "-Wall", "-I" + inc, "-I" + HOST, "-I" + os.path.join(ROOT, "include")]
            + list(extra)
            + [os.path.join(HOST, "run_zustand.cpp"), os.path.join(HOST, "params_from_header.cpp"), main_src,
               "-L" + HERE, "-lbdk_hip", "-Wl,-rpath," + HERE, "-o", exe])


def build_driver(param_header=None):
    """bdk_sim: main.cpp's run loop (main.cpp:130-300) over the C-ABI;
    bdk_refcalls: the driver of tests/drop_in that uses only the reference's
    calls and parameter.h globals; and bdk_refdefault: the same driver
    against the reference's default values (tests/drop_in/refdefault)."""
    run(driver_cmd(os.path.join(HOST, "bdk_sim_main.cpp"), os.path.join(HERE, "bdk_sim"), param_header))
    ref_main = os.path.join(ROOT, "tests", "drop_in", "reference_calls_main.cpp")
    if os.path.exists(ref_main):
        run(driver_cmd(ref_main, os.path.join(HERE, "bdk_refcalls"), param_header))
        ref_hdr = os.path.join(ROOT, "tests", "drop_in", "refdefault", "parameter.h")
        if os.path.exists(ref_hdr):
            run(driver_cmd(ref_main, os.path.join(HERE, "bdk_refdefault"), ref_hdr))


if __name__ == "__main__":
    hdr = None
    if "--param-header" in sys.argv:
        hdr = sys.argv[sys.argv.index("--param-header") + 1]
    build_lib()
    if os.path.exists(os.path.join(HOST, "run_zustand.cpp")):
        build_driver(hdr)
