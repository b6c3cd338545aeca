"""Build libtahoe_gpu.so for gfx950 with hipcc (in-tree, no JIT cache).

Usage: python -m tahoe_amd.build   (or __graft_entry__.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtahoe_gpu.so")
HOST = os.path.join(HERE, "host")
HOST_LIB = os.path.join(HERE, "libtahoe_host.so")
HOST_EXE = os.path.join(HERE, "bin", "tahoe_hip")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
# C++ host layer (plugin-shaped classes + XML/geometry front end) over the C ABI
HOST_SOURCES = ["TotalLagrangianHIPT.cpp", "HIPMSRMatrixT.cpp", "NLSolverHIP.cpp", "InputDeck.cpp",
                "FEManagerHIP.cpp"]

SOURCES = ["tg_setup.cpp", "tg_amg.cpp", "tg_api.cpp", "k_element.hip", "k_q1p0.hip", "k_gather.hip", "k_solver.hip",
           "k_pcg.hip", "k_amg.hip", "k_gmres.hip", "k_graph.hip"]
HEADERS = ["tg_internal.h", "tg_kinematics.h", "tg_element.h", "../../include/tahoe_gpu.h"]
# per-source extra flags.  The PCG kernels are built without FMA
# contraction: every variant of the iteration (SpMV forms, fused direction,
# persistent workgroup) then rounds each expression as written, so variants
# that write the same expressions give bitwise the same iterates whatever
# code surrounds them (with contraction the compiler may fuse a*b + c*d
# either way round, kernel by kernel).  The kernels are memory-bound.
EXTRA = {"k_pcg.hip": ["-ffp-contract=off"]}

ARCH = os.environ.get("TAHOE_GPU_ARCH", "gfx950")
# packed FP32 instructions are disabled for all device code
FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-x", "hip",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if p and (os.path.isabs(p) and os.path.exists(p) or not os.path.isabs(p)):
            return p
    return "hipcc"


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    objs = []
    for s in SOURCES:
        o = os.path.join(OBJ, os.path.splitext(s)[0] + ".o")
        objs.append(o)
        cmd = [hipcc()] + FLAGS + EXTRA.get(s, []) + ["-c", os.path.join(CSRC, s), "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((s, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        for s, out in failed:
            sys.stderr.write(f"--- {s}\n{out}\n")
        raise RuntimeError("hipcc failed for " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose: bool = False) -> str:
    """libtahoe_host.so and the tahoe_hip driver, plain g++ against the C ABI."""
    cxx = os.environ.get("CXX", "g++")
    os.makedirs(os.path.dirname(HOST_EXE), exist_ok=True)
    common = ["-O2", "-std=c++17", "-Wall", "-Wextra", "-I", INCLUDE, "-I", HOST]
    tmp = HOST_LIB + ".tmp"
    cmd = [cxx] + common + ["-fPIC", "-shared", "-o", tmp] + [os.path.join(HOST, f) for f in HOST_SOURCES] + \
        ["-L", HERE, "-ltahoe_gpu", "-Wl,-rpath,$ORIGIN"]
    for c in (cmd,):
        if verbose:
            print(" ".join(c), flush=True)
        r = subprocess.run(c, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("host layer build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, HOST_LIB)
    tmp = HOST_EXE + ".tmp"
    cmd = [cxx] + common + ["-o", tmp, os.path.join(HOST, "tahoe_hip.cpp"), "-L", HERE, "-ltahoe_host",
                            "-ltahoe_gpu", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("tahoe_hip build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, HOST_EXE)
    return HOST_EXE


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
