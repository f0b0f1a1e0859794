"""Build recipe of liblippmannschwinger.so (gfx950, in-tree).

``python -m jaxmaterials_amd.build`` compiles every HIP source of the library
(hipcc) and its host-only C++ sources (g++) into
``jaxmaterials_amd/liblippmannschwinger.so``.  hipcc cross-compiles for gfx950
without a GPU.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "liblippmannschwinger.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

# every HIP source of the library; each compiles to its own object in parallel
SOURCES = ["ls_solver.hip", "microstructure.hip", "conv3d.hip", "derivatives.hip", "stream_probe.hip",
           "phasefield.hip"]
# host-only C++ sources, compiled by g++ (the reference fibre movement is
# compiled by g++ in the reference too, and its results depend on the
# standard library's random engines and distributions)
HOST_SOURCES = ["fibres.cpp"]
# (-fno-math-errno lets sqrt vectorise; it changes no result)
GXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-math-errno"]

HIPCC_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # host arithmetic of the fibre placement must not contract to FMA
    "-Xarch_host", "-ffp-contract=off",
    "-Wno-unused-value",
    "-Wno-unused-result",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; ROCm is required to build liblippmannschwinger.so")
    return exe


def gxx():
    exe = shutil.which("g++")
    if not exe:
        raise RuntimeError("g++ not found; it builds the host sources of liblippmannschwinger.so")
    return exe


def all_library_paths():
    """Every shared library this package loads."""
    return [LIB_PATH]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def build(verbose=False):
    """Compile every source to an object (in parallel), then link the library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs, procs = [], []
    for s in srcs:
        obj = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        flags = [f for f in HIPCC_FLAGS if f != "-shared"]
        procs.append((s, _run([hipcc()] + flags + ["-c", "-o", obj, os.path.join(CSRC, s)], verbose)))
    for s in HOST_SOURCES:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        obj = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        procs.append((s, _run([gxx()] + GXX_FLAGS + ["-c", "-o", obj, src], verbose)))
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(s)
    if failed:
        raise RuntimeError(f"hipcc failed on {failed}")
    tmp = LIB_PATH + ".tmp"
    p = _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    out, _ = p.communicate()
    if p.returncode != 0:
        sys.stderr.write(out)
        raise RuntimeError(f"linking {LIB_NAME} failed")
    os.replace(tmp, LIB_PATH)
    return [LIB_PATH]


if __name__ == "__main__":
    for p in build(verbose=True):
        print("built", p)
