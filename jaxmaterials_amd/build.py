"""Build recipe of liblippmannschwinger.so (gfx950, in-tree).

``python -m jaxmaterials_amd.build`` compiles every HIP source of the library
into ``jaxmaterials_amd/liblippmannschwinger.so``.  hipcc cross-compiles for
gfx950 without a GPU.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "liblippmannschwinger.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

SOURCES = ["ls_solver.hip", "microstructure.hip"]
CONV_SOURCES = ["conv3d.hip"]
CONV_LIB = os.path.join(HERE, "libls_warmstart.so")

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


def _compile(sources, out, verbose):
    srcs = [os.path.join(CSRC, s) for s in sources]
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed building {os.path.basename(out)}")
    os.replace(tmp, out)
    return out


def build(verbose=False, only=None):
    """Compile the solver library (and the warm-start conv library if present)."""
    outs = []
    if only in (None, "solver"):
        outs.append(_compile(SOURCES, LIB_PATH, verbose))
    if only in (None, "conv") and all(os.path.exists(os.path.join(CSRC, s)) for s in CONV_SOURCES):
        outs.append(_compile(CONV_SOURCES, CONV_LIB, verbose))
    return outs


if __name__ == "__main__":
    for p in build(verbose=True, only=(sys.argv[1] if len(sys.argv) > 1 else None)):
        print("built", p)
