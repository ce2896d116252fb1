"""Build recipe for libfvdg.so (hipcc, gfx950) -- run from the repository root
or through ``__graft_entry__.build()`` (which also builds the test checker under oracle/).

    python -m fve_dg_amd.build
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libfvdg.so"

SOURCES = ["fvdg_host.cpp", "fvdg_solver.hip", "fvdg_ensemble.hip", "fvdg_implicit.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    # packed FP32 VALU forms are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def build_library(verbose=False):
    cmd = [HIPCC, *FLAGS, *[str(CSRC / s) for s in SOURCES], "-o", str(LIB) + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{out}")
    if out.strip() and verbose:
        print(out)
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build_library(verbose=True)
    print("built", LIB, file=sys.stderr)
