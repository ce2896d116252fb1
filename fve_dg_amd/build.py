"""Build recipe for libfvdg.so (hipcc, gfx950) -- run from the repository root
or through ``__graft_entry__.build()`` (which also builds the test checker under oracle/).

    python -m fve_dg_amd.build

Each translation unit is compiled to an object in parallel (hipcc -c), then linked.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libfvdg.so"

SOURCES = ["fvdg_host.cpp", "fvdg_solver.hip", "fvdg_ensemble.hip", "fvdg_implicit.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    "--offload-arch=gfx950",
    # packed FP32 VALU forms are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _clean(text):
    return "\n".join(l for l in text.splitlines()
                     if "packed-fp32-ops" not in l and "not a recognized feature" not in l)


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    return res.returncode, _clean(res.stdout + res.stderr)


def build_library(verbose=False):
    with tempfile.TemporaryDirectory(prefix="fvdg_build_") as tmp:
        objs = [str(Path(tmp) / (Path(s).stem + ".o")) for s in SOURCES]
        cmds = [[HIPCC, *FLAGS, "-c", str(CSRC / s), "-o", o] for s, o in zip(SOURCES, objs)]
        if verbose:
            for c in cmds:
                print(" ".join(c), flush=True)
        with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
            results = list(ex.map(_run, cmds))
        for (rc, out), s in zip(results, SOURCES):
            if rc != 0:
                raise RuntimeError(f"hipcc failed on {s} ({rc}):\n{out}")
            if out.strip() and verbose:
                print(out)
        link = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", str(LIB) + ".tmp"]
        rc, out = _run(link)
        if rc != 0:
            raise RuntimeError(f"link failed ({rc}):\n{out}")
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build_library(verbose=True)
    print("built", LIB, file=sys.stderr)
