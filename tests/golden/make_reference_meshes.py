"""Writes reference_meshes.npz: tetrahedral meshes built by the reference's own mesh code
(OrthHexMesh(lb, ub, {Nx,Ny,Nz}) + GeneralMesh::split_hex6_scan / split_hex5_scan +
rebuild_cell_topology + validate_mesh, include/Mesh/OrthHexMesh.h:7-106,
src/Mesh/GeneralMesh.cpp:4-129), through the reference build oracle/_ref/libref.so.

The meshes keep the reference's tet vertex order (from std::unordered_set iteration,
GeneralMesh.cpp:25-47) and face node order, which exercise every face/vertex-order
combination of the residual.  Run in a container that has the reference tree:

    make -C oracle && python3 tests/golden/make_reference_meshes.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle.binding import Reference  # noqa: E402

CASES = {
    "box6": dict(n=(3, 3, 2), lb=(0.0, 0.0, 0.0), ub=(10.0, 10.0, 2.5), split=6),
    "box5": dict(n=(2, 3, 2), lb=(0.0, 0.0, 0.0), ub=(10.0, 10.0, 2.5), split=5),
}


def main():
    ref = Reference()
    out = {}
    for name, c in CASES.items():
        m = ref.box_mesh(c["n"], c["lb"], c["ub"], c["split"])
        for k, v in m.items():
            out[f"{name}_{k}"] = v
        print(name, {k: v.shape for k, v in m.items()})
    np.savez_compressed(Path(__file__).with_name("reference_meshes.npz"), **out)


if __name__ == "__main__":
    main()
