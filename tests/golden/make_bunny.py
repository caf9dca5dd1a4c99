"""Writes tests/golden/bunny_tris.npy from the reference's voxelizer test mesh.

Input: neural_lod/cuda_voxelizer/test_models/bunny.OBJ (5,110 faces; the only
mesh fixture the reference's voxelizer ships).  Output: float32[n_faces][9]
triangle soup (v0 v1 v2, xyz each), vertices parsed from the OBJ text with
float32 rounding of the decimal literals.  Run once in the build container:
    python tests/golden/make_bunny.py
"""
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/neural_lod/cuda_voxelizer/test_models/bunny.OBJ")
DST = Path(__file__).resolve().parent / "bunny_tris.npy"


def main():
    verts, faces = [], []
    for line in SRC.read_text().splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            verts.append([np.float32(p) for p in parts[1:4]])
        elif parts[0] == "f":
            faces.append([int(p.split("/")[0]) - 1 for p in parts[1:4]])
    v = np.array(verts, np.float32)
    f = np.array(faces, np.int64)
    tris = np.ascontiguousarray(v[f].reshape(len(f), 9).astype(np.float32))
    np.save(DST, tris)
    print(DST, tris.shape)


if __name__ == "__main__":
    main()
