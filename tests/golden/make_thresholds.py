"""Copies two of the reference's per-voxel optimal-threshold files (raw float32,
one value per occupied voxel in Morton order) as fixtures:
neural_lod/scenes/arcade/optimal_thresholds/arcade_world_final_lod_{5,6}_optimal_thresholds.bin
(683 and 146 values).  Data only; run from the repository root."""
import shutil
from pathlib import Path

SRC = Path("/root/reference/neural_lod/scenes/arcade/optimal_thresholds")
DST = Path(__file__).resolve().parent

if __name__ == "__main__":
    for lod in (5, 6):
        name = f"arcade_world_final_lod_{lod}_optimal_thresholds.bin"
        shutil.copyfile(SRC / name, DST / name)
        print(name, (DST / name).stat().st_size, "bytes")
