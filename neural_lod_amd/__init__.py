"""neural_lod_amd — MI355X (gfx950) neural-LoD hot path.

Host side of libnlod.so: the voxelizer (``voxelizer``), the tcnn-style
network-with-input-encoding model and its trainer (``tcnn``), and seeded
workload generators (``synthetic``).  Compute runs only in libnlod's HIP
kernels; importing a submodule that needs the library raises if it is not
built.
"""
__all__ = ["voxelizer", "synthetic"]
__version__ = "0.1.0"
