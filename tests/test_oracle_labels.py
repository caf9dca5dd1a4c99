"""The CPU restatement of the visibility training samples (oracle), pinned by
brute force, a closed-form plane intersection and sampling statistics."""
import numpy as np
import pytest

from neural_lod_amd import synthetic as S


def test_lists_equal_brute_force(oracle):
    """Per-voxel triangle lists only accelerate: testing every triangle gives
    the same inputs and labels."""
    tris = S.blob_mesh(20_000, seed=3)
    a, la, n_occ = oracle.visibility_samples(tris, 48, seed=11, first=500, n=4096)
    b, lb, _ = oracle.visibility_samples(tris, 48, seed=11, first=500, n=4096, brute_force=True)
    assert n_occ > 0
    assert np.array_equal(a, b)
    assert np.array_equal(la, lb)
    assert 0 < la.sum() < la.size


def test_plane_closed_form(oracle):
    """A square at z = 0.3 spanning the box: a ray hits it where
    t = (0.3 - o_z) / d_z >= 0, and the label says whether o + t d lies in the
    starting voxel (FP64 closed form; near-boundary cases skipped)."""
    z0 = np.float32(0.3)
    sq = np.array([[0, 0, z0, 1, 0, z0, 1, 1, z0], [0, 0, z0, 1, 1, z0, 0, 1, z0]], np.float32)
    res = 15  # odd: the plane (centred in the cubic box) cuts the middle of a voxel layer
    n = 20000
    x, lab, n_occ = oracle.visibility_samples(sq, res, seed=5, first=0, n=n)
    info = oracle.voxinfo(*oracle.mesh_bbox(sq), res, False, 2)
    bmin = np.array(info.bbox_min, np.float64)
    ext = np.array(info.bbox_max, np.float64) - bmin
    unit = np.array(info.unit, np.float64)
    o = bmin + x[:, :3].astype(np.float64) * ext
    d = x[:, 3:].astype(np.float64) * 2 - 1
    v = np.floor((o - bmin) / unit)
    centre = bmin + (v + 0.5) * unit
    t = (np.float64(z0) - o[:, 2]) / d[:, 2]
    p = o + t[:, None] * d
    margin = (0.5 * unit - np.abs(p - centre)).min(axis=1)
    expect = (t >= 0) & (margin >= 0)
    clear = (np.abs(margin) > 1e-4 * unit[0]) & (np.abs(t) > 1e-4) & (np.abs(o[:, 2] - z0) > 1e-5)
    clear &= np.all(np.abs((o - bmin) / unit - np.round((o - bmin) / unit)) > 1e-4, axis=1)  # origin off faces
    assert clear.mean() > 0.95
    assert np.array_equal(lab[clear, 0] == 1.0, expect[clear])
    assert n_occ == res * res * (1 if (float(z0) - info.bbox_min[2]) / info.unit[2] % 1 else 2)


def test_sampling_statistics(oracle):
    """Directions uniform on the sphere, origins uniform inside their voxel,
    and every sample a function of its index alone."""
    tris = S.blob_mesh(5000, seed=1)
    n = 1 << 16
    x, _, _ = oracle.visibility_samples(tris, 32, seed=2, first=0, n=n)
    d = x[:, 3:].astype(np.float64) * 2 - 1
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert np.all(np.abs(d.mean(axis=0)) < 0.02)
    assert np.all(np.abs((d ** 2).mean(axis=0) - 1 / 3) < 0.01)
    frac = (x[:, :3].astype(np.float64) * 32) % 1.0  # position inside the voxel
    assert np.all(np.abs(frac.mean(axis=0) - 0.5) < 0.01)
    y, _, _ = oracle.visibility_samples(tris, 32, seed=2, first=1000, n=50)
    assert np.array_equal(y, x[1000:1050])
    z, _, _ = oracle.visibility_samples(tris, 32, seed=3, first=1000, n=50)
    assert not np.array_equal(z, y)
