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


def _splitmix_uniforms(seed: int, k: int):
    """The first k draws u = (x >> 40) * 2^-24 of splitmix64 started at `seed`."""
    m = (1 << 64) - 1
    s, out = seed & m, []
    for _ in range(k):
        s = (s + 0x9E3779B97F4A7C15) & m
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        z ^= z >> 31
        out.append(np.float32((z >> 40) * 2.0 ** -24))
    return out


def test_throughput_lists_equal_brute_force(oracle):
    tris = S.blob_mesh(20_000, seed=0)
    a = oracle.throughput_samples(tris, 32, 3, seed=5, first=0, n=384, spp=8)
    b = oracle.throughput_samples(tris, 32, 3, seed=5, first=0, n=384, spp=8, brute_force=True)
    for u, v in zip(a[:3], b[:3]):
        assert np.array_equal(u, v)
    x, t, valid, n_occ = a
    assert n_occ[0] > n_occ[1] > n_occ[2] > 0
    assert 0.1 < valid.mean() < 1.0
    assert np.all(t[valid == 0] == 0) and np.all(t >= 0)
    assert set(np.unique(x[:, 9])) == {0.0, 1.0, 2.0}


def test_throughput_single_plane_closed_form(oracle):
    """One bounce off a flat square: every valid path hits the triangle under
    the voxel, its shadow ray along wi always escapes, so the target is
    albedo / pi * (wi . n) when wi lies on wo's side of the plane, else 0."""
    z0 = np.float32(0.3)
    sq = np.array([[0, 0, z0, 1, 0, z0, 1, 1, z0], [0, 0, z0, 1, 1, z0, 0, 1, z0]], np.float32)
    res = 15
    x, t, valid, _ = oracle.throughput_samples(sq, res, 1, seed=3, first=0, n=2000, spp=4, max_depth=1,
                                               albedo_seed=11)
    info = oracle.voxinfo(*oracle.mesh_bbox(sq), res, False, 2)
    bmin, ext = np.array(info.bbox_min, np.float64), np.array(info.bbox_max, np.float64) - np.array(info.bbox_min)
    unit = np.array(info.unit, np.float64)
    c = bmin + x[:, :3].astype(np.float64) * ext
    lo, hi = c - unit / 2, c + unit / 2
    below = lo[:, 0] > hi[:, 1] + 1e-6       # voxel entirely under the diagonal: triangle 0
    above = lo[:, 1] > hi[:, 0] + 1e-6       # triangle 1
    wo = x[:, 3:6].astype(np.float64) * 2 - 1
    wi = x[:, 6:9].astype(np.float64) * 2 - 1
    side = np.sign(wo[:, 2])
    cos_i = wi[:, 2] * side
    alb = [np.array(_splitmix_uniforms(11 ^ (tri * 0x9E3779B97F4A7C15), 3), np.float64) * 0.8 + 0.1
           for tri in (0, 1)]
    checked = 0
    for mask, a in ((below, alb[0]), (above, alb[1])):
        sel = mask & (valid == 1) & (np.abs(cos_i) > 1e-4)
        expect = np.where(cos_i[sel, None] > 0, a[None, :] / np.pi * cos_i[sel, None], 0.0)
        np.testing.assert_allclose(t[sel], expect, rtol=2e-5, atol=1e-7)
        checked += int(sel.sum())
    assert checked > 500
