"""The CPU restatement of the LoD pyramid and the first-active-voxel march,
pinned by a numpy block-OR and by brute force over every occupied voxel's box
(float64 slab intervals: the first voxel along the ray is the occupied one
with the smallest entry among those the ray still reaches)."""
import numpy as np
import pytest


def _random_table(res, density, seed):
    rng = np.random.default_rng(seed)
    occ = rng.random(res ** 3) < density
    bits = np.zeros(((res ** 3 + 31) // 32) * 32, np.uint8)
    bits[: res ** 3] = occ
    return np.packbits(bits).view(">u4").astype(np.uint32), occ.reshape(res, res, res)  # [z][y][x]


def _unpack(table, res):
    return np.unpackbits(table.astype(">u4").view(np.uint8))[: res ** 3].reshape(res, res, res).astype(bool)


def test_downsample_is_block_or(oracle):
    table, occ = _random_table(16, 0.03, 1)
    parent = oracle.downsample(table, 16)
    expect = occ.reshape(8, 2, 8, 2, 8, 2).any(axis=(1, 3, 5))
    assert np.array_equal(_unpack(parent, 8), expect)


def _brute_force(occ, res, bmin, unit, o, d, t_min, prev):
    zz, yy, xx = np.nonzero(occ)
    lin = xx + yy * res + zz * res * res
    v = np.stack([xx, yy, zz], 1).astype(np.float64)
    lo = bmin + v * unit
    hi = bmin + (v + 1) * unit
    out = np.full(len(o), -1, np.int64)
    near = np.full(len(o), -1.0)
    # the grid's own entry bounds the start, as the march clips to [0, R]^3 first
    glo, ghi = bmin, bmin + res * unit
    for i in range(len(o)):
        with np.errstate(divide="ignore", invalid="ignore"):
            t0 = (lo - o[i]) / d[i]
            t1 = (hi - o[i]) / d[i]
            g0, g1 = (glo - o[i]) / d[i], (ghi - o[i]) / d[i]
        tn = np.nanmax(np.minimum(t0, t1), axis=1)
        tf = np.nanmin(np.maximum(t0, t1), axis=1)
        start = max(t_min[i], np.max(np.minimum(g0, g1)))
        ok = (tn <= tf) & (tf >= start) & (lin != prev[i])
        if ok.any():
            k = np.argmin(np.where(ok, tn, np.inf))
            out[i], near[i] = lin[k], tn[k]
    return out, near


def test_march_matches_brute_force(oracle):
    res = 16
    table, occ = _random_table(res, 0.02, 7)
    bmin = np.array([-1.0, 0.5, 2.0])
    unit = np.array([0.125, 0.125, 0.125])
    rng = np.random.default_rng(3)
    n = 3000
    o = (bmin + rng.uniform(-0.5, res + 0.5, size=(n, 3)) * unit).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    t_min = rng.uniform(-0.2, 0.3, n).astype(np.float32)
    prev = np.where(rng.random(n) < 0.3, rng.integers(0, res ** 3, n), 0xFFFFFFFF).astype(np.uint32)
    vox, tn, tf = oracle.march(table, res, bmin, unit, o, d, t_min, prev)
    exp, near = _brute_force(occ, res, bmin, unit, o.astype(np.float64), d.astype(np.float64),
                             t_min.astype(np.float64), prev.astype(np.int64))
    got = np.where(vox == 0xFFFFFFFF, -1, vox.astype(np.int64))
    assert (got >= 0).sum() > 300 and (got < 0).sum() > 300
    assert np.array_equal(got, exp), f"{int((got != exp).sum())} rays differ"
    hit = got >= 0
    np.testing.assert_allclose(tn[hit], near[hit], rtol=1e-4, atol=1e-5)
    assert np.all(tf[hit] >= tn[hit])
