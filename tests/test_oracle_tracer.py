"""The CPU restatement of the path tracer's loop, pinned without the reference
(the reference renderer needs CUDA, OptiX/Vulkan and nanovdb, none of which
is here):

- the LCG stream is the published Numerical Recipes ranqd1 sequence;
- the seed hash is MurmurHash3_x86_32 (scikit-learn's implementation) up to
  the length word, related through the invertible finaliser;
- sin / cos / exp agree with libm to 1e-15, and uniform-sphere / HG
  directions are unit vectors whose HG cosine has the closed-form mean g;
- the index-space march returns the first occupied voxel other than the
  previous one, as brute force over every voxel box finds it;
- the loop with constant networks reproduces closed forms: never occluded ->
  every path escapes at bounce 0 with the background, after as many
  visibility queries as occupied voxels along the camera ray; always occluded
  with a constant throughput c -> an escaped path at bounce b carries
  (4 pi c)^b times the environment.
"""
import math

import numpy as np
import pytest

from oracle import tracer as T

# Numerical Recipes in C, 2nd ed., section 7.1 ("ranqd1"): idum = 1664525 idum + 1013904223 from 0
NR_RANQD1 = [0x3C6EF35F, 0x47502932, 0xD1CCF6E9, 0xAAF95334, 0x6252E503, 0x9F2EC686, 0x57FE6C2D, 0xA3D95FA8,
             0x81FDBEE7, 0x94F0AF1A, 0xCBF633B1]


def test_lcg_is_numerical_recipes_ranqd1():
    states, floats = T.lcg_stream(0, len(NR_RANQD1))
    assert states == NR_RANQD1
    assert floats == [float(np.float32(s)) * 2.0 ** -32 for s in NR_RANQD1]


def _fmix_inverse(h: int) -> int:
    m = 0xFFFFFFFF
    h ^= h >> 16
    h = (h * pow(0xC2B2AE35, -1, 1 << 32)) & m
    h ^= (h >> 13) ^ (h >> 26)
    h = (h * pow(0x85EBCA6B, -1, 1 << 32)) & m
    h ^= h >> 16
    return h


def test_seed_hash_is_murmur3_body():
    from sklearn.utils import murmurhash3_32
    rng = np.random.default_rng(1)
    for _ in range(200):
        index, frame, linear = (int(v) for v in rng.integers(0, 2 ** 32, 3, dtype=np.uint64))
        key = np.array([linear, index], np.uint32).tobytes()
        ref = murmurhash3_32(key, seed=frame, positive=True)
        # MurmurHash3_x86_32 xors the key length (8) into the state before the finaliser
        assert _fmix_inverse(T.lcg_seed(index, frame, linear)) ^ 8 == _fmix_inverse(ref)


def test_sincos_exp_against_libm():
    xs = np.concatenate([np.linspace(0.0, 2.0 * math.pi, 2001), np.random.default_rng(2).uniform(-7, 7, 500)])
    for x in xs:
        s, c = T.sincos(float(x))
        assert abs(s - math.sin(x)) <= 1e-15 and abs(c - math.cos(x)) <= 1e-15
    for x in np.linspace(-30.0, 30.0, 601):
        assert abs(T.exp(float(x)) - math.exp(x)) <= 4e-16 * math.exp(x)


def test_direction_samplers():
    rng = np.random.default_rng(3)
    u = rng.uniform(size=(4000, 2)).astype(np.float32)
    d = np.array([T.sample_dir(float(a), float(b))[0] for a, b in u])
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    assert np.abs(d.mean(axis=0)).max() < 0.05  # uniform on the sphere
    # the closed form of square_to_uniform_sphere in float64
    phi = 2.0 * np.pi * u[:, 0].astype(np.float64)
    ct = u[:, 1].astype(np.float64) * 2 - 1
    st = np.sqrt(1 - ct * ct)
    assert np.abs(d - np.stack([st * np.cos(phi), st * np.sin(phi), ct], 1)).max() < 2e-6
    assert T.sample_dir(0.3, 0.7)[1] == pytest.approx(4 * math.pi, rel=1e-7)
    # Henyey-Greenstein about the direction -w_o: E[cos(wi, -w_o)] = g; pdf integrates to 1
    g, w_o = 0.5, np.array([0.0, 0.6, 0.8], np.float32)
    cos, inv_pdf = [], []
    for a, b in u:
        w, ip = T.sample_dir(float(a), float(b), hg=True, w_o=w_o, g=g)
        cos.append(float(np.dot(w, -w_o)))
        inv_pdf.append(ip)
    assert abs(np.mean(cos) - g) < 0.02
    assert abs(np.mean(inv_pdf) / (4 * math.pi) - 1.0) < 0.05  # E_pdf[1/pdf] = area of the sphere


def _random_grid(R, density, seed):
    occ = np.random.default_rng(seed).uniform(size=(R, R, R)) < density  # [z][y][x]
    bits = np.packbits(occ.reshape(-1).astype(np.uint8)).view(">u4").astype(np.uint32)
    return occ, bits


def _brute_first(occ, o, d, t_min, prev):
    """First occupied voxel box (other than prev) met by o + t d beyond t_min, float64."""
    R = occ.shape[0]
    best, best_t, second = None, np.inf, np.inf
    zz, yy, xx = np.nonzero(occ)
    lo = np.stack([xx, yy, zz], 1).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (lo - o) / d
        t1 = (lo + 1 - o) / d
    tn = np.nanmax(np.minimum(t0, t1), axis=1)
    tf = np.nanmin(np.maximum(t0, t1), axis=1)
    g0 = np.maximum(np.max(np.minimum(-o / d, (R - o) / d)), t_min)
    for i in range(len(lo)):
        if tuple(lo[i].astype(int)) == tuple(prev) or not (tn[i] < tf[i]) or tf[i] < g0:
            continue
        t = max(tn[i], g0)
        if t < best_t:
            best, second, best_t = tuple(lo[i].astype(int)), best_t, t
        elif t < second:
            second = t
    return best, best_t, second


def test_march_first_occupied_voxel_by_brute_force():
    R = 12
    occ, bits = _random_grid(R, 0.04, 4)
    sc = T.scene(bits, R, (0, 0, 0), (1, 1, 1), 1.0 / R)
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(600):
        o = rng.uniform(-2, R + 2, 3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        prev = tuple(int(v) for v in rng.integers(0, R, 3)) if rng.uniform() < 0.3 else (-1, -1, -1)
        hit, vox, tn, tf = T.march(sc, o, d, prev=prev)
        ref, t_ref, t_second = _brute_first(occ, o.astype(np.float64), d.astype(np.float64), 0.0, prev)
        if ref is not None and t_second - t_ref < 1e-4:
            continue  # an edge or corner tie: either voxel is a valid first hit
        checked += 1
        assert hit == (ref is not None)
        if hit:
            assert vox == ref
            assert abs(max(tn, 0.0) - t_ref) < 1e-4 * (1 + abs(t_ref))
    assert checked > 500


def _scene_and_camera(R=16, seed=6, W=24, H=20):
    occ, bits = _random_grid(R, 0.06, seed)
    sc = T.scene(bits, R, (0, 0, 0), (1, 1, 1), 1.0 / R)
    cam = T.Camera()
    cam.position[:] = [0.5, 0.45, -1.2]
    cam.dir_top_left[:] = [-0.5, 0.5, 1.0]
    cam.du[:] = [1.0, 0.0, 0.0]
    cam.dv[:] = [0.0, -1.0, 0.0]
    cam.width, cam.height = W, H
    return occ, sc, cam


def _params(R, **kw):
    p = T.Params()
    p.lod, p.max_depth = 0, 3
    p.background[:] = [0.25, 0.5, 0.75]
    p.envmap[:] = [1.0, 2.0, 3.0]
    p.inference_min_extent_dilation = p.voxel_extent_dilation_outward = 1e-4
    p.rr_start_bounce, p.hg_g, p.n_thr_inputs, p.grid_res = 1, 0.5, 9, R
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_loop_never_occluded_closed_form():
    occ, sc, cam = _scene_and_camera()
    R = occ.shape[0]
    p = _params(R)
    illum, inf, bnc = T.trace(sc, cam, p, lambda which, x: np.zeros((len(x), 3), np.float32))
    assert np.array_equal(illum, np.tile(np.float32([0.25, 0.5, 0.75]), (len(illum), 1)))
    assert (bnc == 0).all()
    # visibility queries = occupied voxels met by the camera ray: walk it with the march hook
    rng_seed = [T.lcg_seed(0, 0, i) for i in range(cam.width * cam.height)]
    for i in range(0, cam.width * cam.height, 7):
        s = np.uint32(rng_seed[i])
        _, f = T.lcg_stream(int(s), 2)
        px, py = i % cam.width, i // cam.width
        xf = np.float32((np.float32(px) + np.float32(f[0])) / np.float32(cam.width))
        yf = np.float32((np.float32(py) + np.float32(f[1])) / np.float32(cam.height))
        d = (xf * np.float32(cam.du[:]) + yf * np.float32(cam.dv[:])) + np.float32(cam.dir_top_left[:])
        d = (d * (np.float32(1) / np.sqrt(np.float32(np.dot(d, d))))).astype(np.float32)
        o = (np.float32(R) * (np.float32(cam.position[:]) - np.float32(0)) / np.float32(1)).astype(np.float32)
        count, prev, tn, tf = 0, (-1, -1, -1), -1.0, -1.0
        while True:
            hit, prev, tn, tf = T.march(sc, o, d, tn, tf, prev)
            if not hit:
                break
            count += 1
        assert inf[i] == count, i
    assert inf.max() > 1


def test_loop_always_occluded_closed_form():
    occ, sc, cam = _scene_and_camera(seed=7)
    R = occ.shape[0]
    c = np.float32(0.05)
    p = _params(R)

    def net(which, x):
        return np.ones((len(x), 3), np.float32) if which == 0 else np.full((len(x), 3), c, np.float32)

    illum, inf, bnc = T.trace(sc, cam, p, net)
    gain = np.float32(c * np.float32(4 * math.pi))
    env = np.float32([1.0, 2.0, 3.0])
    for i in range(len(illum)):
        b = int(bnc[i])
        if b == 0 and np.array_equal(illum[i], np.float32([0.25, 0.5, 0.75])):
            continue  # escaped straight away
        if np.all(illum[i] == 0):
            assert b == p.max_depth - 1  # stopped in a voxel at the last bounce
            continue
        assert b >= 1
        np.testing.assert_allclose(illum[i], gain ** b * env, rtol=1e-6)
        assert inf[i] == b  # one visibility query per bounce, each occluded
    assert (bnc == 1).any() and (bnc == 2).any()
