"""GPU tests of the path tracer (nlod_tracer_*), the reference's voxel_tracing
loop (wavefront_neural_throughput_visibility_lod.cu:464-720).

Parity: the CPU restatement of the loop (oracle/tracer_oracle.c), fed with
the library's own network outputs for each batch it forms, must give every
pixel's illumination, visibility-query count and final bounce bit for bit --
for the reference schedule in FP32 and FP16.  The persistent FP16 schedule
must equal the reference schedule bit for bit (same per-ray arithmetic; a
query's network output does not depend on its batch), with every option of
the loop, and its hipGraph replays must equal eager runs.  Against an FP64
network the FP32 image may differ only where a visibility lies within
rounding of its threshold."""
import numpy as np
import pytest
import torch

from neural_lod_amd import configs, synthetic
from neural_lod_amd import tracer as TR
from neural_lod_amd import voxelizer as V
from neural_lod_amd.lod_grid import LodGrid
from oracle import model as M
from oracle import tracer as OT

pytestmark = pytest.mark.gpu

G, N_LODS = 64, 3


@pytest.fixture(scope="module")
def scene(cuda):
    tris = synthetic.blob_mesh(30_000, seed=11)
    table, info = V.voxelize(torch.from_numpy(tris).cuda(), G)
    grid = LodGrid(table, info, N_LODS)
    return grid, info


def _models(precision, seed=3, vis_cfg=configs.VISIBILITY, thr_cfg=configs.THROUGHPUT, grid_range=1.0):
    from neural_lod_amd.tcnn import NetworkWithInputEncoding
    out = []
    for cfg, s in ((vis_cfg, seed), (thr_cfg, seed + 1)):
        m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"],
                                     precision)
        L = M.Layout(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"])
        p = M.xavier_uniform(L, s, grid_range=grid_range)
        m.set_params(torch.from_numpy(p).cuda())
        out.append((m, L, p))
    return out


def _camera(info, W=64, H=48):
    lo, hi = np.array(info.bbox_min), np.array(info.bbox_max)
    c, e = (lo + hi) / 2, hi - lo
    return TR.look_at(c + e * np.array([0.35, 0.25, -1.3]), c, (0, 1, 0), 50.0, W, H)


def _oracle_scene(grid, info, lod, thr=None):
    bits = grid.table(lod).cpu().numpy().view(np.uint32)
    return OT.scene(bits, G >> lod, info.bbox_min, info.bbox_max, info.unit[0], thr)


def _gpu_net(vis, thr):
    def net(which, x):
        m = vis if which == 0 else thr
        return m.inference(torch.from_numpy(np.ascontiguousarray(x)).cuda()).cpu().numpy()
    return net


def _render(tr, cam, params, sample_idx=0):
    illum, inf, bnc = tr.render(cam, params, sample_idx=sample_idx)
    torch.cuda.synchronize()
    return illum.cpu().numpy(), inf.cpu().numpy().view(np.uint32), bnc.cpu().numpy().view(np.uint32)


def _assert_same(a, b, what):
    for x, y, name in zip(a, b, ("illumination", "visibility queries", "final bounce")):
        assert np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                              y.view(np.uint32) if y.dtype == np.float32 else y), f"{what}: {name} differs"


@pytest.mark.parametrize("precision", [0, 1])
def test_reference_schedule_equals_oracle_loop(scene, precision):
    grid, info = scene
    (vis, _, _), (thr, _, _) = _models(precision)
    tr = TR.Tracer(grid, vis, thr)
    cam = _camera(info)
    p = TR.default_params(lod=1, max_depth=4, schedule=TR.REFERENCE)
    got = _render(tr, cam, p)
    ref = OT.trace(_oracle_scene(grid, info, 1), OT.camera_from(cam), OT.params_from(p, 10, G), _gpu_net(vis, thr))
    _assert_same(got, ref, f"precision {precision}")
    illum, inf, bnc = got
    # the scene exercises every path ending
    assert (bnc == 0).any() and (bnc >= 1).any() and (inf >= 2).any()
    assert np.isfinite(illum).all()


OPTIONS = {
    "default": {},
    "hg": {"use_hg": 1, "hg_g": 0.6},
    "roulette": {"apply_rr": 1, "rr_start_bounce": 1, "apply_visibility_rr": 1},
    "stochastic": {"stochastic_threshold": 1},
    "log_learning": {"throughput_log_learning": 1, "max_depth": 5},
    "lod0": {"lod": 0, "max_depth": 2},
    "lod2": {"lod": 2},
}


@pytest.mark.parametrize("name", list(OPTIONS))
def test_persistent_equals_reference_schedule(scene, name):
    grid, info = scene
    (vis, _, _), (thr, _, _) = _models(1)
    tr = TR.Tracer(grid, vis, thr)
    cam = _camera(info, 80, 60)
    kw = dict({"lod": 1, "max_depth": 3}, **OPTIONS[name])
    pr = TR.default_params(schedule=TR.REFERENCE, **kw)
    pp = TR.default_params(schedule=TR.PERSISTENT, **kw)
    assert tr.schedule(TR.default_params(**kw)) == TR.PERSISTENT
    for sample in (0, 5):
        got = _render(tr, cam, pp, sample)
        st = tr.stats()
        _assert_same(got, _render(tr, cam, pr, sample), f"{name}, sample {sample}")
        # the frame's statistics count every visibility query once, in 16-slot tiles
        assert st["visibility_queries"] == int(got[1].astype(np.int64).sum())
        assert 16 * st["visibility_tiles"] >= st["visibility_queries"] >= st["visibility_tiles"] > 0
        assert 16 * st["throughput_tiles"] >= st["throughput_queries"] >= st["throughput_tiles"]


def test_persistent_level_bits_in_global_memory(cuda):
    """A 128^3 level (256 KB of occupancy bits) does not fit in LDS beside the
    weights, so the persistent march reads the bits from global memory; it
    must still equal the reference schedule bit for bit."""
    tris = synthetic.blob_mesh(30_000, seed=11)
    table, info = V.voxelize(torch.from_numpy(tris).cuda(), 128)
    grid = LodGrid(table, info, 2)
    (vis, _, _), (thr, _, _) = _models(1, seed=13)
    tr = TR.Tracer(grid, vis, thr)
    cam = _camera(info, 48, 40)
    kw = {"lod": 0, "max_depth": 2}
    got = _render(tr, cam, TR.default_params(schedule=TR.PERSISTENT, **kw), 2)
    _assert_same(got, _render(tr, cam, TR.default_params(schedule=TR.REFERENCE, **kw), 2), "128^3 level")
    assert (got[1] >= 1).any()


def test_options_equal_oracle(scene):
    """The options that draw extra random numbers or change the arithmetic,
    against the CPU loop (FP16 networks, reference schedule)."""
    grid, info = scene
    (vis, _, _), (thr, _, _) = _models(1, seed=9)
    tr = TR.Tracer(grid, vis, thr)
    cam = _camera(info, 40, 30)
    for name in ("hg", "roulette", "stochastic", "log_learning"):
        p = TR.default_params(lod=1, schedule=TR.REFERENCE, **dict({"max_depth": 3}, **OPTIONS[name]))
        got = _render(tr, cam, p, 3)
        ref = OT.trace(_oracle_scene(grid, info, 1), OT.camera_from(cam), OT.params_from(p, 10, G),
                       _gpu_net(vis, thr), sample_idx=3)
        _assert_same(got, ref, name)


def test_per_voxel_thresholds_equal_oracle(scene):
    """Thresholds of the level's occupied voxels in Morton order
    (nlod_lod_grid_set_thresholds), placed independently here."""
    grid, info = scene
    lod = 1
    R = G >> lod
    n = grid.active_voxels(lod)
    t = np.random.default_rng(4).uniform(0.2, 0.8, n).astype(np.float32)
    grid.set_thresholds(lod, torch.from_numpy(t).cuda(), 0.3, 0.7)
    try:
        bits = grid.table(lod).cpu().numpy().view(np.uint32)
        occ = np.unpackbits(bits.astype(">u4").view(np.uint8))[: R ** 3].reshape(R, R, R)  # [z][y][x]
        z, y, x = np.nonzero(occ)

        def spread(v):
            out = np.zeros_like(v, dtype=np.int64)
            for bit in range(10):
                out |= ((v >> bit) & 1) << (3 * bit)
            return out
        order = np.argsort(spread(x) | (spread(y) << 1) | (spread(z) << 2))
        dense = np.full(R ** 3, 0.5, np.float32)
        dense[(x + y * R + z * R * R)[order]] = np.clip(t, 0.3, 0.7)
        (vis, _, _), (thr, _, _) = _models(1, seed=12)
        tr = TR.Tracer(grid, vis, thr)
        cam = _camera(info, 48, 40)
        p = TR.default_params(lod=lod)
        got = _render(tr, cam, p)
        ref = OT.trace(_oracle_scene(grid, info, lod, dense), OT.camera_from(cam), OT.params_from(p, 10, G),
                       _gpu_net(vis, thr))
        _assert_same(got, ref, "per-voxel thresholds")
    finally:
        grid.set_thresholds(lod, torch.full((n,), 0.5, device="cuda"), 0.5, 0.5)


def test_reference_networks_width_128(scene):
    """The reference's own configurations (RepeatedComposite visibility,
    SH degree 8 throughput, width 128): the reference schedule in FP16
    against the CPU loop."""
    grid, info = scene
    (vis, _, _), (thr, _, _) = _models(1, vis_cfg=configs.REFERENCE_VISIBILITY,
                                       thr_cfg=configs.REFERENCE_THROUGHPUT)
    tr = TR.Tracer(grid, vis, thr)
    cam = _camera(info, 40, 30)
    p = TR.default_params(lod=1)
    assert tr.schedule(p) == TR.REFERENCE
    got = _render(tr, cam, p)
    ref = OT.trace(_oracle_scene(grid, info, 1), OT.camera_from(cam), OT.params_from(p, 9, G), _gpu_net(vis, thr))
    _assert_same(got, ref, "reference networks")


def test_persistent_graph_replays(scene):
    grid, info = scene
    (vis, _, _), (thr, _, _) = _models(1, seed=5)
    tr = TR.Tracer(grid, vis, thr)
    cam = _camera(info, 96, 64)
    p = TR.default_params(lod=1, schedule=TR.PERSISTENT)
    eager = _render(tr, cam, p, 7)
    out = tuple(t.clone() for t in tr.render(cam, p, sample_idx=7))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        tr.render(cam, p, sample_idx=7, out=out)  # warm up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        tr.render(cam, p, sample_idx=7, out=out)
    for rep in range(3):
        for t in out:
            t.fill_(-1)
        g.replay()
        torch.cuda.synchronize()
        _assert_same(tuple(t.cpu().numpy() if t.dtype == torch.float32 else t.cpu().numpy().view(np.uint32)
                           for t in out), eager, f"replay {rep}")
        _ = out[0].sum().item()  # eager work between replays


def test_fp32_against_fp64_network(scene):
    """FP32 networks against the FP64 oracle network: the images agree except
    for paths whose visibility lands within rounding of the threshold."""
    grid, info = scene
    (vis, Lv, pv), (thr, Lt, pt) = _models(0, seed=21)
    tr = TR.Tracer(grid, vis, thr)
    cam = _camera(info, 48, 36)
    p = TR.default_params(lod=1)
    got = _render(tr, cam, p)

    def net64(which, x):
        L, prm = (Lv, pv) if which == 0 else (Lt, pt)
        return M.forward(L, prm, x.astype(np.float64))

    ref = OT.trace(_oracle_scene(grid, info, 1), OT.camera_from(cam), OT.params_from(p, 10, G), net64)
    same = (got[1] == ref[1]) & (got[2] == ref[2])
    assert same.mean() >= 0.98
    np.testing.assert_allclose(got[0][same], ref[0][same], rtol=1e-4, atol=1e-6)
