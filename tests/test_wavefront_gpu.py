"""GPU tests of the wavefront engine (visibility, then throughput where visible).

The fused FP16 launch must reproduce the split launches bit for bit (same FP16
features, same MFMA order), the FP32 plain form must match the FP64 oracle at
1e-5 relative, graph replays must equal eager runs, and the autotuner's choice
must give the same results."""
import numpy as np
import pytest
import torch

from neural_lod_amd import configs, synthetic
from oracle import model as M

pytestmark = pytest.mark.gpu

THRESHOLD = 0.5


def _pair(precision, seed=3):
    from neural_lod_amd.tcnn import NetworkWithInputEncoding, Wavefront
    nets = []
    for cfg, s in ((configs.VISIBILITY, seed), (configs.THROUGHPUT, seed + 1)):
        m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"],
                                     precision)
        L = M.Layout(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"])
        p = M.xavier_uniform(L, s, grid_range=1.0)
        pt = torch.from_numpy(p).cuda()
        m.set_params(pt)
        nets.append((m, L, p, pt))
    return Wavefront(nets[0][0], nets[1][0], THRESHOLD), nets


def _inputs(n, seed):
    v = synthetic.visibility_queries(n, seed)
    t = synthetic.throughput_queries(n, seed=seed + 100)
    return torch.from_numpy(v).cuda(), torch.from_numpy(t).cuda()


@pytest.mark.parametrize("n", [1, 17, 1000, 40000])
def test_fused_matches_split_bitwise(cuda, n):
    wf, _ = _pair(1)
    assert wf.max_shape() >= 1
    vi, ti = _inputs(n, 5)
    vs, ts = wf.run(vi, ti, shape=0)
    for shape in range(1, wf.max_shape() + 1):
        vf, tf = wf.run(vi, ti, shape=shape)
        assert torch.equal(vf, vs), f"visibility differs at shape {shape}"
        assert torch.equal(tf, ts), f"throughput differs at shape {shape}"
    vis = vs[:, 0]
    assert ((vis >= THRESHOLD) | (ts == 0).all(dim=1)).all(), "throughput must be 0 where not visible"


def test_fp16_vs_oracle(cuda):
    wf, nets = _pair(1)
    n = 4096
    vi, ti = _inputs(n, 6)
    vo, to = wf.run(vi, ti)
    (_, Lv, pv, _), (_, Lt, pt_, _) = nets
    ref_v = M.forward(Lv, pv, vi.cpu().numpy())[:, :1]
    ref_t = M.forward(Lt, pt_, ti.cpu().numpy())[:, :3]
    assert np.abs(vo.cpu().numpy() - ref_v).max() <= 2e-3
    visible = (vo.cpu().numpy()[:, 0] >= THRESHOLD) & (ref_v[:, 0] >= THRESHOLD)
    got = to.cpu().numpy()[visible]
    # exponential outputs: 2e-3 absolute on the pre-activation, i.e. relative on the output
    assert np.all(np.abs(np.log(got) - np.log(ref_t[visible])) <= 2e-3)


def test_fp32_plain_vs_oracle(cuda):
    wf, nets = _pair(0)
    assert wf.max_shape() == 0
    n = 3000
    vi, ti = _inputs(n, 7)
    vo, to = wf.run(vi, ti)
    (_, Lv, pv, _), (_, Lt, pt_, _) = nets
    ref_v = M.forward(Lv, pv, vi.cpu().numpy())[:, :1]
    ref_t = M.forward(Lt, pt_, ti.cpu().numpy())[:, :3]
    np.testing.assert_allclose(vo.cpu().numpy(), ref_v, rtol=1e-5, atol=1e-7)
    keep = (ref_v[:, 0] >= THRESHOLD) & (vo.cpu().numpy()[:, 0] >= THRESHOLD)
    np.testing.assert_allclose(to.cpu().numpy()[keep], ref_t[keep], rtol=1e-5)
    assert (to.cpu().numpy()[~keep & (vo.cpu().numpy()[:, 0] < THRESHOLD)] == 0).all()


@pytest.mark.parametrize("precision", [0, 1])
def test_batches_graph_replay(cuda, precision):
    """Uneven wavefronts back to back; the graph replay equals eager runs,
    several replays with eager host work in between."""
    wf, _ = _pair(precision)
    rng = np.random.default_rng(8)
    counts = [int(c) for c in np.exp2(rng.uniform(6, 14, 7)).astype(np.int64)]
    n = sum(counts)
    vi, ti = _inputs(n, 9)
    ve, te = wf.run_batches(counts, vi, ti, graph=False)
    out = (torch.full_like(ve, -1.0), torch.full_like(te, -1.0))
    for rep in range(3):
        out[0].fill_(-1.0)
        out[1].fill_(-1.0)
        vg, tg = wf.run_batches(counts, vi, ti, out=out, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(vg, ve) and torch.equal(tg, te), f"replay {rep} differs"
        _ = vi.sum().item()  # eager work between replays


@pytest.mark.parametrize("precision", [0, 1])
def test_batches_graph_after_rebind_and_growth(cuda, precision):
    """A captured sequence holds the models' weight and scratch addresses.
    After new parameters are bound to one model, and after a larger
    inference call moves the other's scratch buffers, a graph call must give
    what an eager run with the new parameters gives."""
    wf, nets = _pair(precision)
    rng = np.random.default_rng(12)
    counts = [int(c) for c in np.exp2(rng.uniform(6, 13, 5)).astype(np.int64)]
    vi, ti = _inputs(sum(counts), 13)
    v0, t0 = wf.run_batches(counts, vi, ti, graph=False)
    out = (torch.empty_like(v0), torch.empty_like(t0))
    for _ in range(2):  # capture, then replay
        wf.run_batches(counts, vi, ti, out=out, graph=True)
    torch.cuda.synchronize()
    (mv, _, _, _), (mt, Lt, pt_, _) = nets
    p2 = torch.from_numpy(M.xavier_uniform(Lt, 77, grid_range=1.0)).cuda()
    mt.set_params(p2)
    big = torch.from_numpy(synthetic.visibility_queries(1 << 17, 14)).cuda()
    _ = mv.inference(big)
    ve, te = wf.run_batches(counts, vi, ti, graph=False)
    assert not torch.equal(te, t0), "the new throughput weights must change the outputs"
    for rep in range(3):
        out[0].fill_(-1.0)
        out[1].fill_(-1.0)
        vg, tg = wf.run_batches(counts, vi, ti, out=out, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(vg, ve) and torch.equal(tg, te), f"call {rep} after the rebind differs"


def test_autotuner(cuda):
    wf, _ = _pair(1)
    vi, ti = _inputs(1 << 16, 10)
    ref = wf.run(vi, ti, shape=0)
    wf.tune(1 << 16)
    for n in (300, 5000, 1 << 16):
        c = wf.choice(n)
        assert 0 <= c <= wf.max_shape()
    got = wf.run(vi, ti)
    assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])


@pytest.mark.parametrize("precision", [0, 1])
def test_per_segment_thresholds(cuda, precision):
    """One threshold per segment (each segment's voxel threshold): the throughput
    is kept where visibility >= that segment's threshold and 0 elsewhere, bit for
    bit equal to the unthresholded throughput; constant thresholds equal the
    engine's own value; every launch shape agrees."""
    from neural_lod_amd.tcnn import Wavefront
    wf, nets = _pair(precision)
    n = 5003
    print(f"case per-segment thresholds precision={precision} n={n}", flush=True)
    vi, ti = _inputs(n, 8)
    open_wf = Wavefront(nets[0][0], nets[1][0], -1.0)  # every segment visible
    v_all, t_all = open_wf.run(vi, ti)
    g = torch.Generator().manual_seed(9)
    thr = torch.rand(n, generator=g).cuda()
    thr[::7] = v_all[::7, 0]  # ties: visibility == threshold counts as visible
    shapes = [None, 0] + (list(range(1, wf.max_shape() + 1)) if precision == 1 else [])
    for shape in shapes:
        vo, to = wf.run(vi, ti, shape=shape, thresholds=thr)
        assert torch.equal(vo, v_all), f"visibility differs at shape {shape}"
        keep = (vo[:, 0] >= thr)[:, None]
        assert torch.equal(to, torch.where(keep, t_all, torch.zeros_like(t_all))), f"shape {shape}"
        assert 0.05 < keep.float().mean().item() < 0.95
        vc, tc = wf.run(vi, ti, shape=shape, thresholds=torch.full((n,), THRESHOLD, device="cuda"))
        vs, ts = wf.run(vi, ti, shape=shape)
        assert torch.equal(vc, vs) and torch.equal(tc, ts), f"constant thresholds differ at shape {shape}"


@pytest.mark.parametrize("precision", [0, 1])
def test_batches_per_segment_thresholds(cuda, precision):
    """Uneven wavefronts with per-segment thresholds: each wavefront equals a
    single-wavefront run with its slice of the thresholds, eagerly and over
    several graph replays."""
    wf, _ = _pair(precision)
    rng = np.random.default_rng(12)
    counts = [int(c) for c in np.exp2(rng.uniform(6, 13, 6)).astype(np.int64)]
    n = sum(counts)
    print(f"case batches thresholds precision={precision} n={n}", flush=True)
    vi, ti = _inputs(n, 13)
    thr = torch.from_numpy(rng.uniform(0.0, 1.0, n).astype(np.float32)).cuda()
    parts, off = [], 0
    for c in counts:
        parts.append(wf.run(vi[off:off + c].contiguous(), ti[off:off + c].contiguous(),
                            thresholds=thr[off:off + c].contiguous()))
        off += c
    ve, te = torch.cat([p[0] for p in parts]), torch.cat([p[1] for p in parts])
    v0, t0 = wf.run_batches(counts, vi, ti, graph=False, thresholds=thr)
    assert torch.equal(v0, ve) and torch.equal(t0, te), "eager batches differ from single runs"
    out = (torch.full_like(ve, -1.0), torch.full_like(te, -1.0))
    for rep in range(3):
        out[0].fill_(-1.0)
        out[1].fill_(-1.0)
        vg, tg = wf.run_batches(counts, vi, ti, out=out, graph=True, thresholds=thr)
        torch.cuda.synchronize()
        assert torch.equal(vg, ve) and torch.equal(tg, te), f"replay {rep} differs"
        _ = vi.sum().item()
