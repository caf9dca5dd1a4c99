"""Benchmark of the neural-LoD hot path on one MI355X.

Headline (the JSON line's `value`): queries/s of the FP16-MFMA fused
encode + MLP inference of the visibility network (16-level hash grid, 2 FP32
features/level, 2^19 entries/level, base 16, growth 1.5, + degree-4 SH of the
direction; MLP 64 wide, 4 hidden ReLU layers, sigmoid output) on a batch of
2^20 synthetic queries resident in HBM.  One step = one inference of the
whole batch.  The same run also measures the plain FP32 GPU form, the error
of both against the FP64 CPU oracle, and a CPU baseline (the oracle's
straightforward FP32 C loop, OpenMP over queries) on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

N_QUERIES = 1 << 20
TRAIN_BATCH = 1 << 18
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# algorithmic bytes per query of the FP16 encoding pass (the dominant kernel):
# 16 levels x 8 corners x 2 features x 2 B gathered (FP16 table copy),
# 6 x 4 B input, 48 FP16 features written
ENC_BYTES_PER_QUERY = 16 * 8 * 2 * 2 + 6 * 4 + 48 * 2
# whole step (encode + MLP): + features read back + 1 x 4 B output
BYTES_PER_QUERY = ENC_BYTES_PER_QUERY + 48 * 2 + 4
FLOP_PER_QUERY = 2 * (48 * 64 + 3 * 64 * 64 + 64 * 1)
# HBM bytes of one grid backward at 2^18 queries of the visibility network,
# from FETCH_SIZE and WRITE_SIZE passes calibrated by a 256 MiB copy
# (bench.py --profile-grid-backward; 345.9 MB fetched + 358.5 MB written: the
# 16-byte edge records are written by the bin pass and read by the reduce)
GRID_BWD_TRAFFIC_BYTES = 704.4e6
# HBM bytes of one encoding pass (its two launches) at 2^20 queries, from
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes calibrated by a 256 MiB copy
# (bench.py --profile-encode; 420.5 MB fetched + 150.3 MB written)
ENC_TRAFFIC_BYTES = 570.8e6


# L2 view of the encoding pass (its binding roof: the level rows keep the L2
# channels busy 96% of the pass, 93% of requests hit): 60.2 L2 read requests
# per query (TCP_TCC_READ_REQ_sum over both launches / queries, profiled at 2^20
# queries), against MI355X's 8 XCDs x 16 L2 channels x one 128-B request per
# channel per clock at 2.4 GHz
ENC_L2_REQ_PER_QUERY = 60.2
L2_PEAK_REQ_PER_S = 8 * 16 * 2.4e9
# training step of the visibility network (config 2), per query: the forward
# MLP FLOPs and twice them for backward (dL/dx and dW); 16 levels x 8 corners x
# 2 features x 4 B gathered and the same scattered back as FP32 gradients; and
# per parameter per step Adam's 28 B (gradient, weight, two moments, step count
# read and written, FP16 copy)
TRAIN_FLOP_PER_QUERY = 3 * FLOP_PER_QUERY
TRAIN_GATHER_BYTES_PER_QUERY = 16 * 8 * 2 * 4
TRAIN_SCATTER_BYTES_PER_QUERY = 16 * 8 * 2 * 4
ADAM_BYTES_PER_PARAM = 28
MFMA_FP16_PEAK_TFLOPS = 2500.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def running_mean(v, w):
    c = np.cumsum(np.insert(np.asarray(v, np.float64), 0, 0.0))
    return (c[w:] - c[:-w]) / w


def curve_gap(c32, c16, window=50, tail=100):
    """Loss-curve agreement of two runs at equal steps: the gap of their means
    over the last `tail` steps and the largest gap of their `window`-step
    running means, in percent of the FP32 curve."""
    r32, r16 = running_mean(c32, window), running_mean(c16, window)
    gaps = np.abs(r16 - r32) / np.abs(r32) * 100.0
    m32, m16 = float(np.mean(c32[-tail:])), float(np.mean(c16[-tail:]))
    return {"fp32_mean_last": m32, "fp16_mean_last": m16, "last_steps": tail,
            "loss_gap_pct": abs(m16 - m32) / abs(m32) * 100.0,
            "max_running_gap_pct": float(gaps.max()), "running_window": window,
            "argmax_running_gap_step": int(np.argmax(gaps)) + window,
            "fp32_first": float(c32[0]), "fp16_first": float(c16[0])}


WAVEFRONT_LODS = 5
WAVEFRONT_BOUNCES = 4


def wavefront_bench(dev, args):
    """Uneven wavefronts: 5 LoD levels x 4 bounces, active rays per wavefront
    log-uniform in [2^10, 2^20] (seeded); per segment the visibility network,
    then the throughput network where visibility >= 0.5.  FP16: tuned fused
    launches, the 20-wavefront sequence replayed from a hipGraph.  FP32 plain:
    separate launches per network and wavefront, no graph."""
    import torch
    from neural_lod_amd import configs, synthetic
    from neural_lod_amd.tcnn import FP16, FP32, NetworkWithInputEncoding, Wavefront
    from oracle import model as M
    rng = np.random.default_rng(4)
    counts = [int(c) for c in np.exp2(rng.uniform(10, 20, WAVEFRONT_LODS * WAVEFRONT_BOUNCES))]
    total = sum(counts)
    vis_in = torch.from_numpy(synthetic.visibility_queries(total, 41)).to(dev)
    thr_q = synthetic.throughput_queries(total, seed=42)
    # LoD input of each wavefront: its level index / 4 (levels 0..4, 4 bounces each)
    off = 0
    for i, c in enumerate(counts):
        thr_q[off:off + c, 9] = (i // WAVEFRONT_BOUNCES) / 4.0
        off += c
    thr_in = torch.from_numpy(thr_q).to(dev)
    res = {"wavefronts": len(counts), "counts": counts, "segments": total, "threshold": 0.5}
    for name, prec in (("fp32", FP32), ("fp16", FP16)):
        nets = []
        for cfg, seed in ((configs.VISIBILITY, 5), (configs.THROUGHPUT, 6)):
            m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], prec)
            L = M.Layout(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"])
            p = torch.from_numpy(M.xavier_uniform(L, seed, grid_range=1.0)).to(dev)
            m.set_params(p)
            nets.append((m, p))
        wf = Wavefront(nets[0][0], nets[1][0], 0.5)
        if prec == FP16:
            wf.tune(max(counts))
        vo = torch.empty((total, 1), device=dev)
        to = torch.empty((total, 3), device=dev)
        use_graph = prec == FP16
        for _ in range(2):  # warm (and capture)
            wf.run_batches(counts, vis_in, thr_in, out=(vo, to), graph=use_graph)
        torch.cuda.synchronize()
        reps = max(3, args.steps // 5)
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            wf.run_batches(counts, vis_in, thr_in, out=(vo, to), graph=use_graph)
        e1.record(stream)
        torch.cuda.synchronize()
        seq_ms = e0.elapsed_time(e1) / reps
        # per-wavefront latency: each wavefront alone, median of 5
        lat = []
        off = 0
        for c in counts:
            sl = slice(off, off + c)
            v, t = vis_in[sl], thr_in[sl]
            ms = []
            for _ in range(5):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                wf.run(v, t, out=(vo[sl], to[sl]))
                b.record(stream)
                torch.cuda.synchronize()
                ms.append(a.elapsed_time(b))
            lat.append(float(np.median(ms)))
            off += c
        res[name] = {"queries_per_s": total / (seq_ms * 1e-3), "sequence_ms": seq_ms,
                     "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
                     "graph": use_graph}
        if prec == FP16:
            res[name]["shapes"] = [wf.choice(c) for c in counts]
        log(f"wavefront {name}: {seq_ms:.3f} ms / sequence, {total / seq_ms / 1e6:.2f} Mq/ms..., "
            f"p50 {res[name]['latency_ms_p50']:.3f} ms p99 {res[name]['latency_ms_p99']:.3f} ms")
        del wf, nets
    res["speedup_fp16_vs_fp32"] = res["fp32"]["sequence_ms"] / res["fp16"]["sequence_ms"]
    return res


LABEL_RES = 64          # LoD grid of the visibility training labels
VOX_RES = 256
VOX_TRIS = 200_000       # the CPU-runnable voxelization workload (blob mesh, 256^3 solid)
VOX_TRIS_LARGE = 20_000_000
VOX_TRIS_SWEEP = ((2_000_000, "2M"), (VOX_TRIS_LARGE, "20M"))  # larger meshes of the same construction


THR_RES = 256     # finest level of the throughput-target pyramid (256^3 .. 16^3)
THR_LODS = 5
THR_SPP = 64


THR_POOL = 16  # batches of Monte-Carlo targets the throughput network cycles through
THR_SEEDS = (1337, 1338, 1339, 1340)  # initialisations of the matched-loss runs


def throughput_train_bench(dev, args):
    """Prefiltered-throughput network (hash grid + SH4(wo) + SH4(wi) + LoD
    scalar, MLP 64 x 5 ReLU, 3 exp outputs), RelativeL2, batch 2^18 of valid
    Monte-Carlo targets (64 paths each) of the 200k-triangle blob mesh over a
    5-level pyramid (256^3 .. 16^3), a pool of 16 batches cycled: FP16-MFMA and
    FP32 plain training steps, timed, then both trained from each of four
    initialisations for the same steps, every step's loss read back and the
    loss of the EMA weights over the whole pool taken at checkpoints over the
    second half."""
    import torch
    from neural_lod_amd import configs, synthetic
    from neural_lod_amd.labels import Labeler
    from neural_lod_amd.tcnn import FP16, FP32, NetworkWithInputEncoding, Trainer, loss_evaluate
    cfg = configs.THROUGHPUT
    nb = TRAIN_BATCH
    mesh = torch.from_numpy(synthetic.blob_mesh(VOX_TRIS, seed=0)).to(dev)
    lb = Labeler(mesh, THR_RES, THR_LODS)
    gen = 1 << 18
    xs, ts, made, first = [], [], 0, 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while made < THR_POOL * nb:
        x, t, v = lb.throughput(gen, seed=200, first=first, spp=THR_SPP, max_depth=3)
        keep = v > 0
        xs.append(x[keep])
        ts.append(t[keep])
        made += int(keep.sum())
        first += gen
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    x_all, t_all = torch.cat(xs)[:THR_POOL * nb].contiguous(), torch.cat(ts)[:THR_POOL * nb].contiguous()
    del xs, ts
    data = [(x_all[k * nb:(k + 1) * nb], t_all[k * nb:(k + 1) * nb]) for k in range(THR_POOL)]
    out = {"batch": nb, "spp": THR_SPP, "lods": THR_LODS, "pool_batches": THR_POOL,
           "targets": f"Monte-Carlo voxel-local throughput, {THR_SPP} paths, 200k-triangle blob mesh, "
                      f"{THR_RES}^3 .. {THR_RES >> (THR_LODS - 1)}^3 pyramid, valid samples only",
           "target_gen_ms_per_1k_samples": gen_s * 1e3 / (first / 1024), "valid_fraction": made / first}
    # Adam wrapped in an EMA of the weights (decay 0.95), the structure of the
    # reference's throughput optimizer (Ema over ExponentialDecay over Adam), and
    # the EMA weights evaluated.  lr 5e-4: at 1e-2 the exp outputs fall to 0,
    # where RelativeL2 has no gradient, and at 1e-3 single steps spike the loss
    # 2-3x, so that the runs' losses at equal steps scatter by tens of percent
    adam = {"otype": "Ema", "decay": 0.95,
            "nested": {"otype": "Adam", "learning_rate": 5e-4, "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-15,
                       "l2_reg": 1e-8}}
    out["optimizer"] = adam
    def pool_loss(m):
        """RelativeL2 of the network over the whole pool (its own forward form)."""
        tot = 0.0
        for xb, tb in data:
            values, _ = loss_evaluate(1, m.inference(xb), tb)
            tot += float(values.double().sum())
        return tot / len(data)

    checkpoints = [args.loss_steps - k * args.loss_steps // 8 for k in (4, 3, 2, 1, 0)]
    curves = {"fp32": [], "fp16": []}
    pool = {"fp32": [], "fp16": []}
    for name, prec in (("fp32", FP32), ("fp16", FP16)):
        m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], prec)
        tr = Trainer(m, adam, {"otype": "RelativeL2"}, seed=1337, graphs=True)
        for k in range(max(args.warmup, 2)):
            tr.training_step(*data[k % THR_POOL])
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(args.train_steps):
            tr.training_step(*data[k % THR_POOL])
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.train_steps
        out[name] = {"ms_per_step": ms, "steps_per_s": 1e3 / ms}
        log(f"throughput train {name}: {ms:.4f} ms/step")
        del tr
        # loss curves: fresh trainers from THR_SEEDS' initialisations; the
        # loss over the whole pool at the last checkpoints of each
        for seed in THR_SEEDS:
            tr = Trainer(m, adam, {"otype": "RelativeL2"}, seed=seed, graphs=True)
            losses = np.empty(args.loss_steps)
            at = []
            for k in range(args.loss_steps):
                tr.training_step(*data[k % THR_POOL])
                losses[k] = tr.status()["loss"]
                if k + 1 in checkpoints:
                    at.append(pool_loss(m))
            curves[name].append(losses)
            pool[name].append(float(np.mean(at)))
            log(f"throughput {name} seed {seed}: loss {losses[0]:.5f} -> {np.mean(losses[-100:]):.5f} over "
                f"{args.loss_steps} steps, pool loss {pool[name][-1]:.5f}")
            del tr
        out[name]["final_loss"] = float(curves[name][0][-1])
        del m
    out["speedup_fp16_vs_fp32"] = out["fp32"]["ms_per_step"] / out["fp16"]["ms_per_step"]
    p32, p16 = float(np.mean(pool["fp32"])), float(np.mean(pool["fp16"]))
    out["matched_loss"] = {
        "statistic": f"loss of the EMA weights over all {THR_POOL} pool batches after steps {checkpoints} "
                     f"(their mean), averaged over the initialisations {list(THR_SEEDS)}; the per-step losses "
                     f"cycle through the pool's batches, whose fit differs, so single curves scatter",
        "fp32_pool_loss": p32, "fp16_pool_loss": p16,
        "loss_gap_pct": abs(p16 - p32) / p32 * 100.0,
        "fp32_per_seed": pool["fp32"], "fp16_per_seed": pool["fp16"],
        "fp32_seed_spread_pct": (max(pool["fp32"]) - min(pool["fp32"])) / p32 * 100.0,
        "curves_per_seed": [curve_gap(c32, c16) for c32, c16 in zip(curves["fp32"], curves["fp16"])],
        "steps": args.loss_steps, "data": f"{THR_POOL} batches of 2^18 valid targets, cycled"}
    log(f"throughput matched loss: {out['matched_loss']}")
    del lb, mesh, data, x_all, t_all
    torch.cuda.empty_cache()
    return out


TRACE_RES, TRACE_LODS, TRACE_LOD = 256, 5, 2   # 256^3 voxelization, pyramid to 16^3, traced at 64^3
TRACE_W, TRACE_H = 1024, 1024                   # 2^20 camera paths per frame
TRACE_DEPTH = 3


def tracer_scene(dev):
    """The 200k-triangle blob mesh voxelized at 256^3, its LoD pyramid and a
    camera outside the box looking at its centre."""
    import torch
    from neural_lod_amd import synthetic, voxelizer as V
    from neural_lod_amd import tracer as TR
    from neural_lod_amd.lod_grid import LodGrid
    tris = torch.from_numpy(synthetic.blob_mesh(VOX_TRIS, seed=0)).to(dev)
    table, info = V.voxelize(tris, TRACE_RES)
    grid = LodGrid(table, info, TRACE_LODS)
    lo, hi = np.array(info.bbox_min), np.array(info.bbox_max)
    c, e = (lo + hi) / 2, hi - lo
    cam = TR.look_at(c + e * np.array([0.35, 0.25, -1.3]), c, (0, 1, 0), 45.0, TRACE_W, TRACE_H)
    return grid, cam


def tracer_bench(dev, args):
    """The path tracer's frame (voxel_tracing): 2^20 camera paths through the
    64^3 level of the blob mesh's pyramid, up to 3 bounces, the visibility
    network (hash grid + SH4, 64 x 4) per voxel segment and the throughput
    network (64 x 5) per bounce.  FP16: the persistent schedule (one launch,
    device ray queue) replayed from a hipGraph, and the reference's schedule
    (compaction rounds with a host count readback each).  FP32 plain: the
    reference's schedule.  Random-initialised networks (grid tables U[-1, 1])."""
    import torch
    from neural_lod_amd import configs
    from neural_lod_amd import tracer as TR
    from neural_lod_amd.tcnn import FP16, FP32, NetworkWithInputEncoding
    from oracle import model as M
    grid, cam = tracer_scene(dev)
    n = TRACE_W * TRACE_H
    res = {"paths": n, "lod_res": TRACE_RES >> TRACE_LOD, "max_depth": TRACE_DEPTH,
           "scene": f"200k-triangle blob mesh, {TRACE_RES}^3 voxelization, level {TRACE_LOD}"}
    images = {}
    for prec_name, prec in (("fp16", FP16), ("fp32", FP32)):
        nets = []
        for cfg_, seed in ((configs.VISIBILITY, 5), (configs.THROUGHPUT, 6)):
            m = NetworkWithInputEncoding(cfg_["n_input_dims"], cfg_["n_output_dims"], cfg_["encoding"],
                                         cfg_["network"], prec)
            L_ = M.Layout(cfg_["n_input_dims"], cfg_["n_output_dims"], cfg_["encoding"], cfg_["network"])
            p_ = torch.from_numpy(M.xavier_uniform(L_, seed, grid_range=1.0)).to(dev)
            m.set_params(p_)
            nets.append((m, p_))
        tr = TR.Tracer(grid, nets[0][0], nets[1][0])
        schedules = (("persistent", TR.PERSISTENT), ("reference", TR.REFERENCE)) if prec == FP16 else \
            (("reference", TR.REFERENCE),)
        for sname, sched in schedules:
            p = TR.default_params(lod=TRACE_LOD, max_depth=TRACE_DEPTH, schedule=sched)
            out = tr.render(cam, p)
            torch.cuda.synchronize()
            reps = max(3, args.steps // 4) if sched == TR.PERSISTENT else 2
            graph = None
            if sched == TR.PERSISTENT:
                s_ = torch.cuda.Stream()
                s_.wait_stream(torch.cuda.current_stream())
                with torch.cuda.stream(s_):
                    tr.render(cam, p, out=out)
                torch.cuda.current_stream().wait_stream(s_)
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph):
                    tr.render(cam, p, out=out)
                graph.replay()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                if graph is not None:
                    graph.replay()
                else:
                    tr.render(cam, p, out=out)
            e1.record()
            torch.cuda.synchronize()
            wall_ms = (time.perf_counter() - t0) * 1e3 / reps
            ms = e0.elapsed_time(e1) / reps
            illum, inf, bnc = (t.clone() for t in out)
            key = f"{prec_name}_{sname}"
            images[key] = (illum, inf, bnc)
            vis_q = int(inf.sum().item())
            thr_q = int(bnc.sum().item())  # one throughput query per bounce taken (no roulette, exp outputs > 0)
            res[key] = {"ms_per_frame": ms, "wall_ms_per_frame": wall_ms, "paths_per_s": n / (ms * 1e-3),
                        "network_queries_per_s": (vis_q + thr_q) / (ms * 1e-3),
                        "visibility_queries": vis_q, "throughput_queries": thr_q,
                        "graph": graph is not None}
            if sched == TR.PERSISTENT:
                res[key]["network_evaluations"] = tr.stats()
            log(f"tracer {key}: {ms:.3f} ms/frame ({wall_ms:.3f} wall), {n / ms / 1e3:.2f} Mpaths/s, "
                f"{vis_q} vis + {thr_q} thr queries")
            del graph
        if prec == FP16:
            a, b = images["fp16_persistent"], images["fp16_reference"]
            res["fp16_persistent_equals_reference"] = all(torch.equal(u, v) for u, v in zip(a, b))
            inf, bnc = a[1].float(), a[2].float()
            res["path_stats"] = {"mean_visibility_queries": float(inf.mean()),
                                 "max_visibility_queries": int(inf.max()),
                                 "bounce_histogram": [int((a[2] == k).sum()) for k in range(TRACE_DEPTH)]}
        del tr, nets
    res["speedup_fp16_persistent_vs_fp32_reference"] = res["fp32_reference"]["ms_per_frame"] / \
        res["fp16_persistent"]["ms_per_frame"]
    res["speedup_fp16_persistent_vs_fp16_reference"] = res["fp16_reference"]["ms_per_frame"] / \
        res["fp16_persistent"]["ms_per_frame"]
    log(f"tracer: {res['speedup_fp16_persistent_vs_fp32_reference']:.2f}x vs FP32 reference schedule, "
        f"{res['speedup_fp16_persistent_vs_fp16_reference']:.2f}x vs FP16 reference schedule, "
        f"identical images: {res.get('fp16_persistent_equals_reference')}")
    del grid
    torch.cuda.empty_cache()
    return res


def voxelize_bench(dev, args):
    """Solid and surface voxelization of the seeded 200k-triangle blob mesh at
    256^3 (bit-exact check against the CPU restatement, which is also timed as
    the CPU baseline), plus 2M- and 20M-triangle meshes of the same
    construction for the rate at scale.  Bytes per mesh: 36 B per triangle read + the table
    written once."""
    import ctypes as C
    import torch
    from neural_lod_amd import synthetic, voxelizer as V
    out = {"res": VOX_RES, "mesh": "blob mesh (64 Gaussian blobs, sigma 0.05, vertex offsets U[-0.01, 0.01])"}
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def time_vox(tris, solid, reps):
        table, info = V.voxelize(tris, VOX_RES, solid=solid)
        fl = V._flags(solid, False, False)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            V._lib.call("nlod_voxelize", stream, C.byref(info), C.c_void_p(tris.data_ptr()),
                        C.c_void_p(table.data_ptr()), fl)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps, table

    table_bytes = VOX_RES ** 3 // 8
    tris_np = synthetic.blob_mesh(VOX_TRIS, seed=0)
    tris = torch.from_numpy(tris_np).to(dev)
    for mode, solid in (("solid", True), ("surface", False)):
        ms, table = time_vox(tris, solid, max(args.steps, 20))
        nbytes = VOX_TRIS * 36 + table_bytes
        out[mode] = {"us_per_mesh": ms * 1e3, "triangles_per_s": VOX_TRIS / (ms * 1e-3),
                     "achieved_GBs": nbytes / (ms * 1e-3) / 1e9, "frac_hbm_peak": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        gpu_table = table.cpu().numpy().view(np.uint32)
        if not args.no_cpu:
            from oracle import oracle as O
            n_threads = int(O.lib().nlo_omp_max_threads())
            for threads, key in ((1, "cpu_baseline"), (n_threads, "cpu_baseline_omp")):
                t0 = time.perf_counter()
                reps = 0
                while True:
                    cpu_table, _ = O.voxelize(tris_np, VOX_RES, solid=solid, threads=threads)
                    reps += 1
                    if time.perf_counter() - t0 > (4.0 if solid else 2.0):
                        break
                dt = (time.perf_counter() - t0) / reps
                out[mode]["bit_exact_vs_cpu"] = bool(np.array_equal(gpu_table, cpu_table))
                out[mode][key] = {"value": VOX_TRIS / dt, "unit": "triangles/s", "cores": threads, "kind": "port",
                                  "sample": f"{reps} x the same 200k-triangle mesh at {VOX_RES}^3 (C restatement of "
                                            f"cpu_voxelizer.cpp with its omp critical bit writes, {threads} "
                                            f"OpenMP thread{'s' if threads > 1 else ''})"}
        log(f"voxelize {mode} 200k @ {VOX_RES}^3: {ms * 1e3:.1f} us  {out[mode]}")
    del tris
    for n_big, tag in VOX_TRIS_SWEEP:
        big = torch.from_numpy(synthetic.blob_mesh(n_big, seed=0)).to(dev)
        for mode, solid in (("solid", True), ("surface", False)):
            ms, _ = time_vox(big, solid, 3)
            nbytes = n_big * 36 + table_bytes
            out[mode + "_" + tag] = {"ms_per_mesh": ms, "triangles_per_s": n_big / (ms * 1e-3),
                                     "achieved_GBs": nbytes / (ms * 1e-3) / 1e9,
                                     "frac_hbm_peak": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
            log(f"voxelize {mode} {tag} @ {VOX_RES}^3: {ms:.3f} ms")
        del big
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-train", action="store_true", help="skip the training-step measurement")
    ap.add_argument("--train-steps", type=int, default=50)
    ap.add_argument("--loss-steps", type=int, default=2000,
                    help="steps of the matched-loss curves (FP16 vs FP32 from the same initialisation)")
    ap.add_argument("--no-tracer", action="store_true", help="skip the path-tracer measurement")
    ap.add_argument("--profile-grid-backward", action="store_true",
                    help="only run the grid backward of a 2^18 training batch and a 256 MiB copy (counter profiles)")
    ap.add_argument("--profile-tracer", action="store_true",
                    help="only render the path-tracer frames (FP16 persistent) for kernel traces")
    ap.add_argument("--profile-mlp", action="store_true",
                    help="only run the FP16 MLP (mlp_fwd_kernel) on pre-encoded 2^20-query features")
    ap.add_argument("--profile-train", action="store_true",
                    help="only run FP16 training steps of the visibility network on one 2^18 mesh-label batch")
    ap.add_argument("--profile-voxelize", action="store_true",
                    help="only voxelize the 200k- and 20M-triangle blob meshes at 256^3, solid and surface")
    ap.add_argument("--no-wavefront", action="store_true", help="skip the wavefront-batch measurement")
    ap.add_argument("--no-voxelize", action="store_true", help="skip the voxelization measurement")
    ap.add_argument("--profile-encode", action="store_true",
                    help="only run the FP16 encoding pass and a 256 MiB device copy (for counter profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from neural_lod_amd import configs, synthetic
    from neural_lod_amd.tcnn import FP16, FP32, NetworkWithInputEncoding
    from oracle import model as M

    dev = torch.device("cuda:0")
    cfg = configs.VISIBILITY
    L = M.Layout(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"])
    # Xavier MLP; grid table drawn U[-1, 1] so the FP16 error check is not vacuous
    params_np = M.xavier_uniform(L, seed=0, grid_range=1.0)
    x_np = synthetic.visibility_queries(N_QUERIES, seed=1)
    params = torch.from_numpy(params_np).to(dev)
    x = torch.from_numpy(x_np).to(dev)
    log(f"workload: {N_QUERIES} queries, {L.n_params} params ({L.n_params * 4 / 2**20:.1f} MiB FP32)")

    if args.profile_encode:
        # the dominant inference kernel alone, then a plain copy of known size that
        # calibrates the HBM byte counters (FETCH_SIZE / WRITE_SIZE)
        m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], FP16)
        m.set_params(params)
        feat = m.encode_fragments(x)
        for _ in range(args.steps):
            m.encode_fragments(x, feat)
        src = torch.empty(1 << 26, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"profile-encode: {args.steps} encoding passes of {N_QUERIES} queries, {args.steps} copies of 256 MiB")
        return

    if args.profile_grid_backward:
        # the dominant kernel of the FP16 training step alone: grid backward of one
        # 2^18 batch of mesh-label inputs, then a 256 MiB copy for the byte counters
        from neural_lod_amd.labels import VisibilityLabeler
        mesh = torch.from_numpy(synthetic.blob_mesh(VOX_TRIS, seed=0)).to(dev)
        xb, _ = VisibilityLabeler(mesh, LABEL_RES).sample(TRAIN_BATCH, seed=100)
        m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], FP16)
        m.set_params(params)
        d_enc = torch.randn((m.encoded_width, TRAIN_BATCH), device=dev) * 1e-3
        d_par = torch.empty(m.n_params, device=dev)
        for _ in range(args.steps):
            m.encode_backward(xb, d_enc, d_par)
        src = torch.empty(1 << 26, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"profile-grid-backward: {args.steps} grid backwards of {TRAIN_BATCH} queries, {args.steps} copies")
        return

    if args.profile_mlp:
        m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], FP16)
        m.set_params(params)
        feat = m.encode_fragments(x)
        out = torch.empty((N_QUERIES, 1), device=dev)
        for _ in range(args.steps):
            m.mlp_fragments(feat, N_QUERIES, out)
        torch.cuda.synchronize()
        log(f"profile-mlp: {args.steps} MLP passes over {N_QUERIES} queries")
        return

    if args.profile_train:
        from neural_lod_amd.labels import VisibilityLabeler
        from neural_lod_amd.tcnn import Trainer
        mesh = torch.from_numpy(synthetic.blob_mesh(VOX_TRIS, seed=0)).to(dev)
        xb, yb = VisibilityLabeler(mesh, LABEL_RES).sample(TRAIN_BATCH, seed=100)
        m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], FP16)
        adam = {"otype": "Adam", "learning_rate": 1e-2, "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-15,
                "l2_reg": 1e-8}
        tr = Trainer(m, adam, {"otype": "BinaryCrossEntropy"}, seed=1337, graphs=False)
        for _ in range(args.steps):
            tr.training_step(xb, yb)
        torch.cuda.synchronize()
        log(f"profile-train: {args.steps} FP16 training steps of {TRAIN_BATCH} queries (eager launches)")
        return

    if args.profile_voxelize:
        import ctypes as C
        from neural_lod_amd import voxelizer as V
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        for n_tri in (VOX_TRIS, VOX_TRIS_LARGE):
            tris = torch.from_numpy(synthetic.blob_mesh(n_tri, seed=0)).to(dev)
            for solid in (True, False):
                table, info = V.voxelize(tris, VOX_RES, solid=solid)
                for _ in range(args.steps):
                    V._lib.call("nlod_voxelize", stream, C.byref(info), C.c_void_p(tris.data_ptr()),
                                C.c_void_p(table.data_ptr()), V._flags(solid, False, False))
            torch.cuda.synchronize()
            del tris
        log(f"profile-voxelize: {args.steps} + 1 runs each of 200k and 20M triangles, solid then surface")
        return

    if args.profile_tracer:
        from neural_lod_amd import tracer as TR
        grid, cam = tracer_scene(dev)
        nets = []
        for cfg_, seed in ((configs.VISIBILITY, 5), (configs.THROUGHPUT, 6)):
            m = NetworkWithInputEncoding(cfg_["n_input_dims"], cfg_["n_output_dims"], cfg_["encoding"],
                                         cfg_["network"], FP16)
            L_ = M.Layout(cfg_["n_input_dims"], cfg_["n_output_dims"], cfg_["encoding"], cfg_["network"])
            p_ = torch.from_numpy(M.xavier_uniform(L_, seed, grid_range=1.0)).to(dev)
            m.set_params(p_)
            nets.append((m, p_))
        tr = TR.Tracer(grid, nets[0][0], nets[1][0])
        p = TR.default_params(lod=TRACE_LOD, max_depth=TRACE_DEPTH, schedule=TR.PERSISTENT)
        out = tr.render(cam, p)
        for _ in range(args.steps):
            tr.render(cam, p, out=out)
        torch.cuda.synchronize()
        log(f"profile-tracer: {args.steps + 1} persistent frames of {TRACE_W}x{TRACE_H} paths")
        return

    models = {}
    for name, prec in (("fp16", FP16), ("fp32", FP32)):
        m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], prec)
        m.set_params(params)
        models[name] = m
    outs = {k: torch.empty((N_QUERIES, 1), device=dev) for k in models}

    def run(name, k):
        m, o = models[name], outs[name]
        for _ in range(k):
            m.inference(x, o)

    def time_steps(fn, k):
        fn()
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    results = {}
    for name in ("fp32", "fp16"):
        run(name, args.warmup)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        run(name, args.steps)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = e0.elapsed_time(e1) / args.steps
        results[name] = {"ms": ms, "qps": N_QUERIES / (ms * 1e-3), "wall_s": wall}
        log(f"{name}: {ms:.4f} ms/step  {N_QUERIES / ms / 1e6:.2f} Mq/ms...")

    # the dominant kernel of the FP16 step: the per-level encoding pass, timed alone
    feat = models["fp16"].encode_fragments(x)
    enc_ms = time_steps(lambda: models["fp16"].encode_fragments(x, feat), args.steps)
    log(f"fp16 encoding pass alone: {enc_ms:.4f} ms")
    # the fused MLP alone on those features (mlp_fwd_kernel), timed with events
    mlp_out = torch.empty((N_QUERIES, 1), device=dev)
    mlp_ms = time_steps(lambda: models["fp16"].mlp_fragments(feat, N_QUERIES, mlp_out), args.steps)
    torch.cuda.synchronize()
    assert torch.equal(mlp_out, outs["fp16"]), "encode + MLP alone must equal the inference call"
    log(f"fp16 MLP alone: {mlp_ms:.4f} ms ({FLOP_PER_QUERY * N_QUERIES / (mlp_ms * 1e-3) / 1e12:.0f} TFLOP/s)")

    # accuracy vs the FP64 oracle on a sample of the same batch
    n_chk = 1 << 14
    ref = M.forward(L, params_np, x_np[:n_chk])[:, :1]
    err = {k: float(np.abs(outs[k][:n_chk].cpu().numpy() - ref).max()) for k in outs}
    log(f"max abs error vs FP64 oracle ({n_chk} queries): {err}")

    # training step (forward + BCE + backward + FP64 gradient check + Adam), batch 2^18
    train = None
    if not args.no_train:
        from neural_lod_amd.tcnn import Trainer
        adam = {"otype": "Adam", "learning_rate": 1e-2, "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-15,
                "l2_reg": 1e-8}
        from neural_lod_amd.labels import VisibilityLabeler
        nb = TRAIN_BATCH
        # labels: visibility rays from the occupied 64^3 voxels of the 200k-triangle
        # blob mesh (nearest hit inside the voxel), drawn on the GPU
        mesh = torch.from_numpy(synthetic.blob_mesh(VOX_TRIS, seed=0)).to(dev)
        labeler = VisibilityLabeler(mesh, LABEL_RES)
        data = [labeler.sample(nb, seed=100, first=k * nb) for k in range(4)]  # resident in HBM, cycled
        label_ms = time_steps(lambda: labeler.sample(nb, seed=101, out=data[0]), 10)
        data[0] = labeler.sample(nb, seed=100, first=0, out=data[0])
        positive = float(torch.cat([d[1] for d in data]).mean())
        log(f"labels: {labeler.n_occupied} occupied voxels, {positive:.3f} positive, {label_ms:.3f} ms per batch")
        train = {}
        for name, prec in (("fp32", FP32), ("fp16", FP16)):
            m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], prec)
            tr = Trainer(m, adam, {"otype": "BinaryCrossEntropy"}, seed=1337, graphs=True)
            for k in range(max(args.warmup, 2)):
                tr.training_step(*data[k % 4])
            torch.cuda.synchronize()
            stream = torch.cuda.current_stream()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for k in range(args.train_steps):
                tr.training_step(*data[k % 4])
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / args.train_steps
            st = tr.status()
            train[name] = {"ms_per_step": ms, "steps_per_s": 1e3 / ms, "final_loss": st["loss"],
                           "loss_scale": st["loss_scale"], "skipped_steps": st["skipped_steps"]}
            log(f"train {name}: {ms:.4f} ms/step, loss {st['loss']:.4f}, scale {st['loss_scale']}")
            del tr, m
        train["speedup_fp16_vs_fp32"] = train["fp32"]["ms_per_step"] / train["fp16"]["ms_per_step"]
        train["batch"] = nb
        # matched loss at the configuration's scale: both forms from the same
        # initialisation, fresh labels every step (sample k of the seeded stream
        # at step k, identical for both), every step's loss read back
        curves = {}
        for name, prec in (("fp32", FP32), ("fp16", FP16)):
            m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], prec)
            tr = Trainer(m, adam, {"otype": "BinaryCrossEntropy"}, seed=1337, graphs=True)
            buf = labeler.sample(nb, seed=300, first=0)
            losses = np.empty(args.loss_steps)
            for k in range(args.loss_steps):
                labeler.sample(nb, seed=300, first=k * nb, out=buf)
                tr.training_step(*buf)
                losses[k] = tr.status()["loss"]
            curves[name] = losses
            log(f"train {name} loss {losses[0]:.5f} -> {np.mean(losses[-100:]):.5f} over {args.loss_steps} "
                f"fresh-label steps")
            del tr, m
        train["matched_loss"] = dict(curve_gap(curves["fp32"], curves["fp16"]), steps=args.loss_steps,
                                     data="fresh mesh labels every step (2^18 samples of one seeded stream)")
        log(f"train matched loss: {train['matched_loss']}")
        # roofline of the FP16 step: algorithmic bytes and FLOPs per step, and the
        # dominant kernel, the grid backward (binned scatter of dL/d(features)),
        # timed alone on the same batch with random feature gradients
        m16 = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], FP16)
        m16.set_params(params)
        xb = data[0][0]
        d_enc = torch.randn((m16.encoded_width, nb), device=dev) * 1e-3
        d_par = torch.empty(m16.n_params, device=dev)
        gb_ms = time_steps(lambda: m16.encode_backward(xb, d_enc, d_par), max(args.steps, 20))
        del m16, d_enc, d_par
        step_bytes = nb * (TRAIN_GATHER_BYTES_PER_QUERY + TRAIN_SCATTER_BYTES_PER_QUERY) + \
            ADAM_BYTES_PER_PARAM * L.n_params
        step_flop = nb * TRAIN_FLOP_PER_QUERY
        ms16 = train["fp16"]["ms_per_step"]
        gb_bytes = nb * TRAIN_SCATTER_BYTES_PER_QUERY
        train["roofline"] = {
            "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "step_bytes": step_bytes, "step_flop": step_flop,
            "achieved": step_bytes / (ms16 * 1e-3) / 1e9, "frac": step_bytes / (ms16 * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "mfma_tflops": step_flop / (ms16 * 1e-3) / 1e12,
            "mfma_frac": step_flop / (ms16 * 1e-3) / 1e12 / MFMA_FP16_PEAK_TFLOPS,
            "dominant_kernel": "grid backward (GridBackward: bin, scans, reduce, apply)",
            "dominant_ms": gb_ms, "dominant_bytes": gb_bytes,
            "dominant_achieved": gb_bytes / (gb_ms * 1e-3) / 1e9,
            "dominant_frac": gb_bytes / (gb_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "dominant_traffic": GRID_BWD_TRAFFIC_BYTES,
            "dominant_traffic_ratio": None if GRID_BWD_TRAFFIC_BYTES is None else GRID_BWD_TRAFFIC_BYTES / gb_bytes,
            "basis": "per query: 16 levels x 8 corners x 2 features x 4 B gathered and scattered, 3 x 30,848 "
                     "MLP FLOP; per parameter 28 B of Adam",
        }
        log(f"train roofline: {train['roofline']}")
        train["labels"] = (f"mesh visibility rays: {VOX_TRIS}-triangle blob mesh, origin in one of the "
                           f"{labeler.n_occupied} occupied {LABEL_RES}^3 voxels, direction on the sphere, "
                           "1 = nearest hit inside the voxel")
        train["labels_positive_fraction"] = positive
        train["label_gen_ms_per_batch"] = label_ms
        if not args.no_cpu:
            # the oracle's training step (FP64 forward/backward, float32 Adam; OpenMP
            # over queries) on 2^14-query batches of the same workload
            n_c = 1 << 14
            xc = data[0][0][:n_c].cpu().numpy()
            yc = data[0][1][:n_c].cpu().numpy()
            p0 = M.xavier_uniform(L, seed=1337)
            M.train(L, p0, [(xc, yc)], adam, "BinaryCrossEntropy")  # warm threads
            t0 = time.perf_counter()
            reps = 0
            while time.perf_counter() - t0 < 8.0:
                M.train(L, p0, [(xc, yc)] * 2, adam, "BinaryCrossEntropy")
                reps += 2
            dt = (time.perf_counter() - t0) / reps
            train["cpu_baseline"] = {"value": n_c / dt / nb, "unit": f"steps/s at batch {nb}",
                                     "cores": int(M._lib().nlo_omp_threads()), "kind": "port",
                                     "sample": f"{reps} steps of {n_c} queries (FP64 C forward/backward + Adam), "
                                               "scaled to the batch"}
            log(f"train cpu baseline: {train['cpu_baseline']}")

    throughput = None if args.no_train else throughput_train_bench(dev, args)
    wavefront = None if args.no_wavefront else wavefront_bench(dev, args)
    tracer = None if args.no_tracer else tracer_bench(dev, args)
    voxelize = None if args.no_voxelize else voxelize_bench(dev, args)

    # CPU baseline: the oracle's FP32 C loop on a bounded sample
    cpu = None
    if not args.no_cpu:
        threads = M._lib().nlo_omp_threads()
        n_cpu = 1 << 17
        xs = x_np[:n_cpu]
        M.infer_f32(L, params_np, xs[:1024])  # warm caches / threads
        t0 = time.perf_counter()
        reps = 0
        while True:
            M.infer_f32(L, params_np, xs)
            reps += 1
            if time.perf_counter() - t0 > 10.0:
                break
        dt = time.perf_counter() - t0
        cpu = {"value": n_cpu * reps / dt, "unit": "queries/s", "cores": int(threads), "kind": "port",
               "sample": f"{reps} x {n_cpu} queries of the same workload (FP32 C loop, OpenMP over queries)"}
        log(f"cpu baseline: {cpu['value']:.3e} queries/s on {threads} threads")

    fp16 = results["fp16"]
    achieved = ENC_BYTES_PER_QUERY * N_QUERIES / (enc_ms * 1e-3) / 1e9
    line = {
        "metric": "encode+MLP queries/s & train steps/s, FP16-MFMA vs FP32 plain, <=2e-3 err",
        "value": fp16["qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp16["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16-mfma/fp32-acc",
        "data": "synthetic",
        "config": {"workload": "visibility network inference: hash grid 16x2 (2^19/level, base 16, x1.5) + SH4, "
                               "MLP 64x4 ReLU, sigmoid; batch 2^20 queries",
                   "global_batch": N_QUERIES, "parallelism": "single GPU"},
        "plain_fp32": {"value": results["fp32"]["qps"], "ms_per_step": results["fp32"]["ms"],
                       "max_abs_err_vs_fp64": err["fp32"]},
        "fp16_max_abs_err_vs_fp64": err["fp16"],
        "speedup_vs_plain_fp32": results["fp32"]["ms"] / fp16["ms"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": ENC_TRAFFIC_BYTES,
                     "kernel": "FP16 encoding pass: encode_lds_kernel<2> (levels 0-1 from LDS) + encode_kernel<2,4,half> (levels 2-15 and SH)",
                     "kernel_ms": enc_ms, "bytes_per_query": ENC_BYTES_PER_QUERY,
                     # the pass's binding roof is the L2 request rate (tables L2/MALL-resident)
                     "l2": {"requests_per_query": ENC_L2_REQ_PER_QUERY,
                            "achieved_req_per_s": ENC_L2_REQ_PER_QUERY * N_QUERIES / (enc_ms * 1e-3),
                            "peak_req_per_s": L2_PEAK_REQ_PER_S,
                            "frac": ENC_L2_REQ_PER_QUERY * N_QUERIES / (enc_ms * 1e-3) / L2_PEAK_REQ_PER_S},
                     "mlp_kernel": "mlp_fwd_kernel (FP16 MFMA 16x16x32, FP32 accumulate), timed alone",
                     "mlp_kernel_ms": mlp_ms, "mlp_flop_per_query": FLOP_PER_QUERY,
                     "mlp_tflops": FLOP_PER_QUERY * N_QUERIES / (mlp_ms * 1e-3) / 1e12,
                     "mlp_frac_mfma_peak": FLOP_PER_QUERY * N_QUERIES / (mlp_ms * 1e-3) / 1e12 / MFMA_FP16_PEAK_TFLOPS},
        "cpu_baseline": cpu,
        "train": train,
        "throughput_train": throughput,
        "wavefront": wavefront,
        "tracer": tracer,
        "voxelize": voxelize,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
