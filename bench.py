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
# HBM bytes of one encoding pass (its two launches) at 2^20 queries, from
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes calibrated by a 256 MiB copy
# (bench.py --profile-encode; 420.5 MB fetched + 150.3 MB written)
ENC_TRAFFIC_BYTES = 570.8e6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def throughput_train_bench(dev, args):
    """Prefiltered-throughput network (hash grid + SH4(wo) + SH4(wi) + LoD
    scalar, MLP 64 x 5 ReLU, 3 exp outputs), RelativeL2, batch 2^18 of valid
    Monte-Carlo targets (64 paths each) of the 200k-triangle blob mesh over a
    5-level pyramid (256^3 .. 16^3): FP16-MFMA and FP32 plain training steps."""
    import torch
    from neural_lod_amd import configs, synthetic
    from neural_lod_amd.labels import Labeler
    from neural_lod_amd.tcnn import FP16, FP32, NetworkWithInputEncoding, Trainer
    cfg = configs.THROUGHPUT
    nb = TRAIN_BATCH
    mesh = torch.from_numpy(synthetic.blob_mesh(VOX_TRIS, seed=0)).to(dev)
    lb = Labeler(mesh, THR_RES, THR_LODS)
    gen = 1 << 18
    xs, ts, made, first = [], [], 0, 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while made < 2 * nb:
        x, t, v = lb.throughput(gen, seed=200, first=first, spp=THR_SPP, max_depth=3)
        keep = v > 0
        xs.append(x[keep])
        ts.append(t[keep])
        made += int(keep.sum())
        first += gen
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    x_all, t_all = torch.cat(xs)[:2 * nb].contiguous(), torch.cat(ts)[:2 * nb].contiguous()
    data = [(x_all[:nb], t_all[:nb]), (x_all[nb:], t_all[nb:])]
    out = {"batch": nb, "spp": THR_SPP, "lods": THR_LODS,
           "targets": f"Monte-Carlo voxel-local throughput, {THR_SPP} paths, 200k-triangle blob mesh, "
                      f"{THR_RES}^3 .. {THR_RES >> (THR_LODS - 1)}^3 pyramid, valid samples only",
           "target_gen_ms_per_1k_samples": gen_s * 1e3 / (first / 1024), "valid_fraction": made / first}
    # lr 1e-3: at 1e-2 the exp outputs fall to 0, where RelativeL2 has no gradient
    adam = {"otype": "Adam", "learning_rate": 1e-3, "beta1": 0.9, "beta2": 0.99, "epsilon": 1e-15, "l2_reg": 1e-8}
    out["optimizer"] = adam
    for name, prec in (("fp32", FP32), ("fp16", FP16)):
        m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], prec)
        tr = Trainer(m, adam, {"otype": "RelativeL2"}, seed=1337, graphs=True)
        for k in range(max(args.warmup, 2)):
            tr.training_step(*data[k % 2])
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(args.train_steps):
            tr.training_step(*data[k % 2])
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.train_steps
        st = tr.status()
        out[name] = {"ms_per_step": ms, "steps_per_s": 1e3 / ms, "final_loss": st["loss"]}
        log(f"throughput train {name}: {ms:.4f} ms/step, loss {st['loss']:.5f}")
        del tr, m
    out["speedup_fp16_vs_fp32"] = out["fp32"]["ms_per_step"] / out["fp16"]["ms_per_step"]
    del lb, mesh
    torch.cuda.empty_cache()
    return out


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
            t0 = time.perf_counter()
            reps = 0
            while True:
                cpu_table, _ = O.voxelize(tris_np, VOX_RES, solid=solid, threads=1)
                reps += 1
                if time.perf_counter() - t0 > (4.0 if solid else 2.0):
                    break
            dt = (time.perf_counter() - t0) / reps
            out[mode]["bit_exact_vs_cpu"] = bool(np.array_equal(gpu_table, cpu_table))
            out[mode]["cpu_baseline"] = {"value": VOX_TRIS / dt, "unit": "triangles/s", "cores": 1, "kind": "port",
                                         "sample": f"{reps} x the same 200k-triangle mesh at {VOX_RES}^3 "
                                                   "(C restatement of cpu_voxelizer.cpp, one thread)"}
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
                     "mlp_kernel_ms": fp16["ms"] - enc_ms, "mlp_flop_per_query": FLOP_PER_QUERY,
                     "mlp_tflops": FLOP_PER_QUERY * N_QUERIES / ((fp16["ms"] - enc_ms) * 1e-3) / 1e12},
        "cpu_baseline": cpu,
        "train": train,
        "throughput_train": throughput,
        "wavefront": wavefront,
        "voxelize": voxelize,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
