"""Benchmark of the reconstruction hot path on one MI355X.

One step = tlr_build (reprojection, bucketing, leaves, hierarchy) + tlr_reconstruct
(all output samples, per-pixel combine) over a synthetic sample buffer that is
already resident in HBM.  Prints one JSON line.

Workload (default, --config 4): 1920x1080, 16 input spp (33,177,600 records),
16 output samples per pixel, one held-out seeded scene of the 960x540 scene
type (50 quads, defocus C0=-15.24 C1=19.05, perspective motion): the plain
path is `value`; the learned gate (parameters from gate_params.f32, trained by
`python3 bench.py --gate-train`) is timed on the same scene beside it, with its
error against the plain output.  --config 2 selects 960x540 (16 out-spp),
--config 3 1920x1080 with 400 quads + 5000 occluders at 32 out-spp, --config 1
the 128x128 4-spp case, --config 5 the 64-buffer batch.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__ as ge  # noqa: E402

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    2: dict(preset=2, W=960, H=540, spp=16, out_spp=16,
            name="960x540 16spp in, 16 out-spp, 50 quads, defocus+perspective motion"),
    3: dict(preset=3, W=1920, H=1080, spp=16, out_spp=32,
            name="1920x1080 16spp in, 32 out-spp, 400 quads + 5000 occluders, strong defocus, mixed motion"),
    1: dict(preset=1, W=128, H=128, spp=4, out_spp=16,
            name="128x128 4spp in, 16 out-spp, 3 quads, defocus only"),
    4: dict(preset=2, W=1920, H=1080, spp=16, out_spp=16, seed=300000,
            name="1920x1080 16spp in, 16 out-spp, one held-out seeded scene of the 960x540 scene type "
                 "(50 quads, defocus C0=-15.24 C1=19.05, perspective motion); plain path, learned gate beside it"),
    5: dict(preset=2, W=0, H=0, spp=0, out_spp=16,
            name="64 buffers, resolution uniform in {128^2, 256^2, 512^2, 1024^2, 960x540, 1920x1080}, "
                 "input spp uniform in {1,4,8,16,32}, scene type of the 960x540 workload, 16 out-spp"),
}

BATCH_RES = [(128, 128), (256, 256), (512, 512), (1024, 1024), (960, 540), (1920, 1080)]
BATCH_SPP = [1, 4, 8, 16, 32]


def algorithmic_bytes(st, n_in, W, H):
    """Bytes the reference algorithm must move (oracle Stats counters):
    inputs 48 B/sample, tree points written+read 2x40 B/sample, leaves 2x64 B,
    per query 64 B per leaf visited + 28 B per sample tested + 12 B per sample
    fetched, output 16 B/pixel."""
    b_in = 48.0 * n_in
    b_build = 80.0 * st["num_samples"] + 128.0 * st["num_leaves"]
    b_query = 64.0 * st["leaf_nodes"] + 28.0 * st["samples_in_leaf_nodes"] + 12.0 * st["samples_fetched"]
    b_out = 16.0 * W * H
    return b_in, b_build, b_query, b_out


def read_traffic(kernel, workload_key):
    path = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (kernel, workload_key))
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def cpu_baseline(cfg, sbuf, sample_rows, method="plain"):
    orc = ge.load_oracle()
    W, H = cfg["W"], cfg["H"]
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except Exception:
        pass
    o = orc.Oracle(W, H, cfg["spp"], sbuf.coc[0], sbuf.coc[1], sbuf.records)
    t0 = time.perf_counter()
    o.build(threads=threads)
    t_build = time.perf_counter() - t0
    stride = max(1, H // sample_rows) if sample_rows > 0 else 1
    t0 = time.perf_counter()
    if method == "refgpu":
        o.reconstruct_refgpu(out_spp=cfg["out_spp"], threads=threads, row_stride=stride)
    else:
        o.reconstruct(out_spp=cfg["out_spp"], threads=threads, row_stride=stride)
    t_filter = time.perf_counter() - t0
    rows = (H + stride - 1) // stride
    est = t_build + t_filter * (H / rows)
    o.close()
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": W * H / est / 1e6, "unit": "Mpix/s", "cores": threads, "kind": "port",
            "cpu_model": model, "build_s": t_build, "filter_s_per_frame": t_filter * (H / rows),
            "sample": "full build (%.2f s) + filter of every %d-th scanline (%d of %d rows, %.2f s); "
                      "frame time %.2f s" % (t_build, stride, rows, H, t_filter, est)}


def copy_calibration(mib):
    """Library kernels that move a known number of bytes, for scaling
    rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes: a streaming copy and
    gathers of 32-byte and 64-byte records (the query kernels' sample record
    and hierarchy node) in a scattered order (tlr_calibrate_traffic)."""
    tlr = ge.load_package()
    ctx = tlr.Ctx(0)
    for kind, name in ((0, "k_calib_stream"), (1, "k_calib_gather<2>"), (2, "k_calib_gather<4>")):
        r, w = ctx.calibrate_traffic(kind, mib << 20)
        print(json.dumps({"kernel": name, "bytes_read": r, "bytes_written": w}), flush=True)


GATE_PARAMS = os.path.join(ROOT, "temporal-lightfield-reconstruction_amd", "gate_params.f32")
GATE_META = os.path.join(ROOT, "temporal-lightfield-reconstruction_amd", "gate_params.json")


def load_gate():
    """Committed gate parameters (5,825 little-endian float32) and their
    threshold, or None."""
    if not (os.path.exists(GATE_PARAMS) and os.path.exists(GATE_META)):
        return None
    params = np.fromfile(GATE_PARAMS, dtype="<f4")
    with open(GATE_META) as f:
        meta = json.load(f)
    if params.size != 5825:
        return None
    return params, meta


def bench_batch(args):
    """64 uneven buffers: one batched call (builds + one captured query phase)
    against the same buffers built and reconstructed one at a time.  Inputs
    are uploaded to device slots before any timing."""
    import concurrent.futures as cf
    tlr = ge.load_package()
    cfg = CONFIGS[5]
    rng = np.random.default_rng(args.seed)
    shapes = [(BATCH_RES[rng.integers(len(BATCH_RES))], BATCH_SPP[rng.integers(len(BATCH_SPP))])
              for _ in range(64)]
    ctx = tlr.Ctx(0)
    t0 = time.perf_counter()

    def gen(k):
        (W, H), spp = shapes[k]
        return tlr.make_scene(cfg["preset"], W, H, spp, seed=args.seed * 1000 + k)

    with cf.ThreadPoolExecutor(8) as ex:
        futs = {ex.submit(gen, k): k for k in range(64)}
        for fu in cf.as_completed(futs):
            k = futs[fu]
            ctx.batch_set_samples(k, fu.result())
    print("generated and uploaded 64 buffers in %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    M = cfg["out_spp"]
    px = [W * H for (W, H), _ in shapes]
    total_px = float(sum(px))
    recs = [W * H * spp for (W, H), spp in shapes]
    # sequential plain path, device time per buffer from the library's events
    seq_dev = []
    t0 = time.perf_counter()
    for k in range(64):
        ctx.use_slot(k)
        ctx.build()
        _, st = ctx.reconstruct(out_spp=M)
        seq_dev.append(st["build_ms"] + st["query_ms"] + st["combine_ms"])
    seq_wall = time.perf_counter() - t0
    print("sequential: %.1f ms wall, %.1f ms device" % (seq_wall * 1e3, sum(seq_dev)), file=sys.stderr, flush=True)
    jobs = [(k, M) for k in range(64)]
    outs = [np.empty((H, W, 4), np.float32) for (W, H), _ in shapes]
    for i in range(args.warmup):
        _, _, bst = ctx.reconstruct_batch(jobs, outs)
        print("batch warmup %d: %s" % (i, bst), file=sys.stderr, flush=True)
    walls, sts, dones = [], [], []
    for i in range(args.steps):
        done = []
        t0 = time.perf_counter()
        _, _, bst = ctx.reconstruct_batch(jobs, outs, done=done)
        walls.append(time.perf_counter() - t0)
        sts.append(bst)
        dones.append(done)
    dt = float(np.mean(walls))
    bdev = float(np.mean([b["total_ms"] for b in sts]))
    seq_cum = np.cumsum(seq_dev)
    done_ms = np.mean(np.array(dones), axis=0)  # per buffer: batch start -> its image complete
    plain_imgs = [o.copy() for o in outs]
    # the same batch with the learned gate on every buffer (committed parameters)
    gated = None
    g = load_gate()
    if g is not None:
        params, meta = g
        ctx.gate_set_params(params)
        ctx.gate_set_threshold(meta["threshold"])
        for i in range(args.warmup):
            ctx.reconstruct_batch(jobs, outs, gate=True)
        gw, gs, gd = [], [], []
        for i in range(args.steps):
            done = []
            t0 = time.perf_counter()
            _, _, bst = ctx.reconstruct_batch(jobs, outs, gate=True, done=done)
            gw.append(time.perf_counter() - t0)
            gs.append(bst)
            gd.append(done)
        gdt = float(np.mean(gw))
        err = float(sum(np.abs(o[..., :3] - p[..., :3]).astype(np.float64).sum() for o, p in zip(outs, plain_imgs)) /
                    (3.0 * total_px))
        gdone = np.mean(np.array(gd), axis=0)
        gated = {"mpix_s": total_px / gdt / 1e6, "ms_per_step": gdt * 1e3,
                 "device_ms": {"batch_build": float(np.mean([b["build_ms"] for b in gs])),
                               "batch_query": float(np.mean([b["query_ms"] for b in gs])),
                               "batch_total": float(np.mean([b["total_ms"] for b in gs]))},
                 "speedup_vs_batch_plain_wall": dt / gdt,
                 "speedup_vs_sequential_plain_wall": seq_wall / gdt,
                 "mean_abs_vs_plain": err,
                 "gate_accepted_fraction_of_hand_offs": float(gs[-1]["gate_accepted"] / max(gs[-1]["general_queries"], 1)),
                 "latency_ms_p50": float(np.percentile(gdone, 50)), "latency_ms_p99": float(np.percentile(gdone, 99)),
                 "chained": gs[-1]["chained"], "params": os.path.relpath(GATE_PARAMS, ROOT),
                 "threshold": meta["threshold"]}
    res = {
        "metric": "reconstructed Mpix/s @16 out-spp, 1 GPU, batched (64 uneven buffers)",
        "value": total_px / dt / 1e6,
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (64 seeded scenes generated on the host, uploaded to device slots before timing)",
        "config": {"workload": cfg["name"], "buffers": 64, "pixels": int(total_px), "records": int(sum(recs)),
                   "out_spp": M, "seed": args.seed},
        "device_ms": {"batch_build": float(np.mean([b["build_ms"] for b in sts])),
                      "batch_query": float(np.mean([b["query_ms"] for b in sts])), "batch_total": bdev,
                      "sequential_total": float(sum(seq_dev))},
        "sequential": {"wall_ms": seq_wall * 1e3, "mpix_per_s": total_px / seq_wall / 1e6,
                       "latency_ms_p50": float(np.percentile(seq_cum, 50)),
                       "latency_ms_p99": float(np.percentile(seq_cum, 99))},
        "batch_latency_ms_p50": float(np.percentile(done_ms, 50)),
        "batch_latency_ms_p99": float(np.percentile(done_ms, 99)),
        "batch_latency_note": "per buffer: device wall clock from the batch's first build to that buffer's image",
        "gated_batch": gated,
        "speedup_vs_sequential_wall": seq_wall / dt,
        "speedup_vs_sequential_device": float(sum(seq_dev)) / bdev,
        "tuned": {k: sts[-1][k] for k in ("chained", "fast_leaf_cap", "fast_block", "general_blocks")},
        "graph_reused": sts[-1]["graph_reused"],
    }
    print(json.dumps(res))


def bench_gate(args):
    """--gate-train N: trains the gate on N seeded 960x540 16-spp scenes with
    the library's kernels, calibrates its threshold for a 1e-3 mean-abs budget
    on a validation scene (half of it kept in reserve), evaluates plain
    against gated on --gate-eval held-out 1920x1080 16-spp scenes, and writes
    the parameters to gate_params.f32 / gate_params.json."""
    import importlib
    tlr = ge.load_package()
    gt = importlib.import_module(tlr.__name__ + ".gate_train")
    log = lambda m: print(m, file=sys.stderr, flush=True)
    ctx = tlr.Ctx(0)
    budget = 1e-3
    if args.gate_eval_only:
        # the committed parameters and threshold, evaluated only
        params = np.fromfile(GATE_PARAMS, dtype="<f4")
        with open(GATE_META) as f:
            meta0 = json.load(f)
        thr, losses, t_train = meta0["threshold"], meta0["losses"], None
        ctx.gate_set_params(params)
    else:
        t0 = time.perf_counter()
        train = [(2, 960, 540, 16, 100000 + k) for k in range(args.gate_train)]
        params, losses = gt.train_gate(tlr, ctx, train, out_spp=16, stride=args.gate_stride,
                                       epochs=args.gate_epochs, batch=4096, lr=2e-3, log=log)
        t_train = time.perf_counter() - t0
        log("trained on %d scenes in %.1f s, losses %s" % (args.gate_train, t_train, losses))
        val = tlr.make_scene(2, 960, 540, 16, seed=200000)
        thr, curve = gt.calibrate_threshold(tlr, ctx, val, budget)
        log("threshold %.4f; validation curve (thr, accepted hand-offs, mean-abs): %s" % (
            thr, [(round(a, 3), round(b, 3), "%.2e" % c) for a, b, c in curve[-6:]]))
    ctx.gate_set_threshold(thr)
    cfg = CONFIGS[4]
    W, H, M = cfg["W"], cfg["H"], cfg["out_spp"]
    rows = []
    for k, sb in enumerate(gt._scenes(tlr, [(2, W, H, 16, 300000 + k) for k in range(args.gate_eval)],
                                      workers=4)):
        ctx.set_samples(sb)
        ctx.build()
        for _ in range(max(args.warmup, 1)):
            ctx.reconstruct(out_spp=M)
            ctx.reconstruct(out_spp=M, gate=True)
        tp, tg, sgs = [], [], []
        for _ in range(args.steps):
            plain, sp = ctx.reconstruct(out_spp=M)
            gated, sg = ctx.reconstruct(out_spp=M, gate=True)
            tp.append(sp["query_ms"] + sp["combine_ms"])
            tg.append(sg["query_ms"] + sg["combine_ms"])
            sgs.append(sg)
        err = np.abs(gated - plain)[..., :3]
        rows.append(dict(plain_ms=float(np.mean(tp)), gated_ms=float(np.mean(tg)), mean_abs=float(err.mean()),
                         max_abs=float(err.max()), build_ms=sp["build_ms"],
                         handoffs=sg["gate_candidates"], rejected=sg["gate_rejected"],
                         gate_ms=float(np.mean([x["gate_ms"] for x in sgs])),
                         general_ms=float(np.mean([x["general_ms"] for x in sgs])),
                         plain_general_ms=sp["general_ms"]))
        log("eval %d: %s" % (k, {a: round(b, 5) for a, b in rows[-1].items()}))
    mean = lambda k: float(np.mean([r[k] for r in rows]))
    gated_ms = mean("gated_ms") + mean("build_ms")
    plain_ms = mean("plain_ms") + mean("build_ms")
    if not args.gate_eval_only:
        cmd = "python3 bench.py --gate-train %d --gate-stride %d --gate-epochs %d --gate-eval %d --steps %d" % (
            args.gate_train, args.gate_stride, args.gate_epochs, args.gate_eval, args.steps)
        params.astype("<f4").tofile(GATE_PARAMS)
        meta = {"threshold": thr, "command": cmd, "train_scenes": args.gate_train, "stride": args.gate_stride,
                "epochs": args.gate_epochs, "losses": [float(x) for x in losses], "budget_mean_abs": budget,
                "validation_curve_tail": curve[-8:],
                "label": "provisional first-surface colour within 0.001 of the plain colour in every channel"}
        with open(GATE_META, "w") as f:
            json.dump(meta, f, indent=1)
    res = {
        "metric": "reconstructed Mpix/s @16 out-spp, 1 GPU; learned-gate vs plain at <=1e-3 err",
        "value": W * H / (gated_ms * 1e-3) / 1e6,
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gated_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (gate MLP: bf16 operands, f32 accumulation)",
        "data": "synthetic (seeded scenes generated on the host, uploaded before timing)",
        "config": {"workload": "gate %s, evaluated on %d held-out 1920x1080 16spp scenes at 16 out-spp" % (
                       "parameters and threshold as committed (gate_params.f32 / .json)" if args.gate_eval_only else
                       "trained on %d seeded 960x540 16spp scenes" % args.gate_train, args.gate_eval),
                   "width": W, "height": H, "input_spp": 16, "out_spp": M},
        "plain_mpix_per_s": W * H / (plain_ms * 1e-3) / 1e6,
        "speedup_vs_plain": plain_ms / gated_ms,
        "device_ms_per_frame": {"build": mean("build_ms"), "plain_query": mean("plain_ms"),
                                "gated_query": mean("gated_ms"), "gated_mlp_and_route": mean("gate_ms"),
                                "gated_general": mean("general_ms"), "plain_general": mean("plain_general_ms")},
        "gate": {"threshold": thr, "hand_offs_per_frame": mean("handoffs"),
                 "accepted_fraction_of_hand_offs": 1.0 - mean("rejected") / max(mean("handoffs"), 1.0),
                 "mean_abs_vs_plain": mean("mean_abs"), "worst_scene_mean_abs": max(r["mean_abs"] for r in rows),
                 "max_abs_vs_plain": max(r["max_abs"] for r in rows), "budget": budget,
                 "training_s": t_train, "final_loss": float(losses[-1])},
        "command": "python3 bench.py " + " ".join(sys.argv[1:]),
    }
    print(json.dumps(res))


def bench_refgpu(args):
    """--path 2: the reference's GPU-variant method (recKernel, failKernel,
    nearestKernel) on the same workload; step = build + reconstruct."""
    cfg = CONFIGS[args.config]
    tlr = ge.load_package()
    W, H, spp, M = cfg["W"], cfg["H"], cfg["spp"], cfg["out_spp"]
    print("generating scene %d (%dx%d, %d spp)" % (args.config, W, H, spp), file=sys.stderr, flush=True)
    sbuf = tlr.make_scene(cfg["preset"], W, H, spp, seed=args.seed)
    ctx = tlr.Ctx(0)
    ctx.set_samples(sbuf)
    img = np.empty((H, W, 4), np.float32)

    def step():
        ctx.build()
        return ctx.reconstruct(out_spp=M, path=2, out=img, device_only=True)[1]

    for i in range(args.warmup):
        st = step()
        print("warmup %d: build %.2f ms query %.2f ms" % (i, st["build_ms"], st["query_ms"]),
              file=sys.stderr, flush=True)
    tlr.lib().tlr_sync(ctx.h)
    sts = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        sts.append(step())
    tlr.lib().tlr_sync(ctx.h)
    dt = (time.perf_counter() - t0) / max(args.steps, 1)
    print("timed: %.2f ms/step" % (dt * 1e3), file=sys.stderr, flush=True)
    stages = ctx.query_stages(M)
    mean = lambda k: float(np.mean([s[k] for s in sts]))
    # counters from one more (untimed) call with statistics
    ctx.build()
    st = ctx.reconstruct(out_spp=M, path=2, collect_stats=True, out=img)[1]
    first_ms = mean("fast_ms")
    b_first = 64.0 * st["leaf_nodes"] + 28.0 * st["samples_in_leaf_nodes"] + 12.0 * st["samples_within_1r"]
    achieved = b_first / (first_ms * 1e-3) / 1e9
    res = {
        "metric": "reconstructed Mpix/s @%d out-spp, 1 GPU (reference GPU-variant method)" % M,
        "value": W * H / dt / 1e6,
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded quad scene, generated on the host, uploaded before timing)",
        "config": {"workload": cfg["name"], "width": W, "height": H, "input_spp": spp, "out_spp": M,
                   "records": W * H * spp, "seed": args.seed, "method": "path 2"},
        "device_ms": {"build": mean("build_ms"), "query": mean("query_ms"), "query_1r": mean("fast_ms"),
                      "query_2r_nearest": mean("general_ms"), "combine": mean("combine_ms")},
        "stage_fractions": {"1r": float(np.mean(stages == 0)), "2r": float(np.mean(stages == 1)),
                            "nearest": float(np.mean(stages == 2))},
        "roofline": {"bound": "hbm", "kernel": "k_ref_first", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic("k_ref_first", "cfg%d" % args.config),
                     "algorithmic_bytes_per_launch": b_first, "avg_launch_ms": first_ms},
    }
    if not args.no_cpu:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline(cfg, sbuf, args.cpu_rows, method="refgpu")
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res))


def bench_single(args, cfg):
    """One buffer: step = build + reconstruct, image left in HBM."""
    tlr = ge.load_package()
    W, H, spp, M = cfg["W"], cfg["H"], cfg["spp"], cfg["out_spp"]
    seed = args.seed
    log = lambda m: print(m, file=sys.stderr, flush=True)
    log("generating %dx%d, %d spp (preset %d, seed %d)" % (W, H, spp, cfg["preset"], seed))
    sbuf = tlr.make_scene(cfg["preset"], W, H, spp, seed=seed)
    n_in = W * H * spp
    ctx = tlr.Ctx(0)
    # host -> device upload (AoS -> SoA transpose on the host + copies), timed
    # apart: the timed steps start with the inputs resident in HBM
    t0 = time.perf_counter()
    ctx.set_samples(sbuf)
    upload_s = time.perf_counter() - t0
    img = np.empty((H, W, 4), np.float32)

    def step(stats=False, to_host=False, gate=False):
        ctx.build()
        return ctx.reconstruct(out_spp=M, collect_stats=stats, gate=gate, out=img, device_only=not to_host)[1]

    def timed(gate=False):
        for i in range(args.warmup):
            st = step(gate=gate)
            log("warmup %d%s: build %.2f ms query %.2f ms" % (i, " (gate)" if gate else "", st["build_ms"],
                                                             st["query_ms"]))
        tlr.lib().tlr_sync(ctx.h)
        sts = []
        t0 = time.perf_counter()
        for i in range(args.steps):
            sts.append(step(gate=gate))
        tlr.lib().tlr_sync(ctx.h)
        return (time.perf_counter() - t0) / max(args.steps, 1), sts

    dt, sts = timed()
    log("timed: %.2f ms/step" % (dt * 1e3))
    # the same step with the image copied to host memory (PCIe included)
    t0 = time.perf_counter()
    for i in range(3):
        step(to_host=True)
    tlr.lib().tlr_sync(ctx.h)
    dt_host = (time.perf_counter() - t0) / 3
    # counters come from one more (untimed) pass with statistics enabled;
    # they do not depend on timing
    st = step(stats=True)
    mean = lambda k, ss=None: float(np.mean([x[k] for x in (ss or sts)]))
    q_ms, b_ms = mean("query_ms"), mean("build_ms")
    # fast_ms spans both 1R lane passes; wide_ms is the second's part
    wide_ms = mean("wide_ms")
    fast_ms, gen_ms = mean("fast_ms") - wide_ms, mean("general_ms")
    b_in, b_build, b_query, b_out = algorithmic_bytes(st, n_in, W, H)
    b_fast = (64.0 * st["fast_leaf_nodes"] + 28.0 * st["fast_samples_in_leaf_nodes"] +
              12.0 * st["fast_samples_within_1r"])
    b_wide = (64.0 * st["wide_leaf_nodes"] + 28.0 * st["wide_samples_in_leaf_nodes"] +
              12.0 * st["wide_samples_within_1r"])
    b_gen = b_query - b_fast - b_wide
    wk = "cfg%d" % args.config
    achieved = b_fast / (fast_ms * 1e-3) / 1e9
    traffic = read_traffic("k_query_fast", wk)
    res = {
        "metric": "reconstructed Mpix/s @%d out-spp, 1 GPU (plain path)" % M,
        "value": W * H / dt / 1e6,
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded quad scene, generated on the host, uploaded before timing)",
        "config": {"workload": cfg["name"], "width": W, "height": H, "input_spp": spp, "out_spp": M,
                   "records": n_in, "seed": seed},
        "output_samples_per_s": W * H * M / dt,
        "upload": {"ms": upload_s * 1e3, "device_bytes": 48 * n_in,
                   "note": "host AoS->SoA transpose + pageable copies, once per buffer, outside the timed steps",
                   "mpix_s_including_upload_each_step": W * H / (dt + upload_s) / 1e6},
        "with_image_to_host": {"ms_per_step": dt_host * 1e3, "mpix_s": W * H / dt_host / 1e6},
        "device_ms": {"build": b_ms, "query": q_ms, "query_fast": fast_ms, "query_wide": wide_ms,
                      "query_general": gen_ms, "combine": mean("combine_ms")},
        "wide_pass_entries": st["wide_queries"],
        "general_queries": st["general_queries"],
        "stats_per_output_sample": {k: st[k] / st["count"] for k in (
            "leaf_nodes", "samples_in_leaf_nodes", "samples_within_1r", "samples_fetched", "surfaces",
            "num_2r", "surfaces_merged", "at_least_one")},
        "roofline": {"bound": "hbm", "kernel": "k_query_fast", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_gbs": traffic / (fast_ms * 1e-3) / 1e9 if traffic else None,
                     "algorithmic_bytes_per_launch": b_fast, "avg_launch_ms": fast_ms},
        "roofline_wide": {"kernel": "k_query_fast_wide", "algorithmic_bytes_per_launch": b_wide,
                          "avg_launch_ms": wide_ms, "achieved": b_wide / max(wide_ms * 1e-3, 1e-9) / 1e9},
        "roofline_general": {"kernel": "k_query_general", "algorithmic_bytes_per_launch": b_gen,
                             "avg_launch_ms": gen_ms,
                             "achieved": b_gen / max(gen_ms * 1e-3, 1e-9) / 1e9},
        "path_roofline": {"algorithmic_bytes": b_in + b_build + b_query + b_out,
                          "roofline_ms": (b_in + b_build + b_query + b_out) / (HBM_PEAK_GBS * 1e9) * 1e3,
                          "frac": (b_in + b_build + b_query + b_out) / (HBM_PEAK_GBS * 1e9) / dt},
    }
    if args.config == 4:
        res["gate"] = gate_leg(args, ctx, sbuf, step, timed, dt, log)
    if not args.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(cfg, sbuf, args.cpu_rows)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res))


def gate_leg(args, ctx, sbuf, step, timed, dt_plain, log):
    """The learned gate on the same buffer: committed parameters, timed steps,
    error of the gated image against the plain one."""
    g = load_gate()
    if g is None:
        return {"status": "no committed parameters (python3 bench.py --gate-train)"}
    params, meta = g
    ctx.gate_set_params(params)
    ctx.gate_set_threshold(meta["threshold"])
    W, H = sbuf.width, sbuf.height
    dt, sts = timed(gate=True)
    log("gated: %.2f ms/step" % (dt * 1e3))
    M = CONFIGS[args.config]["out_spp"]
    ctx.build()
    plain, _ = ctx.reconstruct(out_spp=M)
    gated, sg = ctx.reconstruct(out_spp=M, gate=True)
    err = np.abs(gated - plain)[..., :3]
    mean = lambda k: float(np.mean([x[k] for x in sts]))
    return {"mpix_s": W * H / dt / 1e6, "ms_per_step": dt * 1e3, "speedup_vs_plain": dt_plain / dt,
            "mean_abs_vs_plain": float(err.mean()), "max_abs_vs_plain": float(err.max()),
            "budget_mean_abs": 1e-3, "threshold": meta["threshold"],
            "hand_offs": sg["gate_candidates"],
            "accepted_fraction_of_hand_offs": 1.0 - sg["gate_rejected"] / max(sg["gate_candidates"], 1.0),
            "device_ms": {"build": mean("build_ms"), "query": mean("query_ms"), "fast": mean("fast_ms"),
                          "gate": mean("gate_ms"), "general": mean("general_ms")},
            "params": os.path.relpath(GATE_PARAMS, ROOT), "trained_by": meta.get("command")}


def bench_shadows(args):
    """reconstructDofMotionShadows (camera pass with extended results,
    shadow rays, light pass in shadow mode, shading) on the workload's scene
    with a second seeded scene as the light's buffer; the light is the camera
    turned 3 degrees about the y axis.  Both builds are outside the timed
    region (the reference builds both TreeGathers before the call)."""
    tlr = ge.load_package()
    cfg = CONFIGS[args.config]
    W, H, spp, M = cfg["W"], cfg["H"], cfg["spp"], cfg["out_spp"]
    cam_sb = tlr.make_scene(cfg["preset"], W, H, spp, seed=args.seed)
    light_sb = tlr.make_scene(cfg["preset"], W, H, spp, seed=args.seed + 1)
    cam, light = tlr.Ctx(0), tlr.Ctx(0)
    for c, sb in ((cam, cam_sb), (light, light_sb)):
        c.set_samples(sb)
        c.set_shadow_inputs(sb.records[:, 14:16], None)
        c.build()
    f = 1.0 / np.tan(np.radians(60) / 2)
    P = np.array([[f * H / W, 0, 0, 0], [0, f, 0, 0], [0, 0, -100.1 / 99.9, -20.0 / 99.9], [0, 0, -1, 0]])
    a = np.radians(3.0)
    W2L = np.eye(4)
    W2L[:3, :3] = [[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]]
    sp = tlr.shadow_setup(P, np.eye(4), P, W2L, W, H, cam_sb.coc, light_sb.coc, 100.0)
    for _ in range(args.warmup):
        cam.reconstruct_dof_motion_shadows(light, sp, out_spp=M)
    walls, sts = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        _, st = cam.reconstruct_dof_motion_shadows(light, sp, out_spp=M)
        walls.append(time.perf_counter() - t0)
        sts.append(st)
    dt = float(np.mean(walls))
    dev = {k: float(np.mean([s[k] for s in sts])) for k in ("general_ms", "fast_ms", "combine_ms", "query_ms")}
    res = {
        "metric": "reconstructed Mpix/s @%d out-spp with shadows, 1 GPU" % M,
        "value": W * H / dt / 1e6,
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (two seeded scenes: camera and light buffers)",
        "config": {"workload": cfg["name"] + "; light buffer seed + 1, light turned 3 degrees", "width": W,
                   "height": H, "input_spp": spp, "out_spp": M, "seed": args.seed},
        "device_ms": {"camera_pass": dev["general_ms"], "light_pass": dev["fast_ms"], "shading": dev["combine_ms"],
                      "total": dev["query_ms"] + dev["combine_ms"]},
        "light_queries": float(np.mean([s["light_queries"] for s in sts])),
        "output_samples": W * H * M,
        "lane_pass_hand_offs": float(np.mean([s["wide_queries"] for s in sts])),
        "general_queries": float(np.mean([s["general_queries"] for s in sts])),
    }
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=4, choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=None, help="scene seed (default: the workload's)")
    ap.add_argument("--cpu-rows", type=int, default=-1,
                    help="scanlines in the CPU baseline sample (0: the whole frame; default: a quarter at "
                         "1920x1080, the whole frame otherwise)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gate-train", type=int, default=0, metavar="SCENES",
                    help="train the gate on SCENES 960x540 scenes, calibrate it, evaluate it and write "
                         "gate_params.f32 / gate_params.json")
    ap.add_argument("--gate-eval", type=int, default=20, help="--gate-train: held-out evaluation scenes")
    ap.add_argument("--gate-eval-only", action="store_true",
                    help="evaluate the committed gate parameters on --gate-eval held-out scenes (no training)")
    ap.add_argument("--gate-stride", type=int, default=2, help="--gate-train: keep every n-th hand-off")
    ap.add_argument("--gate-epochs", type=int, default=4, help="--gate-train: epochs")
    ap.add_argument("--path", type=int, default=0, choices=(0, 2),
                    help="2: the reference's GPU-variant method instead of the plain one")
    ap.add_argument("--shadows", action="store_true",
                    help="reconstructDofMotionShadows on the workload: a second seeded scene as the light's "
                         "buffer, the light turned 3 degrees about the camera")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the counter-calibration kernels over MIB MiB")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.seed is None:
        args.seed = CONFIGS[args.config].get("seed", 2011)
    if args.cpu_rows < 0:
        args.cpu_rows = CONFIGS[args.config]["H"] // 4 if CONFIGS[args.config]["H"] >= 1080 else 0
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gate_train or args.gate_eval_only:
        bench_gate(args)
        return
    if args.config == 5:
        bench_batch(args)
        return
    if args.path == 2:
        bench_refgpu(args)
        return
    if args.shadows:
        bench_shadows(args)
        return
    bench_single(args, CONFIGS[args.config])


if __name__ == "__main__":
    main()
