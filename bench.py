"""Benchmark of the 3DGS training step on one MI355X.

Workload: 3,000,000 Gaussians (means U[-3,3]^3, log-scales N(-4,0.5), uniform rotations, opacity
logits N(0,1), SH degree 3 N(0,0.3); seeded), 1920x1080 pinhole views on a radius-8 ring, 16x16
tiles, L1 loss against targets rendered from a seeded ground-truth set, Adam over every parameter.

Two step kinds are timed:
  plain   : FP32 storage, one view per step, kernels launched one by one (the plain FP32 GPU step);
  batched : BF16 storage of SH / log-scales / opacity logits (FP32 blending, FP32 Adam masters),
            V views per step in one launch per stage, the whole step replayed as one hipGraph.
``value`` is the batched step's throughput in view-steps/s (1080p views trained per second:
V x optimizer steps/s), directly comparable with the plain step's steps/s (one view per step).

Prints one JSON line (rank 0). --gpus > 1 is not supported on this build and exits.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "train steps/s at 3M Gaussians 1080p, batched BF16 vs plain FP32; HBM GB/s"
HBM_PEAK_GBPS = 8000.0


def algorithmic_bytes(n, views, visible, instances, pixels, tiles, sh_bytes, lp, sh_quant=False):
    """Per-stage algorithmic HBM bytes of one step (each logical array counted once per kernel
    that must touch it; element sizes of the stored types). The per-Gaussian backward reads and
    writes the 59 parameters and the Adam state: FP32 moments (8 B per value) for all of them, or,
    with the quantised SH state, for the 11 structure values plus 48 x (2 B fp16 update + 1 B code)
    + 16 B bounds for the SH row."""
    s = sh_bytes
    sort_passes_depth = 4
    tile_bits = max(1, int(tiles - 1).bit_length())
    sort_passes_tile = (tile_bits + 7) // 8
    out = {
        "preprocess": n * (12 + (6 if lp else 12) + 16 + (2 if lp else 4)) + visible * (16 * 3 * s)
        + visible * (8 + 16 + 12 + 4) + n * views * (4 + 4),
        "compact+depth_sort": n * views * 4 + visible * 12 + sort_passes_depth * 24 * visible,
        "tile_offsets+emit": visible * 8 + visible * 32 + 12 * instances,
        "tile_sort+ranges": sort_passes_tile * 24 * instances + 8 * instances + 8 * tiles,
        "blend_forward": 40 * instances + 20 * pixels,
        "loss": 36 * pixels,
        "blend_backward": 40 * instances + 24 * pixels + 44 * visible,
        "gaussian_backward+adam": n * (59 * 4 * 2 + (11 * 8 + 48 * 3 + 16 if sh_quant else 59 * 8) * 2
                                       + (52 * 2 if lp else 0)) + 44 * visible,
    }
    return out


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE / WRITE_SIZE summaries under
    profiles/ (already scaled by the calibration copy), or None when no summary covers it."""
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            if not line.startswith("traffic:"):
                continue
            kv = dict(x.split("=", 1) for x in line[len("traffic:"):].split())
            if kv.get("kernel") == kernel and kv.get("workload") == workload:
                return int(kv["fetch_bytes"]) + int(kv["write_bytes"]), f.name
    return None, None


def make_targets(n, cams, opts, device, seed):
    from photara_studio_amd.rasterizer import Rasterizer
    from photara_studio_amd.scenes import model_from, synthetic_torch
    gt = model_from(synthetic_torch(n, 3.0, seed, device), device)
    r = Rasterizer()
    out = []
    for c in cams:
        res = r.forward_views(gt, [c], opts)
        out.append(res[0].color.clone())
        del res
    del gt, r
    torch.cuda.synchronize()
    return out


def timed(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def cpu_baseline(arrays, cams, opts, seconds):
    """The CPU oracle's training step (forward + L1 + backward + Adam) on the host cores, over a
    bounded sample: whole steps of one 1080p view of the same scene until ~`seconds` elapse."""
    import ctypes as C
    import numpy as np
    from oracle import pyoracle as po
    from photara_studio_amd import _abi
    from photara_studio_amd.rasterizer import make_camera
    threads = int(os.environ.get("PHT_CPU_THREADS", "0")) or min(16, os.cpu_count() or 1)
    scene = po.Scene(arrays)
    lib = po.load()
    cam = cams[0]
    g = scene.gaussians()
    s = po.settings_for(scene, opts)
    cams_c = (_abi.pht_camera * 1)(make_camera(cam, opts))
    color = np.zeros((3, cam.height, cam.width), np.float32)
    alpha = np.zeros((cam.height, cam.width), np.float32)
    outs = (_abi.pht_view_outputs * 1)()
    outs[0].color, outs[0].alpha = color.ctypes.data, alpha.ctypes.data
    target = np.full_like(color, 0.5)
    grad = np.zeros_like(color)
    tp = (C.c_void_p * 1)(target.ctypes.data)
    gp = (C.c_void_p * 1)(grad.ctypes.data)
    st = {k: np.zeros_like(scene.a[k]) for k in ("means", "log_scales", "quaternions", "opacity_logits", "sh")}
    st2 = {k: np.zeros_like(scene.a[k]) for k in st}
    a = _abi.pht_adam()
    a.means, a.log_scales = scene.a["means"].ctypes.data, scene.a["log_scales"].ctypes.data
    a.quaternions, a.opacity_logits, a.sh = (scene.a["quaternions"].ctypes.data,
                                             scene.a["opacity_logits"].ctypes.data, scene.a["sh"].ctypes.data)
    a.m_means, a.v_means = st["means"].ctypes.data, st2["means"].ctypes.data
    a.m_log_scales, a.v_log_scales = st["log_scales"].ctypes.data, st2["log_scales"].ctypes.data
    a.m_quaternions, a.v_quaternions = st["quaternions"].ctypes.data, st2["quaternions"].ctypes.data
    a.m_opacity_logits, a.v_opacity_logits = st["opacity_logits"].ctypes.data, st2["opacity_logits"].ctypes.data
    a.m_sh, a.v_sh = st["sh"].ctypes.data, st2["sh"].ctypes.data
    a.lr_means, a.lr_scales, a.lr_quaternions, a.lr_opacity, a.lr_sh0, a.lr_sh_rest = 1.6e-4, 5e-3, 1e-3, 5e-2, 2.5e-3, 1.25e-4
    a.beta1, a.beta2, a.epsilon = 0.9, 0.999, 1e-15
    a.min_log_scale, a.max_log_scale = -float("inf"), float("inf")
    steps, t0 = 0, time.perf_counter()
    while True:
        steps += 1
        a.correction1 = float(np.float32(1) - np.float32(0.9) ** np.float32(steps))
        a.correction2 = float(np.float32(1) - np.float32(0.999) ** np.float32(steps))
        lib.oracle_train_step(C.byref(g), cams_c, 1, C.byref(s), outs, tp, gp, C.byref(a), threads, 0, 0.2, 1.0)
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    return {"value": steps / el, "unit": "view-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} full CPU-oracle training steps (1 view 1920x1080, {scene.n} Gaussians) "
                      f"in {el:.1f}s"}


def densify_bench(model, cams, targets, opts, refines=3, seed=3, sh_optimizer="fp32"):
    """Densification costs on the benchmark model: the batched graph step with the statistics
    accumulated inside it, the standalone statistics launch for all views, and ADC-IGS
    refinements on the statistics the step gathered (prune / recycle, Gumbel and ranked parent
    selection, row remap of the model, bf16 copies, FP32 moments and quantised SH state, split,
    decay, re-binding the step), wall clock including the host readbacks the refinement branches
    on, each followed by graph steps on the grown model (re-capture in their warm-up)."""
    from photara_studio_amd import densify as D
    from photara_studio_amd.train import LearningRates, TrainStep
    dev = model.means.device
    step = TrainStep(model, cams, targets, opts, LearningRates(), densify=True, sh_optimizer=sh_optimizer)
    step.ensure_capacity()
    step.capture()
    step_s = timed(step.step_graph, 10, 2)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    n = model.size()
    stats = D.make_densification_stats(n, dev)
    obs = []
    for v, cam in enumerate(cams):
        rw = torch.rand((n,), generator=gen, device=dev) * 0.01
        obs.append(D.Observation(rw, step.visibility[v], step.radii[v], cam.width, cam.height, True, True, 1.0, 0.5,
                                 rw, v))
    D.accumulate_views(stats, obs)
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        D.accumulate_views(stats, obs)
    e1.record()
    torch.cuda.synchronize()
    acc_ms = e0.elapsed_time(e1) / reps
    del stats, obs
    dopts = D.apply_strategy_defaults(D.DensifyOptions(densification_cap=n + 400_000 * refines))
    rows = []
    for r in range(refines):
        before = step.model.size()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        counts = step.refine(200 * (r + 2), dopts, 1.0, [0.0, 0.0, 0.0], D.Mt19937(42 + r))
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        after_s = timed(step.step_graph, 5, 1)
        rows.append({"ms": round(ms, 2), "rows_in": before, "pruned": counts.pruned, "grown": counts.grown,
                     "step_ms_after": round(after_s * 1e3, 3)})
    return {"step_ms_with_statistics": round(step_s * 1e3, 4), "accumulate_views_ms": round(acc_ms, 4),
            "views": len(cams), "refines": rows,
            "refine_ms_median": sorted(r["ms"] for r in rows)[len(rows) // 2]}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--gaussians", type=int, default=3_000_000)
    p.add_argument("--views", type=int, default=8)
    p.add_argument("--plain-steps", type=int, default=0, help="plain-step count (default: 4 x steps)")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-densify", action="store_true", help="skip the densification timings")
    p.add_argument("--no-alt", action="store_true",
                   help="skip timing the batched step with the other SH optimizer (profiling runs)")
    p.add_argument("--sh-optimizer", choices=("fp32", "quant"), default="fp32",
                   help="SH optimizer of the batched step: FP32 moments, the plain step's (default), or the "
                        "reference trainer's quantised state (also timed beside it as batched_quant_sh)")
    p.add_argument("--workload", choices=("ring", "c2"), default="ring",
                   help="ring: --views 1920x1080 views on a radius-8 ring; c2: 2x1080p + 3x720p + 3x540p")
    p.add_argument("--tile", type=int, default=16)
    p.add_argument("--batch", type=int, default=64)
    p.add_argument("--autotune", action="store_true",
                   help="sweep tile {16,8} x batch {32,64,128,256} for the batched step and time the fastest")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="first copy 1 GiB device-to-device (a kernel of known traffic for counter scaling)")
    args = p.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (multi-GPU view sharding is a design note, "
              "not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    global torch
    import torch
    from photara_studio_amd.rasterizer import RasterizeOptions
    from photara_studio_amd.scenes import model_from, synthetic_torch, workload_cameras
    from photara_studio_amd.train import LearningRates, TrainStep, to_low_precision

    dev = torch.device("cuda:0")
    n, nv = args.gaussians, args.views
    if args.calibrate_copy:
        src = torch.ones((1 << 28,), dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] calibration copy: 1073741824 bytes read + written", flush=True)
        del src, dst
    cams = workload_cameras("c2" if args.workload == "c2" else f"bench{nv}")
    nv = len(cams)
    res_label = ",".join(sorted({f"{c.width}x{c.height}" for c in cams}, reverse=True))
    opts = RasterizeOptions(active_sh_degree=3, kernel_size=0.3, tile_size=args.tile, batch=args.batch)
    opts.rasterization.compensate_2d = False
    print(f"[bench] {n} gaussians, {nv} views ({res_label}), tile {args.tile}, batch {args.batch}", flush=True)
    targets = make_targets(n, cams, opts, dev, seed=1)
    arrays = synthetic_torch(n, 3.0, 0, dev)

    # plain FP32 step: one view per step, eager launches
    plain = TrainStep(model_from(arrays, dev), [cams[0]], [targets[0]], opts, LearningRates(), sh_optimizer="fp32")
    plain.ensure_capacity()
    plain_steps = args.plain_steps or 4 * args.steps
    plain_s = timed(plain.step_eager, plain_steps, args.warmup)
    pc = plain.counts()
    if pc.overflow:
        raise RuntimeError("plain step overflowed its instance capacity")
    print(f"[bench] plain fp32: {plain_s * 1e3:.3f} ms/step ({pc.instances} instances)", flush=True)
    del plain
    torch.cuda.empty_cache()

    # batched BF16 graph step: all views per step
    model = to_low_precision(model_from(arrays, dev), torch.bfloat16)
    bopts = opts
    tune_rows = None
    if args.autotune:
        from photara_studio_amd.autotune import autotune
        bopts, tune_rows = autotune(model, cams, targets, opts, LearningRates(), steps=3, warmup=1,
                                    log=lambda m: print(m, flush=True))
        print(f"[bench] autotuned: tile {bopts.tile_size}, batch {bopts.batch}", flush=True)
    batched = TrainStep(model, cams, targets, bopts, LearningRates(), sh_optimizer=args.sh_optimizer)
    batched.ensure_capacity()
    batched.capture()
    step_s = timed(batched.step_graph, args.steps, args.warmup)
    bc = batched.counts()
    assert not bc.overflow
    print(f"[bench] batched bf16 graph: {step_s * 1e3:.3f} ms/step, {nv / step_s:.2f} view-steps/s", flush=True)

    # per-stage kernel times (hipEvents on the launch stream) over a few eager steps
    batched.ws.set_profiling(True)
    acc = {}
    prof_steps = 3
    for _ in range(prof_steps):
        batched.step_eager()
        for k, v in batched.ws.stage_times().items():
            acc[k] = acc.get(k, 0.0) + v / prof_steps
    batched.ws.set_profiling(False)
    pixels = sum(c.width * c.height for c in cams)
    bt = int(bopts.tile_size)
    tiles = sum(((c.width + bt - 1) // bt) * ((c.height + bt - 1) // bt) for c in cams)
    ab = algorithmic_bytes(n, nv, int(bc.visible), int(bc.instances), pixels, tiles, 2, True,
                           sh_quant=args.sh_optimizer == "quant")
    dominant = max(acc, key=acc.get)
    dom_gbps = ab[dominant] / (acc[dominant] * 1e-3) / 1e9
    step_bytes = sum(ab.values())
    wname = "c2mixed" if args.workload == "c2" else f"{nv}x1080p"
    workload_key = f"{n // 1_000_000}M-{wname}-bf16" if n % 1_000_000 == 0 else f"{n}-{wname}-bf16"
    traffic, traffic_src = profiled_traffic(dominant, workload_key)
    print(f"[bench] stages ms: " + ", ".join(f"{k}={v:.3f}" for k, v in acc.items()), flush=True)

    # the same batched step with the other SH optimizer, on a fresh model
    other_sh = "quant" if args.sh_optimizer == "fp32" else "fp32"
    del batched
    torch.cuda.empty_cache()
    alt_s = None
    if not args.no_alt:
        alt = TrainStep(to_low_precision(model_from(arrays, dev), torch.bfloat16), cams, targets, bopts,
                        LearningRates(), sh_optimizer=other_sh)
        alt.ensure_capacity()
        alt.capture()
        alt_s = timed(alt.step_graph, args.steps, args.warmup)
        print(f"[bench] batched bf16 graph, {other_sh} SH optimizer: {alt_s * 1e3:.3f} ms/step, "
              f"{nv / alt_s:.2f} view-steps/s", flush=True)
        del alt
        torch.cuda.empty_cache()

    dens = None
    if not args.no_densify:
        # runs last: a refinement replaces the model's arrays
        dens = densify_bench(model, cams, targets, bopts, sh_optimizer=args.sh_optimizer)
        print(f"[bench] densify: step with statistics {dens['step_ms_with_statistics']:.3f} ms; standalone statistics "
              f"of {dens['views']} views {dens['accumulate_views_ms']:.3f} ms; refinements "
              + ", ".join(f"{r['ms']:.1f} ms ({r['rows_in']} rows, -{r['pruned']} +{r['grown']}; then "
                          f"{r['step_ms_after']:.2f} ms/step)" for r in dens["refines"]), flush=True)
    del model
    torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu_baseline:
        import numpy as np
        host = {k: v.cpu().numpy() for k, v in arrays.items()}
        print(f"[bench] CPU baseline (~{args.cpu_seconds:.0f}s)", flush=True)
        cpu = cpu_baseline(host, cams, opts, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": round(nv / step_s, 3),
        "unit": "view-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3dgs-train-step {n} gaussians x {nv} views {res_label}, bf16 storage, hipGraph",
                   "gaussians": n, "views": nv, "resolution": res_label, "tile": int(bopts.tile_size),
                   "batch": int(bopts.batch), "storage": "bf16 sh/log_scales/opacity, fp32 masters",
                   "sh_optimizer": ("FP32 moments (as the plain step)" if args.sh_optimizer == "fp32" else
                                    "quantised (fp16 update + 8-bit log-second)") + ", fp32 structure moments",
                   "parallelism": "none (1 GPU)"},
        "autotune": [[r.tile_size, r.batch, round(r.ms_per_step, 3)] for r in tune_rows] if tune_rows else None,
        "plain_fp32": {"ms_per_step": round(plain_s * 1e3, 4), "steps_per_s": round(1.0 / plain_s, 3),
                       "instances": int(pc.instances)},
        "speedup_vs_plain_fp32": round((nv / step_s) * plain_s, 3),
        f"batched_{other_sh}_sh": None if alt_s is None else {
            "ms_per_step": round(alt_s * 1e3, 4), "view_steps_per_s": round(nv / alt_s, 3),
            "speedup_vs_plain_fp32": round((nv / alt_s) * plain_s, 3)},
        "step_algorithmic_GBps": round(step_bytes / step_s / 1e9, 1),
        "counts": {"visible": int(bc.visible), "instances": int(bc.instances), "fp64_conics": int(bc.fp64_conics)},
        "stage_ms": {k: round(v, 4) for k, v in acc.items()},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(dom_gbps, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(dom_gbps / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "algorithmic_bytes": int(ab[dominant]),
                     "traffic_source": f"profiles/{traffic_src}" if traffic_src else None},
        "cpu_baseline": cpu,
        "densify": dens,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
