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
        # the L1 loss is fused into the forward's epilogue: + target read and gradient write
        "blend_forward": 40 * instances + 20 * pixels + 24 * pixels,
        "blend_backward": 40 * instances + 24 * pixels + 40 * visible,
        "gaussian_backward+adam": n * (59 * 4 * 2 + (11 * 8 + 48 * 3 + 16 if sh_quant else 59 * 8) * 2
                                       + (52 * 2 if lp else 0) + 4) + 40 * visible,
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


def profiled_counters(kernel, workload):
    """Measured limiter figures of `kernel` (a stage name) from the `counters:` lines of the
    rocprofv3 summaries under profiles/ (VALU busy fraction, LDS bank-conflict fraction, atomic
    rate, ...), or None."""
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            if not line.startswith("counters:"):
                continue
            kv = dict(x.split("=", 1) for x in line[len("counters:"):].split())
            if kv.pop("kernel", None) == kernel and kv.pop("workload", None) == workload:
                out = {}
                for k, v in kv.items():
                    try:
                        out[k] = float(v)
                    except ValueError:
                        out[k] = v
                out["source"] = f"profiles/{f.name}"
                return out
    return None


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


def host_cpu():
    """CPU model name (/proc/cpuinfo) and the number of logical CPUs the process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    return model, nproc


class CpuStepper:
    """The CPU oracle's whole training step (forward + L1 + backward + Adam, float, OpenMP) over
    one Gaussian set and its views: the reference's math restated on the host cores."""

    def __init__(self, arrays, cams, opts, targets, threads):
        import ctypes as C
        import numpy as np
        from oracle import pyoracle as po
        from photara_studio_amd import _abi
        from photara_studio_amd.rasterizer import make_camera
        self.C, self.np = C, np
        self.scene = po.Scene(arrays)
        self.lib = po.load()
        self.threads = threads
        self.g = self.scene.gaussians()
        self.s = po.settings_for(self.scene, opts)
        nv = len(cams)
        self.nv = nv
        self.cams = (_abi.pht_camera * nv)(*[make_camera(c, opts) for c in cams])
        self.color = [np.zeros((3, c.height, c.width), np.float32) for c in cams]
        self.alpha = [np.zeros((c.height, c.width), np.float32) for c in cams]
        self.outs = (_abi.pht_view_outputs * nv)()
        for k in range(nv):
            self.outs[k].color, self.outs[k].alpha = self.color[k].ctypes.data, self.alpha[k].ctypes.data
        self.targets = [np.ascontiguousarray(t, np.float32) for t in targets]
        self.grad = [np.zeros_like(c) for c in self.color]
        self.tp = (C.c_void_p * nv)(*[t.ctypes.data for t in self.targets])
        self.gp = (C.c_void_p * nv)(*[g.ctypes.data for g in self.grad])
        a = self.scene.a
        self.m = {k: np.zeros_like(a[k]) for k in ("means", "log_scales", "quaternions", "opacity_logits", "sh")}
        self.v = {k: np.zeros_like(a[k]) for k in self.m}
        ad = _abi.pht_adam()
        for k in self.m:
            setattr(ad, k, a[k].ctypes.data)
            setattr(ad, "m_" + k, self.m[k].ctypes.data)
            setattr(ad, "v_" + k, self.v[k].ctypes.data)
        ad.lr_means, ad.lr_scales, ad.lr_quaternions, ad.lr_opacity, ad.lr_sh0, ad.lr_sh_rest = \
            1.6e-4, 5e-3, 1e-3, 5e-2, 2.5e-3, 1.25e-4
        ad.beta1, ad.beta2, ad.epsilon = 0.9, 0.999, 1e-15
        ad.min_log_scale, ad.max_log_scale = -float("inf"), float("inf")
        self.adam = ad
        self.t = 0

    def step(self) -> float:
        """One training step; returns its wall-clock seconds (steady clock, no I/O)."""
        np, C = self.np, self.C
        self.t += 1
        self.adam.correction1 = float(np.float32(1) - np.float32(0.9) ** np.float32(self.t))
        self.adam.correction2 = float(np.float32(1) - np.float32(0.999) ** np.float32(self.t))
        t0 = time.perf_counter()
        self.lib.oracle_train_step(C.byref(self.g), self.cams, self.nv, C.byref(self.s), self.outs, self.tp, self.gp,
                                   C.byref(self.adam), self.threads, 0, 0.2, 1.0)
        return time.perf_counter() - t0


def cpu_render(arrays, cams, opts, threads):
    from oracle import pyoracle as po
    return [c.copy() for c in po.OracleRun(po.Scene(arrays), cams, opts, 32, threads=threads, key_cap=1).color]


def cpu_baseline(arrays, cams, opts, seconds):
    """The CPU oracle's training step on the host cores.
    - The bench workload, bounded: whole steps of one 1920x1080 view of the same 3M-Gaussian scene
      (against the GPU run's target of that view) until ~`seconds` elapse -> view-steps/s.
    - The small CPU configuration (10,000 Gaussians, one 256x256 view from (0,0,4), target rendered
      with means perturbed by N(0, 0.01)): median of 10 steps after 2 warm-ups.
    - 1,000,000 Gaussians in [-3,3]^3, one 1920x1080 view at distance 8, target rendered from a
      seeded ground-truth set: 3 steps after 1 warm-up.
    Threads: all CPUs the process may use, at most 16 (a one-GPU machine's CPU share)."""
    import numpy as np
    from photara_studio_amd.scenes import synthetic_arrays, workload_cameras
    model, nproc = host_cpu()
    threads = int(os.environ.get("PHT_CPU_THREADS", "0")) or min(16, nproc)
    cam = cams[0]
    tgt = arrays.pop("_target0")
    st = CpuStepper(arrays, [cam], opts, [tgt], threads)
    times = []
    t_start = time.perf_counter()
    while True:
        times.append(st.step())
        if time.perf_counter() - t_start >= seconds or len(times) >= 50:
            break
    del st
    out = {"value": len(times) / sum(times), "unit": "view-steps/s", "cores": threads, "kind": "port",
           "sample": f"{len(times)} full CPU-oracle training steps (forward + L1 + backward + Adam) of 1 view "
                     f"{cam.width}x{cam.height} of the {len(arrays['means'])}-Gaussian bench scene, "
                     f"{sum(times):.1f}s",
           "cpu_model": model, "nproc": nproc, "threads": threads}
    # the two CPU-sized configurations
    c0 = synthetic_arrays(10_000, 1.0, 0)
    c0_cams = workload_cameras("c0")
    pert = dict(c0)
    pert["means"] = (c0["means"] + np.random.default_rng(1).normal(0, 0.01, c0["means"].shape)).astype(np.float32)
    st = CpuStepper(c0, c0_cams, opts, cpu_render(pert, c0_cams, opts, threads), threads)
    for _ in range(2):
        st.step()
    t0s = sorted(st.step() for _ in range(10))
    out["c0_step_ms_median"] = round(1e3 * (t0s[4] + t0s[5]) / 2, 3)
    c1 = synthetic_arrays(1_000_000, 3.0, 0)
    c1_cams = workload_cameras("c1")
    st = CpuStepper(c1, c1_cams, opts, cpu_render(synthetic_arrays(1_000_000, 3.0, 1), c1_cams, opts, threads),
                    threads)
    st.step()
    t1s = [st.step() for _ in range(3)]
    out["c1_step_ms"] = [round(1e3 * t, 1) for t in t1s]
    return out


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
    p.add_argument("--waves", type=int, default=1, help="blend waves per workgroup (1, 2, 4) without --autotune")
    p.add_argument("--queue", action="store_true", help="persistent work-queue blends without --autotune")
    p.add_argument("--view-groups", type=int, default=2,
                   help="split the batched step's views into this many groups on their own streams")
    p.add_argument("--no-autotune", dest="autotune", action="store_false",
                   help="skip the launch-shape sweep (tile {16,8} x batch {32..256}, then waves {1,2,4} x "
                        "grid/queue) that picks the batched step's shape")
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
    opts = RasterizeOptions(active_sh_degree=3, kernel_size=0.3, tile_size=args.tile, batch=args.batch,
                            blend_waves=args.waves, blend_queue=args.queue)
    opts.rasterization.compensate_2d = False
    print(f"[bench] {n} gaussians, {nv} views ({res_label}), tile {args.tile}, batch {args.batch}", flush=True)
    targets = make_targets(n, cams, opts, dev, seed=1)
    arrays = synthetic_torch(n, 3.0, 0, dev)

    # launch shapes: each step kind gets the fastest shape for its own workload, so the plain step
    # is compared at its best (a one-view step favours other block shapes than an 8-view one)
    model = to_low_precision(model_from(arrays, dev), torch.bfloat16)
    bopts = popts = opts
    tune_rows = plain_tune_rows = None
    if args.autotune:
        from photara_studio_amd.autotune import autotune
        bopts, tune_rows = autotune(model, cams, targets, opts, LearningRates(), steps=3, warmup=1,
                                    log=lambda m: print(m, flush=True))
        print(f"[bench] autotuned (batched): tile {bopts.tile_size}, batch {bopts.batch}, waves {bopts.blend_waves}, "
              f"{'queue' if bopts.blend_queue else 'grid'}", flush=True)
        torch.cuda.empty_cache()
        popts, plain_tune_rows = autotune(model_from(arrays, dev), [cams[0]], [targets[0]], opts, LearningRates(),
                                          steps=6, warmup=2, log=lambda m: print(m.replace("[autotune]", "[autotune plain]"),
                                                                                 flush=True))
        print(f"[bench] autotuned (plain): tile {popts.tile_size}, batch {popts.batch}, waves {popts.blend_waves}, "
              f"{'queue' if popts.blend_queue else 'grid'}", flush=True)
        torch.cuda.empty_cache()

    # plain FP32 step: one view per step, eager launches
    plain = TrainStep(model_from(arrays, dev), [cams[0]], [targets[0]], popts, LearningRates(), sh_optimizer="fp32")
    plain.ensure_capacity()
    plain_steps = args.plain_steps or 4 * args.steps
    plain_s = timed(plain.step_eager, plain_steps, args.warmup)
    pc = plain.counts()
    if pc.overflow:
        raise RuntimeError("plain step overflowed its instance capacity")
    print(f"[bench] plain fp32: {plain_s * 1e3:.3f} ms/step ({pc.instances} instances)", flush=True)
    plain.ws.set_profiling(True)
    plain_acc = {}
    for _ in range(3):
        plain.step_eager()
        for k, v in plain.ws.stage_times().items():
            plain_acc[k] = plain_acc.get(k, 0.0) + v / 3
    plain.ws.set_profiling(False)
    print(f"[bench] plain stages ms: " + ", ".join(f"{k}={v:.3f}" for k, v in plain_acc.items()), flush=True)
    del plain
    torch.cuda.empty_cache()

    # batched BF16 graph step: all views per step
    batched = TrainStep(model, cams, targets, bopts, LearningRates(), sh_optimizer=args.sh_optimizer,
                        view_groups=args.view_groups)
    batched.ensure_capacity()
    batched.capture()
    step_s = timed(batched.step_graph, args.steps, args.warmup)
    bc = batched.counts()
    assert not bc.overflow
    print(f"[bench] batched bf16 graph: {step_s * 1e3:.3f} ms/step, {nv / step_s:.2f} view-steps/s", flush=True)

    # per-stage kernel times (hipEvents on the launch stream) over a few eager steps of the same
    # workload in one view group: with several groups the stages of different groups overlap in
    # time, so per-group events would not add up to kernel time
    staged = batched
    if args.view_groups > 1:
        staged = TrainStep(to_low_precision(model_from(arrays, dev), torch.bfloat16), cams, targets, bopts,
                           LearningRates(), sh_optimizer=args.sh_optimizer)
        staged.ensure_capacity()
    staged.ws.set_profiling(True)
    acc = {}
    prof_steps = 3
    for _ in range(prof_steps):
        staged.step_eager()
        for k, v in staged.ws.stage_times().items():
            acc[k] = acc.get(k, 0.0) + v / prof_steps
    staged.ws.set_profiling(False)
    if staged is not batched:
        del staged
        torch.cuda.empty_cache()
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
    limiter = profiled_counters(dominant, workload_key)
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
        host["_target0"] = targets[0].cpu().numpy()
        print(f"[bench] CPU baseline (~{args.cpu_seconds:.0f}s + the 10k and 1M configurations)", flush=True)
        cpu = cpu_baseline(host, cams, opts, args.cpu_seconds)
        print(f"[bench] CPU: {cpu['value']:.3f} view-steps/s, 10k median {cpu['c0_step_ms_median']} ms, "
              f"1M {cpu['c1_step_ms']} ms ({cpu['threads']} threads, {cpu['cpu_model']})", flush=True)

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
                   "batch": int(bopts.batch), "blend_waves": int(bopts.blend_waves),
                   "blend_launch": "persistent work queue" if bopts.blend_queue else "flat grid",
                   "view_groups": int(args.view_groups),
                   "storage": "bf16 sh/log_scales/opacity, fp32 masters",
                   "sh_optimizer": ("FP32 moments (as the plain step)" if args.sh_optimizer == "fp32" else
                                    "quantised (fp16 update + 8-bit log-second)") + ", fp32 structure moments",
                   "parallelism": "none (1 GPU)"},
        "autotune": [[r.tile_size, r.batch, r.blend_waves, "queue" if r.blend_queue else "grid",
                      round(r.ms_per_step, 3)] for r in tune_rows] if tune_rows else None,
        "plain_fp32": {"ms_per_step": round(plain_s * 1e3, 4), "steps_per_s": round(1.0 / plain_s, 3),
                       "instances": int(pc.instances), "tile": int(popts.tile_size), "batch": int(popts.batch),
                       "blend_waves": int(popts.blend_waves),
                       "blend_launch": "persistent work queue" if popts.blend_queue else "flat grid",
                       "stage_ms": {k: round(v, 4) for k, v in plain_acc.items()},
                       "autotune": [[r.tile_size, r.batch, r.blend_waves, "queue" if r.blend_queue else "grid",
                                     round(r.ms_per_step, 3)] for r in plain_tune_rows] if plain_tune_rows else None},
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
                     "traffic_source": f"profiles/{traffic_src}" if traffic_src else None,
                     "limiter": limiter},
        "cpu_baseline": cpu,
        "densify": dens,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
