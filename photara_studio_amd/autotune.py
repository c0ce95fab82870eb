"""Launch-shape autotuner for the training step: tile edge (8x8 or 16x16), LDS batch depth, and
the blend block shape (waves per workgroup, flat grid or persistent work queue).

The tile edge changes the tile space, the instance count (a splat touches more small tiles), the
tile-key width (and so the number of tile-sort passes) and the blend kernels' quadrant layout; the
batch depth changes how many instances one wave stages in LDS per round; the block shape changes
how (view, tile) work items are grouped into workgroups (1, 2 or 4 waves, each with its own item)
and whether the blends launch one workgroup per item group or a resident grid that pulls items
from a device counter. None changes any pixel's result: parity is defined per tile size against
the oracle run at the same tile size. The reference has a fixed 16x16 tile and 256-thread block
(third_party/splat_drender/include/splat_drender/config.h:10-52); this sweep is this build's own.

The sweep is staged: tile x batch with one-wave workgroups on the flat grid first, then the block
shapes (waves x grid/queue) at the fastest tile and batch.

Every candidate trains on a private copy of the model (so tuning leaves the caller's parameters
untouched) with the whole step captured as one hipGraph, and is timed with hipEvents over
`steps` replays after `warmup` replays. Results can be cached in a JSON file keyed by the
workload's shape.
"""
from __future__ import annotations

import copy
import json
import os
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import Callable, Iterable, List, Optional, Sequence, Tuple

import torch

from .rasterizer import Camera, GaussianModel, RasterizeOptions
from .train import LearningRates, TrainStep

CANDIDATES: Tuple[Tuple[int, int], ...] = ((16, 32), (16, 64), (16, 128), (16, 256), (8, 32), (8, 64), (8, 128))
BLOCK_SHAPES: Tuple[Tuple[int, bool], ...] = ((1, False), (2, False), (4, False), (1, True), (2, True), (4, True))


@dataclass
class TuneResult:
    tile_size: int
    batch: int
    ms_per_step: float
    instances: int
    blend_waves: int = 1
    blend_queue: bool = False


def _apply(options: RasterizeOptions, r: TuneResult) -> RasterizeOptions:
    out = copy.copy(options)
    out.tile_size, out.batch = r.tile_size, r.batch
    out.blend_waves, out.blend_queue = r.blend_waves, r.blend_queue
    return out


def clone_model(model: GaussianModel) -> GaussianModel:
    out = copy.copy(model)
    for name in ("means", "log_scales", "quaternions", "opacity_logits", "sh", "filter_3d", "sh_lp",
                 "log_scales_lp", "opacity_logits_lp"):
        t = getattr(model, name)
        if t is not None:
            setattr(out, name, t.clone())
    return out


def workload_key(model: GaussianModel, cameras: Sequence[Camera], options: RasterizeOptions) -> str:
    lp = "none" if model.sh_lp is None else str(model.sh_lp.dtype).replace("torch.", "")
    res = ",".join(f"{c.width}x{c.height}" for c in cameras)
    return f"n={model.size()};views={res};deg={options.active_sh_degree};lp={lp}"


def time_candidate(model: GaussianModel, cameras: Sequence[Camera], targets: Sequence[torch.Tensor],
                   options: RasterizeOptions, lrs: Optional[LearningRates], steps: int, warmup: int) -> TuneResult:
    step = TrainStep(clone_model(model), cameras, targets, options, lrs)
    instances = step.ensure_capacity()
    step.capture()
    for _ in range(warmup):
        step.step_graph()
    stream = torch.cuda.current_stream()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        step.step_graph()
    end.record(stream)
    end.synchronize()
    ms = start.elapsed_time(end) / steps
    del step
    torch.cuda.empty_cache()
    return TuneResult(int(options.tile_size), int(options.batch), float(ms), int(instances),
                      int(options.blend_waves), bool(options.blend_queue))


def autotune(model: GaussianModel, cameras: Sequence[Camera], targets: Sequence[torch.Tensor],
             options: RasterizeOptions, lrs: Optional[LearningRates] = None,
             candidates: Iterable[Tuple[int, int]] = CANDIDATES,
             block_shapes: Iterable[Tuple[int, bool]] = BLOCK_SHAPES, steps: int = 3, warmup: int = 1,
             cache: Optional[os.PathLike] = None,
             log: Optional[Callable[[str], None]] = None) -> Tuple[RasterizeOptions, List[TuneResult]]:
    """Sweep (tile_size, batch) over `candidates` (one-wave workgroups, flat grid), then the block
    shapes (blend_waves, blend_queue) at the fastest of those; returns options with the fastest
    shape and the per-candidate table (ordered as swept). With `cache`, a previous result for the
    same workload key is reused without timing."""
    key = workload_key(model, cameras, options)
    store = {}
    if cache is not None and Path(cache).exists():
        store = json.loads(Path(cache).read_text())
        if key in store:
            rows = [TuneResult(**r) for r in store[key]]
            return _apply(options, min(rows, key=lambda r: r.ms_per_step)), rows
    rows: List[TuneResult] = []

    def run(o: RasterizeOptions) -> TuneResult:
        r = time_candidate(model, cameras, targets, o, lrs, steps, warmup)
        rows.append(r)
        if log:
            log(f"[autotune] tile {r.tile_size:2d} batch {r.batch:3d} waves {r.blend_waves} "
                f"{'queue' if r.blend_queue else 'grid '}: {r.ms_per_step:8.3f} ms/step ({r.instances} instances)")
        return r

    for ts, batch in candidates:
        if ts == 8 and batch > 128:
            continue
        o = copy.copy(options)
        o.tile_size, o.batch, o.blend_waves, o.blend_queue = int(ts), int(batch), 1, False
        run(o)
    best = min(rows, key=lambda r: r.ms_per_step)
    for waves, queue in block_shapes:
        if waves == 1 and not queue:
            continue  # timed in the first stage
        o = _apply(options, best)
        o.blend_waves, o.blend_queue = int(waves), bool(queue)
        run(o)
    out = _apply(options, min(rows, key=lambda r: r.ms_per_step))
    if cache is not None:
        store[key] = [asdict(r) for r in rows]
        Path(cache).parent.mkdir(parents=True, exist_ok=True)
        Path(cache).write_text(json.dumps(store, indent=1))
    return out, rows
