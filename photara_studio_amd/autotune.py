"""Launch-shape autotuner for the training step: tile edge (8x8 or 16x16) x LDS batch depth.

The tile edge changes the tile space, the instance count (a splat touches more small tiles), the
tile-key width (and so the number of tile-sort passes) and the blend kernels' quadrant layout; the
batch depth changes how many instances one wave stages in LDS per round. Neither changes any
pixel's result: parity is defined per tile size against the oracle run at the same tile size.
The reference has a fixed 16x16 / 256 shape (third_party/splat_drender/include/splat_drender/
config.h:10-52); this sweep is this build's own.

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


@dataclass
class TuneResult:
    tile_size: int
    batch: int
    ms_per_step: float
    instances: int


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
    return TuneResult(int(options.tile_size), int(options.batch), float(ms), int(instances))


def autotune(model: GaussianModel, cameras: Sequence[Camera], targets: Sequence[torch.Tensor],
             options: RasterizeOptions, lrs: Optional[LearningRates] = None,
             candidates: Iterable[Tuple[int, int]] = CANDIDATES, steps: int = 3, warmup: int = 1,
             cache: Optional[os.PathLike] = None,
             log: Optional[Callable[[str], None]] = None) -> Tuple[RasterizeOptions, List[TuneResult]]:
    """Sweep (tile_size, batch) over `candidates`; returns options with the fastest shape and the
    per-candidate table (ordered as swept). With `cache`, a previous result for the same workload
    key is reused without timing."""
    key = workload_key(model, cameras, options)
    store = {}
    if cache is not None and Path(cache).exists():
        store = json.loads(Path(cache).read_text())
        if key in store:
            rows = [TuneResult(**r) for r in store[key]]
            best = min(rows, key=lambda r: r.ms_per_step)
            out = copy.copy(options)
            out.tile_size, out.batch = best.tile_size, best.batch
            return out, rows
    rows: List[TuneResult] = []
    for ts, batch in candidates:
        if ts == 8 and batch > 128:
            continue
        o = copy.copy(options)
        o.tile_size, o.batch = int(ts), int(batch)
        r = time_candidate(model, cameras, targets, o, lrs, steps, warmup)
        rows.append(r)
        if log:
            log(f"[autotune] tile {ts:2d} batch {batch:3d}: {r.ms_per_step:8.3f} ms/step ({r.instances} instances)")
    best = min(rows, key=lambda r: r.ms_per_step)
    out = copy.copy(options)
    out.tile_size, out.batch = best.tile_size, best.batch
    if cache is not None:
        store[key] = [asdict(r) for r in rows]
        Path(cache).parent.mkdir(parents=True, exist_ok=True)
        Path(cache).write_text(json.dumps(store, indent=1))
    return out, rows
