"""ADC-IGS densification on the GPU: statistics, prune / recycle / cap, parent selection, the
random-axis split and the model / optimizer row remap.

Mirrors photara::splat::densification (densification.hpp:17-136, densification.cpp:114-161,
densification_igs.cpp:224-515, densification_adc_plus.cpp:18-196) and
detail::accumulate_densification_stats (adc_ops.cpp:23-62). Every per-row step runs in the HIP
kernels behind the C ABI (``pht_densify_*``, ``pht_select_rows``, ``pht_igs_*``,
``pht_gather_rows``, ``pht_adc_decay``). The host sequences them and reads back the few counts
the refinement branches on, as the reference does; mask algebra on byte masks and the index
compaction of the keep / chosen masks use torch device ops.

Random numbers follow the reference's: ``Mt19937`` is std::mt19937 (the Gumbel seeds are two of
its 32-bit outputs, high word first, igs_weighted_sample_cuda's seed), and ``normal_floats`` draws
the split offsets with the C++ standard library's std::normal_distribution<float> over that
engine's state, as grow_igs_random_gpu does.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import NamedTuple, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _abi
from ._abi import check
from .rasterizer import GaussianModel, ShAdamQuant, StructureAdamUpdate, _stream, lp_format_of


def _vp(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


# ---------------------------------------------------------------- random numbers
class Mt19937:
    """std::mt19937 with the standard seeding (numpy's legacy RandomState seeds the same way)."""

    def __init__(self, seed: int = 5489):
        self._rs = np.random.RandomState(seed & 0xFFFFFFFF)

    def __call__(self) -> int:
        return int(self._rs.randint(0, 1 << 32, dtype=np.uint64))

    def raw(self, count: int) -> np.ndarray:
        return self._rs.randint(0, 1 << 32, size=count, dtype=np.uint64)

    def get_state(self):
        return self._rs.get_state()

    def set_state(self, state) -> None:
        self._rs.set_state(state)


def normal_floats(random: Mt19937, count: int) -> np.ndarray:
    """``count`` draws of a fresh std::normal_distribution<float>(0, 1) over the engine, by the
    standard library itself (pht_host_normal_floats); the engine advances in place."""
    _, key, pos, _, _ = random.get_state()
    key = np.ascontiguousarray(key, np.uint32).copy()
    p = C.c_int32(int(pos))
    out = np.empty(count, np.float32)
    check(_abi.load().pht_host_normal_floats(C.c_void_p(key.ctypes.data), C.byref(p),
                                             C.c_void_p(out.ctypes.data) if count else None, int(count)))
    random.set_state(("MT19937", key, int(p.value), 0, 0.0))
    return out


def gumbel_seed(random: Mt19937) -> int:
    high = random()
    low = random()
    return (high << 32) | low


# ---------------------------------------------------------------- options and schedule
@dataclass
class DensifyOptions:
    """The TrainingOptions fields the ADC-IGS refinement reads (options.hpp:106-224)."""
    iterations: int = 30_000
    enable_densification: bool = True
    densification_cap: int = 3_000_000
    refine_start_iter: int = 0
    refine_stop_iter: int = 0
    refine_stop_num_iter: int = 2_500
    grow_stop_iter: int = 0
    refine_every: int = 0
    densify_gradient_threshold: float = 0.0025
    densify_select_fraction: float = 0.25
    densify_screen_threshold: float = 0.25
    oversize_penalty: float = 0.0
    dense_recycle_fraction: float = 0.01
    prune_opacity: float = 1.0 / 255.0
    opacity_decay: float = 0.004
    scale_decay: float = 0.0


def grow_stop_iteration(options: DensifyOptions) -> int:
    """options.hpp:500-506 (ADC strategies): half of the iterations unless set."""
    return options.grow_stop_iter if options.grow_stop_iter != 0 else max(1, options.iterations // 2)


def apply_strategy_defaults(options: DensifyOptions) -> DensifyOptions:
    """apply_strategy_defaults for adc_igs (densification.cpp:211-230); the trainer's learning-rate
    presets it also sets are outside the refinement."""
    options.grow_stop_iter = grow_stop_iteration(options)
    options.densify_gradient_threshold = 0.0025
    options.densify_select_fraction = 0.25
    options.densify_screen_threshold = 0.5
    return options


def strategy_schedule(options: DensifyOptions) -> Tuple[int, int, int]:
    """strategy_schedule for adc_igs (densification.cpp:114-140): (start, stop, every)."""
    stop_iter = options.refine_stop_iter if options.refine_stop_iter != 0 else 14_000
    from_end = options.iterations - options.refine_stop_num_iter if options.iterations > options.refine_stop_num_iter \
        else 0
    stop = min(max(stop_iter, from_end), options.iterations)
    return options.refine_start_iter, stop, options.refine_every if options.refine_every != 0 else 200


def is_refinement_iteration(iteration: int, options: DensifyOptions) -> bool:
    """densification.cpp:163-174 for adc_igs."""
    start, stop, every = strategy_schedule(options)
    return not (iteration <= start or iteration >= stop or every == 0 or iteration % every != 0)


# ---------------------------------------------------------------- statistics
@dataclass
class DensificationStats:
    """detail::DensificationStats (adc_ops.cpp:8-19)."""
    gradient: torch.Tensor = None
    count: torch.Tensor = None
    max_screen_radius: torch.Tensor = None
    priority: torch.Tensor = None
    max_geometry_gradient: torch.Tensor = None
    first_view: torch.Tensor = None  # int32
    view_support: torch.Tensor = None
    accumulated_error: torch.Tensor = None
    growth_cap: int = 0

    def size(self) -> int:
        return int(self.gradient.shape[0])

    def abi(self) -> _abi.pht_densify_stats:
        s = _abi.pht_densify_stats()
        for name in ("gradient", "count", "max_screen_radius", "priority", "max_geometry_gradient", "first_view",
                     "view_support", "accumulated_error"):
            t = getattr(self, name)
            if t.dtype != (torch.int32 if name == "first_view" else torch.float32) or not t.is_contiguous():
                raise ValueError(f"densification statistic {name} has the wrong dtype or layout")
            setattr(s, name, t.data_ptr())
        return s


def make_densification_stats(count: int, device) -> DensificationStats:
    z = lambda dt=torch.float32: torch.zeros((count,), dtype=dt, device=device)  # noqa: E731
    return DensificationStats(z(), z(), z(), z(), z(), z(torch.int32), z(), z())


@dataclass
class Observation:
    """One view's inputs to the statistics (accumulate_densification_stats arguments)."""
    refine_weight: torch.Tensor
    visibility: torch.Tensor
    radii: torch.Tensor
    width: int
    height: int
    use_maximum: bool = True
    require_contribution_visibility: bool = True
    step_score_power: float = 1.0
    oversize_screen_threshold: float = 0.0
    geometry_gradient: Optional[torch.Tensor] = None
    view_index: int = -1
    image_error: Optional[torch.Tensor] = None
    normalized_size: Optional[torch.Tensor] = None


def accumulate_views(stats: DensificationStats, observations: Sequence[Observation]) -> None:
    """Folds the observations, in order, into the statistics with one launch (the same result as
    one accumulate_densification_stats call per observation)."""
    n = stats.size()
    if n == 0 or not observations:
        return
    if len(observations) > _abi.PHT_MAX_VIEWS:
        raise ValueError("at most 32 observations per launch")
    obs = (_abi.pht_densify_observation * len(observations))()
    for k, o in enumerate(observations):
        for name, t, dt in (("refine_weight", o.refine_weight, torch.float32), ("visibility", o.visibility, torch.float32),
                            ("radii", o.radii, torch.int32), ("geometry_gradient", o.geometry_gradient, torch.float32),
                            ("image_error", o.image_error, torch.float32),
                            ("normalized_size", o.normalized_size, torch.float32)):
            if t is None:
                continue
            if t.dtype != dt or t.numel() != n or not t.is_contiguous() or t.device != stats.gradient.device:
                raise ValueError(f"densification observation {name} must be a contiguous [{n}] {dt} tensor")
            setattr(obs[k], name, t.data_ptr())
        obs[k].width, obs[k].height = int(o.width), int(o.height)
        obs[k].use_maximum = int(bool(o.use_maximum))
        obs[k].require_contribution_visibility = int(bool(o.require_contribution_visibility))
        obs[k].step_score_power = float(o.step_score_power)
        obs[k].oversize_screen_threshold = float(o.oversize_screen_threshold)
        obs[k].view_index = int(o.view_index)
    s = stats.abi()
    check(_abi.load().pht_densify_accumulate(C.byref(s), obs, len(observations), n, C.c_void_p(_stream())))


def accumulate_densification_stats(refine_weight: torch.Tensor, visibility: torch.Tensor, radii: torch.Tensor,
                                   stats: DensificationStats, width: int, height: int, use_maximum: bool,
                                   require_contribution_visibility: bool, step_score_power: float,
                                   oversize_screen_threshold: float, geometry_gradient: Optional[torch.Tensor] = None,
                                   view_index: int = -1, image_error: Optional[torch.Tensor] = None,
                                   normalized_size: Optional[torch.Tensor] = None) -> None:
    """detail::accumulate_densification_stats (adc_ops.cpp:23-62)."""
    accumulate_views(stats, [Observation(refine_weight, visibility, radii, width, height, use_maximum,
                                         require_contribution_visibility, step_score_power,
                                         oversize_screen_threshold, geometry_gradient, view_index, image_error,
                                         normalized_size)])


# ---------------------------------------------------------------- row primitives
class PruneMasks(NamedTuple):
    keep: torch.Tensor       # uint8 [N]
    hard: torch.Tensor       # uint8 [N]
    opacities: torch.Tensor  # float32 [N]


def _fp32_gaussians(model: GaussianModel) -> _abi.pht_gaussians:
    g = _abi.pht_gaussians()
    g.count = model.size()
    for name in ("means", "log_scales", "quaternions", "opacity_logits", "sh"):
        t = getattr(model, name)
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous float32 masters")
        setattr(g, name, t.data_ptr())
    return g


def prune_masks(model: GaussianModel, minimum_opacity: float, maximum_bounds: float,
                scene_center: Sequence[float]) -> PruneMasks:
    """detail::prune_masks (ops.hpp:131-141)."""
    n = model.size()
    dev = model.means.device
    m = PruneMasks(torch.empty((n,), dtype=torch.uint8, device=dev), torch.empty((n,), dtype=torch.uint8, device=dev),
                   torch.empty((n,), dtype=torch.float32, device=dev))
    if n == 0:
        return m
    g = _fp32_gaussians(model)
    center = (C.c_float * 3)(*[float(c) for c in scene_center])
    check(_abi.load().pht_densify_prune_masks(C.byref(g), int(model.sh.numel() // n), float(minimum_opacity),
                                              float(maximum_bounds), center, _vp(m.keep), _vp(m.hard),
                                              _vp(m.opacities), C.c_void_p(_stream())))
    return m


class Selection(NamedTuple):
    eligible: int
    selected: int
    threshold: int  # key of the last selected row (0 when none)


def select_rows(values: torch.Tensor, k: int, mode: int, out: torch.Tensor, out_value: int = 1,
                mask: Optional[torch.Tensor] = None, exclude: Optional[torch.Tensor] = None,
                seed: int = 0) -> Selection:
    """pht_select_rows: marks the selected rows in ``out`` (uint8) and returns the counts."""
    n = int(values.numel())
    dev = values.device
    for t, name in ((mask, "mask"), (exclude, "exclude"), (out, "out")):
        if t is not None and (t.dtype != torch.uint8 or t.numel() != n or not t.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous uint8 tensor of {n} rows")
    lib = _abi.load()
    wsb = int(lib.pht_select_workspace_bytes(n))
    ws = torch.empty((wsb,), dtype=torch.uint8, device=dev)
    res = torch.zeros((4,), dtype=torch.int32, device=dev)  # counts[2] then the threshold key
    check(lib.pht_select_rows(_vp(values), _vp(mask), _vp(exclude), n, int(max(k, 0)), int(mode),
                              C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), _vp(out), int(out_value), C.c_void_p(res.data_ptr()),
                              C.c_void_p(res.data_ptr() + 8), _vp(ws), wsb, C.c_void_p(_stream())))
    host = res.cpu().numpy().view(np.uint32)
    return Selection(int(host[0]), int(host[1]), int(host[2]) | (int(host[3]) << 32))


def key_value(key: int, mode: int) -> float:
    return float(_abi.load().pht_select_key_value(C.c_uint64(key), int(mode)))


def weighted_gumbel_sample(weights: torch.Tensor, requested: int, random: Mt19937, out: torch.Tensor,
                           exclude: Optional[torch.Tensor] = None) -> int:
    """weighted_gumbel_sample, CUDA branch (densification_igs.cpp:168-181): no draw for an empty
    request; otherwise two 32-bit outputs of the engine seed the device keys."""
    if requested == 0 or weights.numel() == 0:
        return 0
    seed = gumbel_seed(random)
    return select_rows(weights, requested, _abi.PHT_SELECT_GUMBEL, out, 1, exclude=exclude, seed=seed).selected


def positive_upper_median(values: torch.Tensor, positive: torch.Tensor, positive_count: int) -> float:
    """positive_upper_median (densification_igs.cpp:146-166): the (count/2 + 1)-th smallest
    positive value, by radix select."""
    if positive_count == 0:
        return 0.0
    scratch = torch.zeros_like(positive)
    sel = select_rows(values, positive_count // 2 + 1, _abi.PHT_SELECT_SMALLEST, scratch, 1, mask=positive)
    return key_value(sel.threshold, _abi.PHT_SELECT_SMALLEST) if sel.selected else 0.0


class IgsSelection(NamedTuple):
    parents: torch.Tensor  # int32, ascending
    replacement: int
    oversized: int
    growth: int


def select_igs_parents(replacement_weights: torch.Tensor, oversize_scores: Optional[torch.Tensor],
                       growth_weights: Optional[torch.Tensor], replacement_slots: int, desired_growth: int,
                       capacity: int, random: Optional[Mt19937] = None) -> IgsSelection:
    """select_igs_parents (densification_igs.cpp:226-277)."""
    random = random if random is not None else Mt19937(0)
    chosen = torch.zeros((replacement_weights.numel(),), dtype=torch.uint8, device=replacement_weights.device)
    replacement = weighted_gumbel_sample(replacement_weights, min(replacement_slots, capacity), random, chosen)
    remaining = capacity - replacement
    oversized = growth = 0
    if remaining and oversize_scores is not None:
        # rank actual severity among positive finite scores not already chosen
        oversized = select_rows(oversize_scores, remaining, _abi.PHT_SELECT_LARGEST_POSITIVE, chosen, 1,
                                exclude=chosen).selected
        remaining -= oversized
    if remaining and desired_growth and growth_weights is not None:
        growth = weighted_gumbel_sample(growth_weights, min(remaining, desired_growth), random, chosen, exclude=chosen)
    parents = torch.nonzero(chosen).squeeze(1).to(torch.int32)
    return IgsSelection(parents, replacement, oversized, growth)


@dataclass
class AdamStates:
    """densification::AdamStates (densification.hpp:34-71): the structure optimizer's moments
    (FP32 SH moments included when it carries them) and the quantised SH state."""
    structure: Optional[StructureAdamUpdate] = None
    sh_quant: Optional[ShAdamQuant] = None


_STRUCTURE_MOMENTS = ("means_first", "means_second", "scales_first", "scales_second", "rotations_first",
                      "rotations_second", "opacity_first", "opacity_second", "sh_first", "sh_second")


def _gather(pairs, rows: torch.Tensor) -> None:
    bufs = (_abi.pht_row_buffer * len(pairs))()
    for k, (src, dst) in enumerate(pairs):
        if not src.is_contiguous() or not dst.is_contiguous() or src.dtype != dst.dtype:
            raise ValueError("row remap needs contiguous arrays of one dtype")
        bufs[k].src, bufs[k].dst = src.data_ptr(), dst.data_ptr()
        bufs[k].row_bytes = int(src.numel() // src.shape[0]) * src.element_size()
    check(_abi.load().pht_gather_rows(bufs, len(pairs), _vp(rows), int(rows.numel()), C.c_void_p(_stream())))


def remap_rows(model: GaussianModel, states: AdamStates, param_rows: torch.Tensor,
               adam_rows: Optional[torch.Tensor] = None) -> None:
    """New model / optimizer arrays: row r of each takes old row param_rows[r] (model) or
    adam_rows[r] (moments; -1 = a zero row). Replaces select_training_rows_gpu + append_model +
    append_zero_adam + zero_adam_rows (densification_adc_plus.cpp:18-95, 148-196)."""
    adam_rows = param_rows if adam_rows is None else adam_rows
    n_new = int(param_rows.numel())
    if model.size() == 0:
        raise ValueError("empty model")
    pairs = []

    def fresh(t):
        if t is None:
            return None
        new = torch.empty((n_new,) + tuple(t.shape[1:]), dtype=t.dtype, device=t.device)
        pairs.append((t, new))
        return new

    updated = {name: fresh(getattr(model, name)) for name in ("means", "log_scales", "quaternions", "opacity_logits",
                                                              "sh", "sh_lp", "log_scales_lp", "opacity_logits_lp")}
    if pairs:
        _gather(pairs, param_rows)
    for name, t in updated.items():
        if t is not None:
            setattr(model, name, t)
    pairs = []
    moments = {}
    if states.structure is not None:
        moments = {name: fresh(getattr(states.structure, name)) for name in _STRUCTURE_MOMENTS}
    quant = {}
    if states.sh_quant is not None and states.sh_quant.active():
        quant = {name: fresh(getattr(states.sh_quant, name)) for name in ("first", "packed", "bounds")}
    if pairs:
        _gather(pairs, adam_rows)
    for name, t in moments.items():
        if t is not None:
            setattr(states.structure, name, t)
    for name, t in quant.items():
        setattr(states.sh_quant, name, t)


def split_gaussians(model: GaussianModel, parents: torch.Tensor, child_base: int, samples: torch.Tensor,
                    mode: int, minimum_opacity: float, split_at_screen_size: float = 0.0,
                    screen_sizes: Optional[torch.Tensor] = None, split_opacity_k: float = 0.0) -> None:
    """detail::split_gaussians (split_gaussians_kernel, cuda_ops.cu:1720-1845) on rows already
    holding the parents' parameters: parent rows ``parents``, child rows child_base + j."""
    count = int(parents.numel())
    if count == 0:
        return
    p = _abi.pht_split_params()
    p.means, p.log_scales = model.means.data_ptr(), model.log_scales.data_ptr()
    p.opacity_logits, p.quaternions = model.opacity_logits.data_ptr(), model.quaternions.data_ptr()
    p.lp_format = lp_format_of(model)
    if p.lp_format:
        p.log_scales_lp, p.opacity_logits_lp = model.log_scales_lp.data_ptr(), model.opacity_logits_lp.data_ptr()
    p.mode, p.minimum_opacity = int(mode), float(minimum_opacity)
    p.split_at_screen_size, p.split_opacity_k = float(split_at_screen_size), float(split_opacity_k)
    check(_abi.load().pht_densify_split(C.byref(p), _vp(parents), count, int(child_base), _vp(samples),
                                        _vp(screen_sizes), C.c_void_p(_stream())))


def adc_decay_factors(opacity_decay: float, scale_decay: float) -> Tuple[float, float]:
    """cuda_backend::apply_adc_decay's host part (cuda_ops.cu:3388-3399)."""
    scale_factor = max(np.float32(1.0) - np.float32(scale_decay), np.float32(1e-6))
    return float(max(np.float32(opacity_decay), np.float32(0.0))), float(np.log(np.float32(scale_factor)))


def apply_adc_decay(model: GaussianModel, opacity_decay: float, scale_decay: float) -> None:
    """detail::apply_adc_decay (ops.cpp:588-598)."""
    n = model.size()
    if n == 0:
        return
    od, lsd = adc_decay_factors(opacity_decay, scale_decay)
    fmt = lp_format_of(model)
    check(_abi.load().pht_adc_decay(_vp(model.log_scales), _vp(model.opacity_logits),
                                    _vp(model.log_scales_lp) if fmt else None,
                                    _vp(model.opacity_logits_lp) if fmt else None, fmt, n, od, lsd,
                                    C.c_void_p(_stream())))


# ---------------------------------------------------------------- the refinement
class RefinementCounts(NamedTuple):
    grown: int
    pruned: int


class IgsStrategy:
    """densification::IgsStrategy (densification_igs.hpp, densification_igs.cpp:279-515)."""

    def __init__(self):
        self.last = {}

    def refine(self, model: GaussianModel, stats: DensificationStats, iteration: int, scene_extent: float,
               scene_center: Sequence[float], options: DensifyOptions, random: Mt19937,
               states: AdamStates) -> Tuple[RefinementCounts, DensificationStats]:
        """One refinement. Returns the counts and the fresh statistics (the reference assigns
        ``stats = make_densification_stats(...)``)."""
        old = model.size()
        if not is_refinement_iteration(iteration, options) or old == 0:
            return RefinementCounts(0, 0), stats
        if options.oversize_penalty > 0.0:
            raise NotImplementedError("the oversize log-scale clip (oversize_penalty > 0) is not built")
        if stats.size() != old:
            raise ValueError("densification statistics do not match the model size")
        dev = model.means.device
        lib = _abi.load()
        masks = prune_masks(model, options.prune_opacity, 100.0 * scene_extent, scene_center)
        # rows below the opacity floor are recycled within a budget, not pruned outright
        recycled = ((masks.keep == 0) & (masks.hard == 0)).to(torch.uint8)
        keep = (masks.hard == 0).to(torch.uint8)
        hard_count = int(masks.hard.sum())
        frac = float(min(max(np.float32(options.dense_recycle_fraction), np.float32(0.0)), np.float32(1.0)))
        recycle_limit = int(math.ceil(float(old) * frac))
        budget = recycle_limit - hard_count if recycle_limit > hard_count else 0
        if budget:
            select_rows(masks.opacities, budget, _abi.PHT_SELECT_SMALLEST, keep, 0, mask=recycled)
        # the most opaque row (lowest index on ties) survives unless hard-pruned
        best = torch.zeros_like(keep)
        select_rows(masks.opacities, 1, _abi.PHT_SELECT_LARGEST, best, 1)
        keep |= best & (masks.hard == 0).to(torch.uint8)
        retained = int(keep.sum())
        if retained > options.densification_cap:
            retained -= select_rows(masks.opacities, retained - options.densification_cap, _abi.PHT_SELECT_SMALLEST,
                                    keep, 0, mask=keep, exclude=best).selected
        keep_rows = None
        if retained != old:
            keep_rows = torch.nonzero(keep).squeeze(1).to(torch.int32)
        pruned = old - retained
        growth_cap = options.densification_cap if stats.growth_cap == 0 else \
            min(options.densification_cap, max(old, stats.growth_cap))
        capacity = growth_cap - retained if growth_cap > retained else 0
        remaining_progress = float(np.float32(1.0) - np.float32(iteration) / max(np.float32(1.0),
                                                                                np.float32(options.iterations)))
        progress = max(remaining_progress, 0.0)

        def decay():
            apply_adc_decay(model, float(np.float32(options.opacity_decay) * np.float32(progress)),
                            float(np.float32(options.scale_decay) * np.float32(progress)))

        if capacity == 0:
            if keep_rows is not None:
                remap_rows(model, states, keep_rows)
            decay()
            self.last = dict(pruned=pruned, retained=retained, grown=0, capacity=0)
            return RefinementCounts(0, pruned), make_densification_stats(model.size(), dev)

        s = stats.abi()
        gthr, sthr = float(options.densify_gradient_threshold), float(options.densify_screen_threshold)
        priority = torch.empty((retained,), dtype=torch.float32, device=dev)
        positive = torch.empty((retained,), dtype=torch.uint8, device=dev)
        counts = torch.zeros((3,), dtype=torch.int32, device=dev)
        check(lib.pht_igs_prepare(C.byref(s), _vp(keep_rows), retained, gthr, sthr, _vp(priority), _vp(positive),
                                  _vp(counts), C.c_void_p(_stream())))
        positive_count, growth_candidates, oversized_count = (int(v) for v in counts.cpu().tolist())
        allow_growth = iteration < grow_stop_iteration(options)
        if not allow_growth:
            growth_candidates = oversized_count = 0
        median = 1.0
        if positive_count:
            median = max(positive_upper_median(priority, positive, positive_count), 1e-9)
        replacement = torch.empty((retained,), dtype=torch.float32, device=dev)
        growth = torch.empty_like(replacement) if allow_growth else None
        oversize = torch.empty_like(replacement) if allow_growth else None
        check(lib.pht_igs_weights(C.byref(s), _vp(keep_rows), retained, _vp(masks.opacities), float(median),
                                  int(allow_growth), gthr, sthr, _vp(replacement), _vp(growth), _vp(oversize),
                                  C.c_void_p(_stream())))
        desired = int(round_half_away(growth_candidates * float(np.float32(options.densify_select_fraction)))) \
            if allow_growth else 0
        sel = select_igs_parents(replacement, oversize, growth, pruned, desired, capacity, random)
        grown = int(sel.parents.numel())
        self._grow(model, sel.parents, keep_rows, retained, old, options, random, states)
        decay()
        self.last = dict(pruned=pruned, retained=retained, grown=grown, capacity=capacity,
                         replacement=sel.replacement, oversized=sel.oversized, growth=sel.growth,
                         growth_candidates=growth_candidates, oversized_candidates=oversized_count,
                         positive=positive_count, median=median)
        return RefinementCounts(grown, pruned), make_densification_stats(model.size(), dev)

    @staticmethod
    def _grow(model: GaussianModel, parents: torch.Tensor, keep_rows: Optional[torch.Tensor], retained: int, old: int,
              options: DensifyOptions, random: Mt19937, states: AdamStates) -> None:
        """grow_igs_random_gpu (densification_adc_plus.cpp:97-146): the retained rows, then one child
        per parent; parents and children start with zero moments; igs_random split."""
        dev = model.means.device
        count = int(parents.numel())
        if count == 0:
            if keep_rows is not None:
                remap_rows(model, states, keep_rows)
            return
        samples = torch.from_numpy(normal_floats(random, 3 * count).reshape(count, 3)).to(dev)
        base = keep_rows if keep_rows is not None else torch.arange(old, dtype=torch.int32, device=dev)
        source = base[parents.long()]
        param_rows = torch.cat([base, source])
        adam_rows = param_rows.clone()
        adam_rows[parents.long()] = -1
        adam_rows[retained:] = -1
        remap_rows(model, states, param_rows, adam_rows)
        split_gaussians(model, parents, retained, samples, _abi.PHT_SPLIT_IGS_RANDOM, options.prune_opacity)


def round_half_away(x: float) -> float:
    """std::llround: nearest integer, halves away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def refine_gaussians(model: GaussianModel, stats: DensificationStats, iteration: int, scene_extent: float,
                     scene_center: Sequence[float], options: DensifyOptions, random: Mt19937,
                     states: AdamStates) -> Tuple[RefinementCounts, DensificationStats]:
    """densification::refine_gaussians for the adc_igs strategy (the trainer's default,
    options.hpp:155)."""
    return IgsStrategy().refine(model, stats, iteration, scene_extent, scene_center, options, random, states)
