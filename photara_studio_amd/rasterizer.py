"""Python mirror of the reference's rasterizer API over the gfx950 C ABI.

Names, argument meaning and error behaviour follow ``photara::splat::Rasterizer``
(photara/include/splat/rasterizer.hpp:118-211) and its types (photara/include/splat/types.hpp:17-111,
options.hpp:32-37): ``Rasterizer.forward(model, camera, options) -> RenderResult`` and
``Rasterizer.backward(model, rendered, grad_color, grad_alpha, grad_depth, grad_normal,
densify_map, sh_adam, structure_adam) -> ModelGradients``. Tensors are torch tensors on the GPU
(contiguous float32, the reference's shapes); invalid arguments raise ValueError (the reference's
std::invalid_argument), a backward without recorded forward state raises RuntimeError (its
std::logic_error), and the depth/normal channels, which this build does not produce, raise
NotImplementedError.

Beyond the reference: ``forward_views`` / ``backward_views`` render and differentiate several
cameras of different resolutions in one batched launch per stage.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import torch

from . import _abi
from ._abi import check


# ---------------------------------------------------------------- types (types.hpp / options.hpp)
@dataclass
class Camera:
    """photara::splat::Camera (types.hpp:17-34): column-major world_to_camera, pinhole model."""
    world_to_camera: Sequence[float] = field(default_factory=lambda: [0.0] * 16)
    position: Sequence[float] = (0.0, 0.0, 0.0)
    fx: float = 1.0
    fy: float = 1.0
    cx: float = 0.0
    cy: float = 0.0
    width: int = 0
    height: int = 0
    model: str = "pinhole"
    orthographic: bool = False


@dataclass
class GaussianModel:
    """photara::splat::GaussianModel (types.hpp:36-56): raw parameters."""
    means: torch.Tensor = None           # [N,3]
    log_scales: torch.Tensor = None      # [N,3]
    quaternions: torch.Tensor = None     # [N,4] scalar-first
    opacity_logits: torch.Tensor = None  # [N,1]
    sh: torch.Tensor = None              # [N,K,3]
    filter_3d: Optional[torch.Tensor] = None  # [N,1]
    sh_degree: int = 3
    # Reduced-precision storage read by the forward instead of the FP32 arrays (bf16/fp16).
    sh_lp: Optional[torch.Tensor] = None
    log_scales_lp: Optional[torch.Tensor] = None
    opacity_logits_lp: Optional[torch.Tensor] = None

    def size(self) -> int:
        return 0 if self.means is None or self.means.dim() == 0 else int(self.means.shape[0])


@dataclass
class RasterizationPolicy:
    """options.hpp:32-37."""
    alpha_clip: float = 0.99
    near_plane: float = 0.2
    pixel_center_offset: float = 0.0
    compensate_2d: bool = True


@dataclass
class RasterizeOptions:
    """rasterizer.hpp:9-54 (pinhole RGB training fields) plus this build's launch tuning."""
    rasterization: RasterizationPolicy = field(default_factory=RasterizationPolicy)
    active_sh_degree: int = 0
    background: Sequence[float] = (0.0, 0.0, 0.0)
    kernel_size: float = 0.0
    scale_modifier: float = 1.0
    # The reference renders median depth and normals by default; this build renders colour and
    # alpha only, so the default is off and True raises NotImplementedError.
    require_depth: bool = False
    record_backward_state: bool = True
    colors_precomp: Optional[torch.Tensor] = None  # [N,3]
    # gfx950 launch shape: tile edge (16 = reference, or 8), LDS batch depth, waves (one (view,
    # tile) item each) per blend workgroup, and whether the blends run as one persistent
    # work-queue launch (resident workgroups pull items from a device counter) or a flat grid.
    tile_size: int = 16
    batch: int = 64
    blend_waves: int = 1
    blend_queue: bool = False
    fp64_conic: bool = True


def configure_rasterization(options: RasterizeOptions, filter_mode: str = "classic") -> RasterizeOptions:
    """Training raster defaults (rasterizer.hpp:56-63): classic = 0.3 px dilation, no compensation."""
    mip2d = filter_mode in ("mip_2d", "mip")
    options.kernel_size = 0.1 if mip2d else 0.3
    options.rasterization.compensate_2d = mip2d
    return options


@dataclass
class RasterContext:
    impl: object = None


@dataclass
class RenderResult:
    """types.hpp:99-111."""
    color: torch.Tensor = None        # [3,H,W]
    alpha: torch.Tensor = None        # [H,W]
    median_depth: Optional[torch.Tensor] = None
    normal: Optional[torch.Tensor] = None
    radii: torch.Tensor = None        # [N] int32
    visibility: torch.Tensor = None   # [N] float
    context: RasterContext = field(default_factory=RasterContext)
    rendered_instances: int = 0


@dataclass
class ModelGradients:
    """types.hpp:71-87 (gradients wrt raw parameters)."""
    means: Optional[torch.Tensor] = None
    log_scales: Optional[torch.Tensor] = None
    quaternions: Optional[torch.Tensor] = None
    opacity_logits: Optional[torch.Tensor] = None
    sh: Optional[torch.Tensor] = None
    colors_precomp: Optional[torch.Tensor] = None
    refine_weight: Optional[torch.Tensor] = None


@dataclass
class StructureAdamUpdate:
    """StructureAdamUpdate (rasterizer.hpp:84-104) with FP32 SH moments added so that one fused
    pass updates every Gaussian parameter (the reference's quantised SH Adam is a separate row)."""
    means_first: torch.Tensor = None
    means_second: torch.Tensor = None
    scales_first: torch.Tensor = None
    scales_second: torch.Tensor = None
    rotations_first: torch.Tensor = None
    rotations_second: torch.Tensor = None
    opacity_first: torch.Tensor = None
    opacity_second: torch.Tensor = None
    sh_first: Optional[torch.Tensor] = None
    sh_second: Optional[torch.Tensor] = None
    means_lr: float = 1.6e-4
    scales_lr: float = 5e-3
    quaternions_lr: float = 1e-3
    opacities_lr: float = 5e-2
    sh_lr: float = 2.5e-3
    sh_rest_lr: float = 1.25e-4
    beta1: float = 0.9
    beta2: float = 0.999
    correction1: float = 1.0
    correction2: float = 1.0
    epsilon: float = 1e-15
    minimum_log_scale: float = -math.inf
    maximum_log_scale: float = math.inf

    @staticmethod
    def zeros_like(model: GaussianModel, sh_moments: bool = True, **kw) -> "StructureAdamUpdate":
        """Zero moments for every parameter; without ``sh_moments`` the SH coefficients are left
        to a quantised SH Adam (SHAdamUpdate)."""
        z = torch.zeros_like
        return StructureAdamUpdate(
            means_first=z(model.means), means_second=z(model.means),
            scales_first=z(model.log_scales), scales_second=z(model.log_scales),
            rotations_first=z(model.quaternions), rotations_second=z(model.quaternions),
            opacity_first=z(model.opacity_logits), opacity_second=z(model.opacity_logits),
            sh_first=z(model.sh) if sh_moments else None, sh_second=z(model.sh) if sh_moments else None, **kw)

    def set_step(self, step: int) -> None:
        """Bias corrections 1 - beta^t in FP32 on the host (cuda_ops.cu:2943-2944)."""
        import numpy as np
        self.correction1 = float(np.float32(1.0) - np.power(np.float32(self.beta1), np.float32(step)))
        self.correction2 = float(np.float32(1.0) - np.power(np.float32(self.beta2), np.float32(step)))


@dataclass
class ShAdamQuant:
    """The reference trainer's SH optimizer state (detail::ShAdamQuant,
    photara/src/splat/sh_adam_quant.hpp:12-24): ``first`` float16 [N, stride] normalised Adam
    update, ``packed`` uint8 [N, stride] log-second codes, ``bounds`` float32 [N, 4] (DC min, DC
    max, rest min, rest max), stride = SH bases * 3."""
    first: torch.Tensor = None
    packed: torch.Tensor = None
    bounds: torch.Tensor = None
    stride: int = 0

    @staticmethod
    def zeros(rows: int, stride: int, device) -> "ShAdamQuant":
        """make_sh_adam_quant (sh_adam_quant.cpp:9-24). This build takes strides up to 48 (SH degree 3)."""
        if stride <= 0 or stride > 48:
            raise ValueError("quantized SH Adam stride must be in [1, 48]")
        return ShAdamQuant(torch.zeros((rows, stride), dtype=torch.float16, device=device),
                           torch.zeros((rows, stride), dtype=torch.uint8, device=device),
                           torch.zeros((rows, 4), dtype=torch.float32, device=device), stride)

    def active(self) -> bool:
        return self.first is not None

    def select_rows(self, indices: torch.Tensor) -> None:
        """sh_adam_quant_select_rows: keep (and reorder / repeat) the given rows, bit for bit."""
        idx = indices.to(device=self.first.device, dtype=torch.long)
        self.first, self.packed, self.bounds = self.first[idx].contiguous(), self.packed[idx].contiguous(), \
            self.bounds[idx].contiguous()

    def zero_rows(self, indices: torch.Tensor) -> None:
        """sh_adam_quant_zero_rows: reset the given rows' state."""
        if indices is None or indices.numel() == 0:
            return
        idx = indices.to(device=self.first.device, dtype=torch.long)
        self.first[idx] = 0
        self.packed[idx] = 0
        self.bounds[idx] = 0

    def append_zeros(self, count: int) -> None:
        """sh_adam_quant_append_zeros: new rows start from a zero state."""
        if count <= 0:
            return
        z = ShAdamQuant.zeros(count, self.stride, self.first.device)
        self.first = torch.cat([self.first, z.first])
        self.packed = torch.cat([self.packed, z.packed])
        self.bounds = torch.cat([self.bounds, z.bounds])


@dataclass
class SHAdamUpdate:
    """photara::splat::SHAdamUpdate (rasterizer.hpp:73-82): the quantised SH Adam applied by the
    backward. ``first``, ``packed`` and ``bounds`` are an ShAdamQuant state's tensors."""
    first: torch.Tensor = None
    lr: float = 0.0
    rest_lr: float = 0.0
    beta1: float = 0.9
    beta2: float = 0.999
    correction1: float = 1.0
    correction2: float = 1.0
    epsilon: float = 1e-8
    regularization_weight: float = 0.0
    packed: torch.Tensor = None
    bounds: torch.Tensor = None

    @staticmethod
    def of(state: ShAdamQuant, **kw) -> "SHAdamUpdate":
        return SHAdamUpdate(first=state.first, packed=state.packed, bounds=state.bounds, **kw)

    def regularization_factor(self, count: int, stride: int) -> float:
        """2 w / (N (stride - 3)) in float32 (rasterizer.cu:598-602), 0 without non-DC columns."""
        import numpy as np
        if stride <= 3 or self.regularization_weight <= 0.0:
            return 0.0
        return float(np.float32(np.float32(2.0) * np.float32(self.regularization_weight)) / np.float32(count * (stride - 3)))


def sh_adam_quant_step(parameter: torch.Tensor, gradient: torch.Tensor, state: ShAdamQuant, active_stride: int,
                       learning_rate: float, rest_learning_rate: float, beta1: float, beta2: float,
                       correction1: float, correction2: float, adam_epsilon: float,
                       regularization_factor: float) -> None:
    """detail::sh_adam_quant_step (photara/src/splat/sh_adam_quant.hpp:34-49) on the gfx950 kernel:
    columns [0, active_stride) updated, the rest re-encoded under the new bounds."""
    rows = parameter.numel() // max(state.stride, 1)
    if parameter.numel() != rows * state.stride or gradient.numel() != parameter.numel():
        raise ValueError("parameter and gradient must hold rows x stride values")
    for t, name, dt in ((parameter, "parameter", torch.float32), (gradient, "gradient", torch.float32),
                        (state.first, "first", torch.float16), (state.packed, "packed", torch.uint8),
                        (state.bounds, "bounds", torch.float32)):
        _require(t, name, dt)
    if not 0 <= active_stride <= state.stride:
        raise ValueError("active_stride must lie in [0, stride]")
    check(_abi.load().pht_sh_adam_quant_step(
        C.c_void_p(parameter.data_ptr()), C.c_void_p(gradient.data_ptr()), C.c_void_p(state.first.data_ptr()),
        C.c_void_p(state.packed.data_ptr()), C.c_void_p(state.bounds.data_ptr()), rows, state.stride,
        int(active_stride), learning_rate, rest_learning_rate, beta1, beta2, correction1, correction2,
        adam_epsilon, regularization_factor, C.c_void_p(_stream())))


# ---------------------------------------------------------------- helpers
def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _require(t: torch.Tensor, name: str, dtype=torch.float32) -> None:
    if t is None or not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous CUDA {dtype} tensor")


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def lp_format_of(model: GaussianModel) -> int:
    if model.sh_lp is None and model.log_scales_lp is None and model.opacity_logits_lp is None:
        return _abi.PHT_LP_NONE
    dt = (model.sh_lp if model.sh_lp is not None else model.log_scales_lp).dtype
    return _abi.PHT_LP_BF16 if dt == torch.bfloat16 else _abi.PHT_LP_FP16


def validate_model(model: GaussianModel, options: RasterizeOptions) -> None:
    """validate_model (rasterizer.cu:306-329)."""
    n = model.size()
    for name in ("means", "log_scales", "quaternions", "opacity_logits", "sh"):
        _require(getattr(model, name), f"model.{name}")
    if model.filter_3d is not None:
        _require(model.filter_3d, "model.filter_3d")
    if (model.means.dim() != 2 or model.means.shape[1] != 3 or model.log_scales.shape != model.means.shape
            or model.quaternions.dim() != 2 or model.quaternions.shape != (n, 4)
            or model.opacity_logits.numel() != n or model.sh.dim() != 3 or model.sh.shape[0] != n
            or model.sh.shape[2] != 3
            or (model.filter_3d is not None and tuple(model.filter_3d.shape) != (n, 1))):
        raise ValueError("Invalid splat model tensor shapes")
    if options.colors_precomp is not None:
        _require(options.colors_precomp, "colors_precomp")
        if tuple(options.colors_precomp.shape) != (n, 3):
            raise ValueError("Precomputed colors must have shape [N,3]")
    for name in ("sh_lp", "log_scales_lp", "opacity_logits_lp"):
        t = getattr(model, name)
        if t is not None:
            if not t.is_cuda or t.dtype not in (torch.bfloat16, torch.float16) or not t.is_contiguous():
                raise ValueError(f"model.{name} must be a contiguous CUDA bf16/fp16 tensor")


def make_settings(model: GaussianModel, options: RasterizeOptions) -> _abi.pht_settings:
    deg = min(int(options.active_sh_degree), int(model.sh_degree))
    s = _abi.pht_settings()
    s.active_sh_degree = deg
    s.sh_bases = int(model.sh.shape[1])
    s.kernel_size = float(options.kernel_size)
    s.scale_modifier = float(options.scale_modifier)
    s.alpha_clip = float(options.rasterization.alpha_clip)
    s.near_plane = float(options.rasterization.near_plane)
    s.compensate_2d = 1 if options.rasterization.compensate_2d else 0
    s.tile_size = int(options.tile_size)
    s.batch = int(options.batch)
    s.blend_waves = int(options.blend_waves)
    s.blend_queue = 1 if options.blend_queue else 0
    s.fp64_conic = 1 if options.fp64_conic else 0
    s.require_depth = 1 if options.require_depth else 0
    return s


def make_camera(camera: Camera, options: RasterizeOptions) -> _abi.pht_camera:
    if camera.orthographic or camera.model != "pinhole":
        raise NotImplementedError("only pinhole cameras are built (fisheye/equirect/ortho are out of scope)")
    if int(camera.width) <= 0 or int(camera.height) <= 0:
        raise ValueError("Splat camera dimensions must be positive")
    c = _abi.pht_camera()
    for k in range(16):
        c.world_to_camera[k] = float(camera.world_to_camera[k])
    for k in range(3):
        c.position[k] = float(camera.position[k])
        c.background[k] = float(options.background[k])
    c.fx, c.fy = float(camera.fx), float(camera.fy)
    off = float(options.rasterization.pixel_center_offset)
    c.cx, c.cy = float(camera.cx) - off, float(camera.cy) - off
    c.width, c.height = int(camera.width), int(camera.height)
    return c


def make_gaussians(model: GaussianModel, options: RasterizeOptions) -> _abi.pht_gaussians:
    g = _abi.pht_gaussians()
    g.count = model.size()
    g.means = _ptr(model.means)
    g.log_scales = _ptr(model.log_scales)
    g.quaternions = _ptr(model.quaternions)
    g.opacity_logits = _ptr(model.opacity_logits)
    g.sh = _ptr(model.sh)
    g.filter_3d = _ptr(model.filter_3d)
    g.colors_precomp = _ptr(options.colors_precomp)
    g.lp_format = lp_format_of(model)
    g.sh_lp = _ptr(model.sh_lp)
    g.log_scales_lp = _ptr(model.log_scales_lp)
    g.opacity_logits_lp = _ptr(model.opacity_logits_lp)
    return g


class Workspace:
    """Caller-owned device bytes (a torch uint8 tensor) carved by the C ABI."""

    def __init__(self, device: torch.device, gaussians: int, views: int, tile_size: int, tiles: int, pixels: int,
                 instances: int):
        lib = _abi.load()
        self.cap = _abi.pht_capacity(gaussians, views, tile_size, tiles, pixels, instances)
        nbytes = lib.pht_workspace_bytes(C.byref(self.cap))
        if nbytes == 0:
            check(_abi.PHT_EINVAL)
        self.buf = torch.empty(nbytes + 256, dtype=torch.uint8, device=device)
        base = (self.buf.data_ptr() + 255) & ~255
        handle = C.c_void_p()
        check(lib.pht_workspace_create(C.c_void_p(base), C.c_size_t(nbytes), C.byref(self.cap), C.byref(handle)))
        self.handle = handle
        self.busy = False
        self.generation = 0

    def fits(self, gaussians: int, views: int, tile_size: int, tiles: int, pixels: int, instances: int) -> bool:
        c = self.cap
        return (c.gaussians >= gaussians and c.tile_size == tile_size and c.tiles >= tiles and c.pixels >= pixels
                and c.instances >= instances and c.views >= views
                and (1 << (max(views, 1) - 1).bit_length()) == (1 << (max(c.views, 1) - 1).bit_length()))

    def counts(self) -> _abi.pht_counts:
        out = _abi.pht_counts()
        check(_abi.load().pht_get_counts(self.handle, C.byref(out), C.c_void_p(_stream())))
        return out

    STAGES = ("preprocess", "compact+depth_sort", "tile_offsets+emit", "tile_sort+ranges", "blend_forward",
              "blend_backward", "gaussian_backward+adam")

    def set_profiling(self, enable: bool) -> None:
        check(_abi.load().pht_set_profiling(self.handle, 1 if enable else 0))

    def stage_times(self) -> dict:
        """Milliseconds of each stage of the last forward/backward recorded with profiling on."""
        buf = (C.c_float * len(self.STAGES))()
        n = _abi.load().pht_stage_times(self.handle, buf, len(self.STAGES))
        if n < 0:
            check(n)
        return {name: float(buf[k]) for k, name in enumerate(self.STAGES) if buf[k] >= 0.0}

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _abi.load().pht_workspace_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class _Context:
    """Forward state a backward needs (RasterContextImpl, rasterizer.cu:33-50)."""

    def __init__(self, ws: Workspace, pool: "Rasterizer", n_views: int, settings, gaussians, options):
        self.ws, self.pool, self.n_views = ws, pool, n_views
        self.settings, self.gaussians, self.options = settings, gaussians, options
        self.generation = ws.generation
        self.backward_consumed = False

    def live(self) -> bool:
        return self.ws.generation == self.generation

    def release(self) -> None:
        if self.live():
            self.ws.busy = False

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def default_instance_capacity(n: int, views: int) -> int:
    return max(1 << 20, 8 * n * views)


def tiles_of(cameras: Sequence[Camera], ts: int) -> int:
    return sum(((int(c.width) + ts - 1) // ts) * ((int(c.height) + ts - 1) // ts) for c in cameras)


class Rasterizer:
    """photara::splat::Rasterizer (rasterizer.hpp:118-211), pinhole RGB training path."""

    def __init__(self):
        self._pool: List[Workspace] = []

    # -- workspace pool: a workspace stays reserved while a RenderResult holds its context
    def _workspace(self, device, n, views, ts, tiles, pixels, instances) -> Workspace:
        for ws in self._pool:
            if not ws.busy and ws.fits(n, views, ts, tiles, pixels, instances):
                return ws
        self._pool = [w for w in self._pool if w.busy]
        ws = Workspace(device, n, views, ts, tiles, pixels, instances)
        self._pool.append(ws)
        return ws

    def forward(self, model: GaussianModel, camera: Camera, options: Optional[RasterizeOptions] = None) -> RenderResult:
        return self.forward_views(model, [camera], options)[0]

    def forward_views(self, model: GaussianModel, cameras: Sequence[Camera],
                      options: Optional[RasterizeOptions] = None, instance_capacity: Optional[int] = None
                      ) -> List[RenderResult]:
        """One batched forward over several pinhole cameras (any resolutions)."""
        options = options or RasterizeOptions()
        if options.require_depth:
            raise NotImplementedError("median-depth / normal channels are not built; set require_depth=False")
        validate_model(model, options)
        if not cameras or len(cameras) > _abi.PHT_MAX_VIEWS:
            raise ValueError("between 1 and 32 cameras per batched forward")
        n = model.size()
        if n == 0:
            raise ValueError("empty model")
        device = model.means.device
        cams = (_abi.pht_camera * len(cameras))(*[make_camera(c, options) for c in cameras])
        settings = make_settings(model, options)
        g = make_gaussians(model, options)
        ts = int(options.tile_size)
        tiles = tiles_of(cameras, ts)
        pixels = sum(int(c.width) * int(c.height) for c in cameras)
        results = []
        outs = (_abi.pht_view_outputs * len(cameras))()
        for k, c in enumerate(cameras):
            r = RenderResult()
            r.color = torch.empty((3, int(c.height), int(c.width)), dtype=torch.float32, device=device)
            r.alpha = torch.empty((int(c.height), int(c.width)), dtype=torch.float32, device=device)
            r.radii = torch.empty((n,), dtype=torch.int32, device=device)
            r.visibility = torch.empty((n,), dtype=torch.float32, device=device)
            outs[k].color, outs[k].alpha = r.color.data_ptr(), r.alpha.data_ptr()
            outs[k].radii, outs[k].visibility = r.radii.data_ptr(), r.visibility.data_ptr()
            results.append(r)
        cap = instance_capacity or default_instance_capacity(n, len(cameras))
        lib = _abi.load()
        while True:
            ws = self._workspace(device, n, len(cameras), ts, tiles, pixels, cap)
            rc = lib.pht_forward(ws.handle, C.byref(g), cams, len(cameras), C.byref(settings), outs, 1,
                                 C.c_void_p(_stream()))
            if rc == _abi.PHT_ECAPACITY:
                cap = int(ws.counts().instances * 1.25) + 1024
                continue
            check(rc)
            break
        ws.busy = True
        ws.generation += 1
        counts = ws.counts()
        ctx = _Context(ws, self, len(cameras), settings, g, options)
        ctx.model_ptrs = (model.means.data_ptr(), model.log_scales.data_ptr(), model.quaternions.data_ptr())
        ctx.counts = counts
        ctx.tiles, ctx.pixels = tiles, pixels
        for r in results:
            r.context = RasterContext(ctx)
            r.rendered_instances = int(counts.instances)
        if not options.record_backward_state:
            ctx.inference_only = True
        return results

    def backward(self, model: GaussianModel, rendered: RenderResult, grad_color: torch.Tensor,
                 grad_alpha: Optional[torch.Tensor] = None, grad_depth: Optional[torch.Tensor] = None,
                 grad_normal: Optional[torch.Tensor] = None, densify_map: Optional[torch.Tensor] = None,
                 sh_adam=None, structure_adam: Optional[StructureAdamUpdate] = None) -> ModelGradients:
        if densify_map is not None and densify_map.numel() != 0:
            raise NotImplementedError("error-map densification statistics are not built")
        return self.backward_views(model, [rendered], [grad_color], [grad_alpha], structure_adam, sh_adam)

    def backward_views(self, model: GaussianModel, rendered: Sequence[RenderResult],
                       grad_colors: Sequence[torch.Tensor], grad_alphas: Optional[Sequence] = None,
                       structure_adam: Optional[StructureAdamUpdate] = None,
                       sh_adam: Optional[SHAdamUpdate] = None) -> ModelGradients:
        """With ``sh_adam`` the SH coefficients take the quantised SH Adam step: fused into the
        per-Gaussian backward when ``structure_adam`` is given too (the reference's fused path,
        render_backward.cu:833-853), else as a step on the returned SH gradient."""
        if not rendered or rendered[0].context is None or rendered[0].context.impl is None:
            raise ValueError("Splat backward requires a live forward context")
        ctx: _Context = rendered[0].context.impl
        if getattr(ctx, "inference_only", False):
            raise RuntimeError("Cannot backpropagate an inference-only render")
        if any(r.context.impl is not ctx for r in rendered) or len(rendered) != ctx.n_views:
            raise ValueError("backward_views needs every view of one batched forward, in order")
        if not ctx.live():
            raise RuntimeError("the forward state of this render was overwritten by a later forward")
        n = model.size()
        if n != ctx.gaussians.count:
            raise ValueError("model size differs from the forward's")
        grad_alphas = list(grad_alphas) if grad_alphas is not None else [None] * len(rendered)
        dl = (_abi.pht_view_grads * len(rendered))()
        keep = []
        for k, (r, gc, ga) in enumerate(zip(rendered, grad_colors, grad_alphas)):
            _require(gc, "grad_color")
            if tuple(gc.shape) != tuple(r.color.shape):
                raise ValueError("grad_color must match the rendered colour shape")
            dl[k].grad_color = gc.data_ptr()
            if ga is not None:
                _require(ga, "grad_alpha")
                if tuple(ga.shape) != tuple(r.alpha.shape):
                    raise ValueError("grad_alpha must match the rendered alpha shape")
                dl[k].grad_alpha = ga.data_ptr()
            keep.append((gc, ga))
        device = model.means.device
        out = ModelGradients()
        mg = _abi.pht_model_grads()
        out.refine_weight = torch.empty((n,), dtype=torch.float32, device=device)
        mg.refine_weight = out.refine_weight.data_ptr()
        adam_ptr = None
        if structure_adam is None:
            out.means = torch.empty_like(model.means)
            out.log_scales = torch.empty_like(model.log_scales)
            out.quaternions = torch.empty_like(model.quaternions)
            out.opacity_logits = torch.empty((n, 1), dtype=torch.float32, device=device)
            mg.means, mg.log_scales = out.means.data_ptr(), out.log_scales.data_ptr()
            mg.quaternions, mg.opacity_logits = out.quaternions.data_ptr(), out.opacity_logits.data_ptr()
            if ctx.options.colors_precomp is not None:
                out.colors_precomp = torch.empty((n, 3), dtype=torch.float32, device=device)
                mg.colors_precomp = out.colors_precomp.data_ptr()
            else:
                out.sh = torch.empty_like(model.sh)
                mg.sh = out.sh.data_ptr()
        else:
            if sh_adam is not None and (sh_adam.beta1, sh_adam.beta2, sh_adam.correction1, sh_adam.correction2,
                                        sh_adam.epsilon) != (structure_adam.beta1, structure_adam.beta2,
                                                             structure_adam.correction1, structure_adam.correction2,
                                                             structure_adam.epsilon):
                raise ValueError("the fused step needs one beta1/beta2/correction/epsilon for SH and structure")
            adam_ptr = C.byref(make_adam(model, structure_adam, sh_adam))
        if sh_adam is not None and ctx.options.colors_precomp is not None:
            raise ValueError("sh_adam needs SH colours (no colors_precomp)")
        lib = _abi.load()
        check(lib.pht_backward(ctx.ws.handle, C.byref(ctx.gaussians), dl, C.byref(mg), adam_ptr,
                               C.c_void_p(_stream())))
        ctx.backward_consumed = True
        if sh_adam is not None and structure_adam is None:
            stride = int(model.sh.shape[1]) * 3
            active = (int(ctx.settings.active_sh_degree) + 1) ** 2 * 3
            state = ShAdamQuant(sh_adam.first, sh_adam.packed, sh_adam.bounds, stride)
            sh_adam_quant_step(model.sh, out.sh, state, active, sh_adam.lr, sh_adam.rest_lr, sh_adam.beta1,
                               sh_adam.beta2, sh_adam.correction1, sh_adam.correction2, sh_adam.epsilon,
                               sh_adam.regularization_factor(n, stride))
            out.sh = None  # consumed by the optimizer, as in the reference's fused SH update
        if structure_adam is not None or sh_adam is not None:
            # parameters changed in place: the recorded forward no longer matches the model
            ctx.release()
            ctx.generation = -1
        return out


def debug_lists(rendered: RenderResult) -> dict:
    """Host copies of the last forward's sorted instance list, as 64-bit keys
    ((global_tile << 32) | depth bits) and gaussian indices, plus per-tile ranges and the per-pixel
    contributor counts / final transmittance of every view (concatenated)."""
    ctx: _Context = rendered.context.impl
    if ctx is None or not ctx.live():
        raise RuntimeError("no live forward state")
    dev = ctx.ws.buf.device
    counts = ctx.ws.counts()
    ni = int(counts.instances)
    keys = torch.empty((max(ni, 1),), dtype=torch.int64, device=dev)
    vals = torch.empty((max(ni, 1),), dtype=torch.int32, device=dev)
    ranges = torch.empty((2 * ctx.tiles,), dtype=torch.int32, device=dev)
    n_contrib = torch.empty((ctx.pixels,), dtype=torch.int32, device=dev)
    final_T = torch.empty((ctx.pixels,), dtype=torch.float32, device=dev)
    check(_abi.load().pht_debug_lists(ctx.ws.handle, C.c_void_p(keys.data_ptr()), C.c_void_p(vals.data_ptr()),
                                      C.c_void_p(ranges.data_ptr()), C.c_void_p(n_contrib.data_ptr()),
                                      C.c_void_p(final_T.data_ptr()), C.c_void_p(_stream())))
    torch.cuda.synchronize()
    import numpy as np
    return {
        "keys": keys.cpu().numpy().view(np.uint64)[:ni],
        "values": vals.cpu().numpy().view(np.uint32)[:ni],
        "ranges": ranges.cpu().numpy().view(np.uint32),
        "n_contrib": n_contrib.cpu().numpy().view(np.uint32),
        "final_T": final_T.cpu().numpy(),
        "counts": counts,
    }


def debug_splat_grads(rendered: RenderResult):
    """Per-(gaussian, view) blend gradients of the last backward, numpy [N, views, 10]
    (d_mean2d.xy, d_conic.xyz, d_opacity, d_color.rgb, refine)."""
    ctx: _Context = rendered.context.impl
    n = int(ctx.gaussians.count)
    out = torch.empty((n, ctx.n_views, 10), dtype=torch.float32, device=ctx.ws.buf.device)
    check(_abi.load().pht_debug_splat_grads(ctx.ws.handle, C.c_void_p(out.data_ptr()), C.c_void_p(_stream())))
    return out.cpu().numpy()


def make_adam(model: GaussianModel, a: StructureAdamUpdate, sh: Optional[SHAdamUpdate] = None) -> _abi.pht_adam:
    s = _abi.pht_adam()
    s.means, s.log_scales = model.means.data_ptr(), model.log_scales.data_ptr()
    s.quaternions, s.opacity_logits = model.quaternions.data_ptr(), model.opacity_logits.data_ptr()
    s.m_means, s.v_means = a.means_first.data_ptr(), a.means_second.data_ptr()
    s.m_log_scales, s.v_log_scales = a.scales_first.data_ptr(), a.scales_second.data_ptr()
    s.m_quaternions, s.v_quaternions = a.rotations_first.data_ptr(), a.rotations_second.data_ptr()
    s.m_opacity_logits, s.v_opacity_logits = a.opacity_first.data_ptr(), a.opacity_second.data_ptr()
    if a.sh_first is not None:
        s.sh, s.m_sh, s.v_sh = model.sh.data_ptr(), a.sh_first.data_ptr(), a.sh_second.data_ptr()
    s.lr_means, s.lr_scales, s.lr_quaternions = a.means_lr, a.scales_lr, a.quaternions_lr
    s.lr_opacity, s.lr_sh0, s.lr_sh_rest = a.opacities_lr, a.sh_lr, a.sh_rest_lr
    s.beta1, s.beta2, s.correction1, s.correction2, s.epsilon = a.beta1, a.beta2, a.correction1, a.correction2, a.epsilon
    s.min_log_scale, s.max_log_scale = a.minimum_log_scale, a.maximum_log_scale
    s.lp_format = lp_format_of(model)
    s.sh_lp = _ptr(model.sh_lp)
    s.log_scales_lp = _ptr(model.log_scales_lp)
    s.opacity_logits_lp = _ptr(model.opacity_logits_lp)
    if sh is not None:
        stride = int(model.sh.shape[1]) * 3
        for t, name, dt in ((sh.first, "sh_adam.first", torch.float16), (sh.packed, "sh_adam.packed", torch.uint8),
                            (sh.bounds, "sh_adam.bounds", torch.float32)):
            _require(t, name, dt)
        if tuple(sh.first.shape) != (model.size(), stride) or tuple(sh.packed.shape) != (model.size(), stride) \
                or tuple(sh.bounds.shape) != (model.size(), 4):
            raise ValueError("sh_adam state must be [N, 3K] (first, packed) and [N, 4] (bounds)")
        s.sh = model.sh.data_ptr()
        s.m_sh = s.v_sh = None
        s.sh_first, s.sh_packed, s.sh_bounds = sh.first.data_ptr(), sh.packed.data_ptr(), sh.bounds.data_ptr()
        s.lr_sh0, s.lr_sh_rest = sh.lr, sh.rest_lr
        s.sh_regularization = sh.regularization_factor(model.size(), stride)
    return s
