"""ctypes binding of the CPU oracle (``oracle/_build/liboracle.so``) — TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker or the
timed CPU baseline. The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Dict, List, Optional, Sequence

import numpy as np

from photara_studio_amd import _abi
from photara_studio_amd.rasterizer import Camera, RasterizeOptions, make_camera

LIB = Path(__file__).resolve().parent / "_build" / "liboracle.so"
_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            from oracle import build as _b
            _b.build(verbose=False)
        lib = C.CDLL(str(LIB))
        vp = C.c_void_p
        lib.oracle_pexpf.argtypes = [vp, vp, C.c_int64]
        lib.oracle_l1_ssim.argtypes = [C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_float, C.c_float, vp]
        lib.oracle_l1_ssim.restype = C.c_double
        lib.oracle_pexp2f_blend.argtypes = [vp, vp, C.c_int64]
        lib.oracle_plog1pf.argtypes = [vp, vp, C.c_int64]
        lib.oracle_pexpm1f.argtypes = [vp, vp, C.c_int64]
        lib.oracle_sh_adam_quant_step.argtypes = [vp, vp, vp, vp, vp, C.c_int64, C.c_int, C.c_int, C.c_float,
                                                  C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                                  C.c_float]
        lib.oracle_plogf.argtypes = [vp, vp, C.c_int64]
        lib.oracle_pair_power2.argtypes = [vp, vp, vp, C.c_int64]
        lib.oracle_lp_roundtrip.argtypes = [vp, vp, vp, C.c_int64, C.c_int]
        lib.oracle_project.argtypes = [C.c_int, C.POINTER(_abi.pht_gaussians), C.POINTER(_abi.pht_camera),
                                       C.POINTER(_abi.pht_settings), vp]
        lib.oracle_enumerate.argtypes = [C.c_float, C.c_float, vp, C.c_float, C.c_int, C.c_int, C.c_int, vp, C.c_int]
        lib.oracle_enumerate.restype = C.c_int
        lib.oracle_forward.argtypes = [C.c_int, C.POINTER(_abi.pht_gaussians), C.POINTER(_abi.pht_camera), C.c_int,
                                       C.POINTER(_abi.pht_settings), C.POINTER(_abi.pht_view_outputs), vp, vp,
                                       C.c_int64, vp, vp, vp, C.POINTER(_abi.pht_counts), C.c_int]
        lib.oracle_backward.argtypes = [C.c_int, C.POINTER(_abi.pht_gaussians), C.POINTER(_abi.pht_settings),
                                        C.POINTER(_abi.pht_view_grads), vp, vp, vp, vp, vp, vp, vp, C.c_int]
        lib.oracle_l1_loss.argtypes = [vp, vp, vp, C.c_int64, C.c_float]
        lib.oracle_l1_loss.restype = C.c_double
        lib.oracle_adam_step.argtypes = [vp, vp, vp, vp, C.c_int64, C.c_int64, C.c_int64, C.c_float, C.c_float,
                                         C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                         C.c_float, C.c_float, C.c_int]
        lib.oracle_train_step.argtypes = [C.POINTER(_abi.pht_gaussians), C.POINTER(_abi.pht_camera), C.c_int,
                                          C.POINTER(_abi.pht_settings), C.POINTER(_abi.pht_view_outputs),
                                          C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(_abi.pht_adam),
                                          C.c_int, C.c_int, C.c_float, C.c_float]
        lib.oracle_train_step.restype = C.c_double
        lib.oracle_splat_grads.argtypes = [C.c_int, C.c_int64, vp]
        lib.oracle_set_inputs64.argtypes = [vp, vp, vp, vp, vp]
        lib.oracle_images.argtypes = [C.c_int, vp, vp]
        lib.oracle_sh_eval.argtypes = [C.c_int, vp, vp, vp, C.c_int64]
        lib.oracle_bruteforce.argtypes = [vp, vp, vp]
        lib.oracle_densify_accumulate.argtypes = [C.POINTER(_abi.pht_densify_stats),
                                                  C.POINTER(_abi.pht_densify_observation), C.c_int, C.c_int64]
        lib.oracle_prune_masks.argtypes = [C.POINTER(_abi.pht_gaussians), C.c_int, C.c_float, C.c_float, vp, vp, vp,
                                           vp]
        lib.oracle_select_keys.argtypes = [vp, vp, vp, C.c_int64, C.c_int, C.c_uint64, vp]
        lib.oracle_densify_split.argtypes = [C.POINTER(_abi.pht_split_params), vp, C.c_int64, C.c_int64, vp, vp]
        lib.oracle_adc_decay.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int64, C.c_float, C.c_float]
        lib.oracle_igs_rows.argtypes = [C.POINTER(_abi.pht_densify_stats), vp, C.c_int64, vp, C.c_float, C.c_int,
                                        C.c_float, C.c_float, vp, vp, vp, vp, vp, vp]
        _lib = lib
    return _lib


def _p(a: Optional[np.ndarray]):
    return None if a is None else C.c_void_p(a.ctypes.data)


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def threads_default() -> int:
    return int(os.environ.get("PHT_ORACLE_THREADS", "0")) or (os.cpu_count() or 1)


class Scene:
    """Host copy of one Gaussian set plus the C structs the oracle reads."""

    def __init__(self, arrays: Dict[str, np.ndarray], sh_degree: int = 3, colors_precomp=None, filter_3d=None,
                 lp: Optional[str] = None):
        self.a = {k: f32(v) for k, v in arrays.items()}
        self.n = int(self.a["means"].shape[0])
        self.sh_bases = int(self.a["sh"].shape[1])
        self.sh_degree = sh_degree
        self.colors_precomp = None if colors_precomp is None else f32(colors_precomp)
        self.filter_3d = None if filter_3d is None else f32(filter_3d).reshape(-1)
        self.lp = lp
        self.lp_arrays = {}
        if lp is not None:
            import torch
            dt = torch.bfloat16 if lp == "bf16" else torch.float16
            for k in ("sh", "log_scales", "opacity_logits"):
                t = torch.from_numpy(self.a[k]).to(dt)
                self.lp_arrays[k] = np.ascontiguousarray(t.view(torch.int16).numpy().view(np.uint16))

    def gaussians(self) -> _abi.pht_gaussians:
        g = _abi.pht_gaussians()
        g.count = self.n
        g.means = self.a["means"].ctypes.data
        g.log_scales = self.a["log_scales"].ctypes.data
        g.quaternions = self.a["quaternions"].ctypes.data
        g.opacity_logits = self.a["opacity_logits"].ctypes.data
        g.sh = self.a["sh"].ctypes.data
        g.filter_3d = None if self.filter_3d is None else self.filter_3d.ctypes.data
        g.colors_precomp = None if self.colors_precomp is None else self.colors_precomp.ctypes.data
        if self.lp is not None:
            g.lp_format = _abi.PHT_LP_BF16 if self.lp == "bf16" else _abi.PHT_LP_FP16
            g.sh_lp = self.lp_arrays["sh"].ctypes.data
            g.log_scales_lp = self.lp_arrays["log_scales"].ctypes.data
            g.opacity_logits_lp = self.lp_arrays["opacity_logits"].ctypes.data
        return g


def settings_for(scene: Scene, options: RasterizeOptions) -> _abi.pht_settings:
    s = _abi.pht_settings()
    s.active_sh_degree = min(int(options.active_sh_degree), scene.sh_degree)
    s.sh_bases = scene.sh_bases
    s.kernel_size = float(options.kernel_size)
    s.scale_modifier = float(options.scale_modifier)
    s.alpha_clip = float(options.rasterization.alpha_clip)
    s.near_plane = float(options.rasterization.near_plane)
    s.compensate_2d = 1 if options.rasterization.compensate_2d else 0
    s.tile_size = int(options.tile_size)
    s.batch = int(options.batch)
    s.fp64_conic = 1 if options.fp64_conic else 0
    return s


class OracleRun:
    """One oracle forward (and optional backward) over a list of cameras. precision: 32 = float
    parity oracle, 64 = double accuracy oracle, 65 = double values on the float run's discrete
    decisions (tile lists, 1/255 skips, early termination, SH clamps), 66 = double values and
    double skip / termination decisions on the float run's tile lists and blend order (images
    only)."""

    def __init__(self, scene: Scene, cameras: Sequence[Camera], options: RasterizeOptions, precision: int = 32,
                 threads: Optional[int] = None, key_cap: int = 1 << 26):
        self.scene, self.cameras, self.options, self.precision = scene, list(cameras), options, precision
        self.threads = threads or threads_default()
        lib = load()
        nv = len(self.cameras)
        self.cams = (_abi.pht_camera * nv)(*[make_camera(c, options) for c in self.cameras])
        self.settings = settings_for(scene, options)
        self.g = scene.gaussians()
        n = scene.n
        self.color = [np.zeros((3, c.height, c.width), np.float32) for c in self.cameras]
        self.alpha = [np.zeros((c.height, c.width), np.float32) for c in self.cameras]
        self.radii = [np.zeros((n,), np.int32) for _ in self.cameras]
        self.visibility = [np.zeros((n,), np.float32) for _ in self.cameras]
        outs = (_abi.pht_view_outputs * nv)()
        for k in range(nv):
            outs[k].color, outs[k].alpha = self.color[k].ctypes.data, self.alpha[k].ctypes.data
            outs[k].radii, outs[k].visibility = self.radii[k].ctypes.data, self.visibility[k].ctypes.data
        ts = options.tile_size
        self.tiles = sum(((c.width + ts - 1) // ts) * ((c.height + ts - 1) // ts) for c in self.cameras)
        self.pixels = sum(c.width * c.height for c in self.cameras)
        counts = _abi.pht_counts()
        if precision == 32:
            keys = np.zeros((key_cap,), np.uint64)
            values = np.zeros((key_cap,), np.uint32)
            ranges = np.zeros((2 * self.tiles,), np.uint32)
            n_contrib = np.zeros((self.pixels,), np.uint32)
            final_T = np.zeros((self.pixels,), np.float32)
            lib.oracle_forward(32, C.byref(self.g), self.cams, nv, C.byref(self.settings), outs, _p(keys), _p(values),
                               key_cap, _p(ranges), _p(n_contrib), _p(final_T), C.byref(counts), self.threads)
            ni = int(counts.instances)
            self.keys, self.values = keys[:ni].copy(), values[:ni].copy()
            self.ranges, self.n_contrib, self.final_T = ranges, n_contrib, final_T
        else:
            lib.oracle_forward(precision, C.byref(self.g), self.cams, nv, C.byref(self.settings), outs, None, None, 0,
                               None, None, None, C.byref(counts), self.threads)
        self.counts = counts
        self._outs = outs
        # images at the run's own precision (double for the accuracy oracle)
        c = np.zeros((3 * self.pixels,), np.float64)
        a = np.zeros((self.pixels,), np.float64)
        lib.oracle_images(precision, _p(c), _p(a))
        self.color_hp, self.alpha_hp = [], []
        off = 0
        for cam in self.cameras:
            P = cam.width * cam.height
            self.color_hp.append(c[3 * off:3 * (off + P)].reshape(3, cam.height, cam.width))
            self.alpha_hp.append(a[off:off + P].reshape(cam.height, cam.width))
            off += P

    def backward(self, grad_colors: Sequence[np.ndarray], grad_alphas: Optional[Sequence] = None) -> Dict[str, np.ndarray]:
        if self.precision == 66:
            raise ValueError("the order-matched FP64 oracle (precision 66) renders images only")
        lib = load()
        nv = len(self.cameras)
        gcs = [f32(g) for g in grad_colors]
        gas = [None if (grad_alphas is None or grad_alphas[k] is None) else f32(grad_alphas[k]) for k in range(nv)]
        dl = (_abi.pht_view_grads * nv)()
        for k in range(nv):
            dl[k].grad_color = gcs[k].ctypes.data
            dl[k].grad_alpha = None if gas[k] is None else gas[k].ctypes.data
        n, s = self.scene.n, self.scene
        dt = np.float64 if self.precision in (64, 65) else np.float32
        out = {
            "means": np.zeros((n, 3), dt), "log_scales": np.zeros((n, 3), dt), "quaternions": np.zeros((n, 4), dt),
            "opacity_logits": np.zeros((n, 1), dt), "sh": np.zeros((n, s.sh_bases, 3), dt),
            "colors_precomp": np.zeros((n, 3), dt), "refine_weight": np.zeros((n,), dt),
        }
        lib.oracle_backward(self.precision, C.byref(self.g), C.byref(self.settings), dl, _p(out["means"]),
                            _p(out["log_scales"]), _p(out["quaternions"]), _p(out["opacity_logits"]),
                            _p(out["sh"]) if s.colors_precomp is None else None,
                            _p(out["colors_precomp"]) if s.colors_precomp is not None else None,
                            _p(out["refine_weight"]), self.threads)
        self._keep = (gcs, gas)
        return out

    def bruteforce(self):
        """Tile-free render of this (float) run: per-view colour, alpha, final T."""
        c = np.zeros((3 * self.pixels,), np.float32)
        a = np.zeros((self.pixels,), np.float32)
        t = np.zeros((self.pixels,), np.float32)
        load().oracle_bruteforce(_p(c), _p(a), _p(t))
        out, off = [], 0
        for cam in self.cameras:
            P = cam.width * cam.height
            out.append((c[3 * off:3 * (off + P)].reshape(3, cam.height, cam.width),
                        a[off:off + P].reshape(cam.height, cam.width), t[off:off + P]))
            off += P
        return out

    def splat_grads(self) -> np.ndarray:
        """Per-(gaussian, view) blend gradients of the last backward, [N, views, 10]."""
        out = np.zeros((self.scene.n, len(self.cameras), 10), np.float64)
        load().oracle_splat_grads(self.precision, self.scene.n, _p(out))
        return out


class inputs64:
    """Context manager: precision-64 runs read these double parameters instead of the float ones."""

    def __init__(self, arrays: Dict[str, np.ndarray]):
        self.a = {k: np.ascontiguousarray(np.asarray(arrays[k], dtype=np.float64))
                  for k in ("means", "log_scales", "quaternions", "opacity_logits", "sh")}

    def __enter__(self):
        a = self.a
        load().oracle_set_inputs64(_p(a["means"]), _p(a["log_scales"]), _p(a["quaternions"]),
                                   _p(a["opacity_logits"]), _p(a["sh"]))
        return self

    def __exit__(self, *exc):
        load().oracle_set_inputs64(None, None, None, None, None)
        return False


def sh_eval(degree: int, dirs: np.ndarray, coef: np.ndarray) -> np.ndarray:
    """Unclamped SH colour (double) of unit directions [n,3] for one coefficient row [16,3]."""
    d = np.ascontiguousarray(dirs, dtype=np.float64)
    c = np.ascontiguousarray(np.asarray(coef, dtype=np.float64).reshape(-1))
    out = np.zeros((d.shape[0], 3), np.float64)
    load().oracle_sh_eval(degree, _p(d), _p(c), _p(out), d.shape[0])
    return out


def pexpf(x: np.ndarray) -> np.ndarray:
    x = f32(x)
    y = np.empty_like(x)
    load().oracle_pexpf(_p(x), _p(y), x.size)
    return y


def pexp2f_blend(x: np.ndarray) -> np.ndarray:
    x = f32(x)
    y = np.empty_like(x)
    load().oracle_pexp2f_blend(_p(x), _p(y), x.size)
    return y


def l1_ssim(color: np.ndarray, target: np.ndarray, mask: Optional[np.ndarray] = None, ssim_weight: float = 0.2,
            photometric_weight: float = 1.0, precision: int = 32):
    """Fused L1 + SSIM of one [3,H,W] view: (loss, gradient wrt color) at the given precision."""
    color, target = f32(color), f32(target)
    m = None if mask is None else f32(mask)
    _, H, W = color.shape
    grad = np.zeros(color.shape, np.float64 if precision == 64 else np.float32)
    loss = load().oracle_l1_ssim(precision, _p(color), _p(target), None if m is None else _p(m), W, H,
                                 ssim_weight, photometric_weight, _p(grad))
    return float(loss), grad


def plogf(x: np.ndarray) -> np.ndarray:
    x = f32(x)
    y = np.empty_like(x)
    load().oracle_plogf(_p(x), _p(y), x.size)
    return y


def plog1pf(x: np.ndarray) -> np.ndarray:
    x = f32(x)
    y = np.empty_like(x)
    load().oracle_plog1pf(_p(x), _p(y), x.size)
    return y


def pexpm1f(x: np.ndarray) -> np.ndarray:
    x = f32(x)
    y = np.empty_like(x)
    load().oracle_pexpm1f(_p(x), _p(y), x.size)
    return y


class ShQuantState:
    """Quantised SH Adam state on the host: first (fp16 bits as uint16) and packed codes
    [rows, stride], bounds float32 [rows, 4]."""

    def __init__(self, rows: int, stride: int):
        self.stride = stride
        self.first = np.zeros((rows, stride), np.uint16)
        self.packed = np.zeros((rows, stride), np.uint8)
        self.bounds = np.zeros((rows, 4), np.float32)

    def first_f32(self) -> np.ndarray:
        return self.first.view(np.float16).astype(np.float32)


def sh_adam_quant_step(p: np.ndarray, g: Optional[np.ndarray], state: ShQuantState, active: int, lr: float,
                       rest_lr: float, b1: float, b2: float, c1: float, c2: float, eps: float,
                       regularization: float = 0.0) -> None:
    """The oracle's quantised SH Adam step, in place (the same contract as pht_sh_adam_quant_step)."""
    assert p.dtype == np.float32 and p.flags.c_contiguous and p.size == state.first.size
    if g is not None:
        assert g.dtype == np.float32 and g.flags.c_contiguous and g.size == p.size
    rows = state.first.shape[0]
    load().oracle_sh_adam_quant_step(_p(p), _p(g), _p(state.first), _p(state.packed), _p(state.bounds), rows,
                                     state.stride, active, lr, rest_lr, b1, b2, c1, c2, eps, regularization)


def pair_power2(conic3: np.ndarray, d2: np.ndarray) -> np.ndarray:
    """The blend's base-2 exponent log2(e) * power of each (conic, offset) pair."""
    conic3, d2 = f32(conic3), f32(d2)
    out = np.empty((conic3.shape[0],), np.float32)
    load().oracle_pair_power2(_p(conic3), _p(d2), _p(out), conic3.shape[0])
    return out


def lp_roundtrip(x: np.ndarray, fmt: int):
    x = f32(x)
    bits = np.empty(x.shape, np.uint16)
    back = np.empty_like(x)
    load().oracle_lp_roundtrip(_p(x), _p(bits), _p(back), x.size, fmt)
    return bits, back


def project(scene: Scene, camera: Camera, options: RasterizeOptions, precision: int = 32) -> np.ndarray:
    """Per-Gaussian projection rows: mean2d(2) conic(3) opacity depth radius tiles rgb(3) clamp fp64 visible."""
    out = np.zeros((scene.n, 16), np.float64)
    cam = make_camera(camera, options)
    s = settings_for(scene, options)
    g = scene.gaussians()
    load().oracle_project(precision, C.byref(g), C.byref(cam), C.byref(s), _p(out))
    return out


def enumerate_tiles(mean2d, conic3, opacity, grid_x, grid_y, ts, cap=4096) -> List[int]:
    conic3 = f32(conic3)
    buf = np.zeros((cap,), np.int32)
    n = load().oracle_enumerate(float(mean2d[0]), float(mean2d[1]), _p(conic3), float(opacity), grid_x, grid_y,
                                ts, _p(buf), cap)
    return buf[:min(n, cap)].tolist()


def l1_loss(color: np.ndarray, target: np.ndarray, weight: float = 1.0):
    color, target = f32(color), f32(target)
    grad = np.empty_like(color)
    loss = load().oracle_l1_loss(_p(color), _p(target), _p(grad), color[0].size, weight)
    return loss, grad


def adam_step(p, g, m, v, rows, stride, active, lr, lr2, secondary_from, b1, b2, c1, c2, eps, lo, hi, prefix):
    for a in (p, g, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    load().oracle_adam_step(_p(p), _p(g), _p(m), _p(v), rows, stride, active, lr, lr2, secondary_from, b1, b2, c1, c2,
                            eps, lo, hi, prefix)


class OracleTrainer:
    """The oracle's training step (forward, photometric loss, backward, fused Adam) over one scene,
    with its own Adam moments; the scene's float arrays are the parameters, updated in place.
    The learning rates and Adam constants mirror photara_studio_amd.train.TrainStep."""

    def __init__(self, scene: Scene, cameras: Sequence[Camera], options: RasterizeOptions, targets, lrs,
                 loss: str = "l1", ssim_weight: float = 0.2, photometric_weight: float = 1.0,
                 threads: Optional[int] = None, sh_optimizer: str = "fp32", sh_regularization: float = 0.0):
        self.scene, self.cameras, self.options = scene, list(cameras), options
        self.lrs, self.threads = lrs, threads or threads_default()
        self.loss_kind = 1 if loss == "l1_ssim" else 0
        self.ssim_weight, self.photometric_weight = ssim_weight, photometric_weight
        nv = len(self.cameras)
        self.cams = (_abi.pht_camera * nv)(*[make_camera(c, options) for c in self.cameras])
        self.settings = settings_for(scene, options)
        self.targets = [f32(t) for t in targets]
        self.color = [np.zeros((3, c.height, c.width), np.float32) for c in self.cameras]
        self.alpha = [np.zeros((c.height, c.width), np.float32) for c in self.cameras]
        self.grads = [np.zeros_like(c) for c in self.color]
        self.outs = (_abi.pht_view_outputs * nv)()
        for k in range(nv):
            self.outs[k].color, self.outs[k].alpha = self.color[k].ctypes.data, self.alpha[k].ctypes.data
        self.tp = (C.c_void_p * nv)(*[t.ctypes.data for t in self.targets])
        self.gp = (C.c_void_p * nv)(*[g.ctypes.data for g in self.grads])
        keys = ("means", "log_scales", "quaternions", "opacity_logits", "sh")
        self.m = {k: np.zeros_like(scene.a[k]) for k in keys}
        self.v = {k: np.zeros_like(scene.a[k]) for k in keys}
        self.sh_state = None
        if sh_optimizer == "quant":
            n, bases = scene.a["sh"].shape[0], scene.a["sh"].shape[1]
            self.sh_state = ShQuantState(n, bases * 3)
        self.sh_regularization = float(sh_regularization)
        self.step_index = 0

    def step(self) -> float:
        self.step_index += 1
        b1, b2 = np.float32(0.9), np.float32(0.999)
        a = _abi.pht_adam()
        sa = self.scene.a
        a.means, a.log_scales = sa["means"].ctypes.data, sa["log_scales"].ctypes.data
        a.quaternions, a.opacity_logits, a.sh = (sa["quaternions"].ctypes.data, sa["opacity_logits"].ctypes.data,
                                                 sa["sh"].ctypes.data)
        a.m_means, a.v_means = self.m["means"].ctypes.data, self.v["means"].ctypes.data
        a.m_log_scales, a.v_log_scales = self.m["log_scales"].ctypes.data, self.v["log_scales"].ctypes.data
        a.m_quaternions, a.v_quaternions = self.m["quaternions"].ctypes.data, self.v["quaternions"].ctypes.data
        a.m_opacity_logits = self.m["opacity_logits"].ctypes.data
        a.v_opacity_logits = self.v["opacity_logits"].ctypes.data
        a.m_sh, a.v_sh = self.m["sh"].ctypes.data, self.v["sh"].ctypes.data
        if self.sh_state is not None:
            a.sh_first, a.sh_packed = self.sh_state.first.ctypes.data, self.sh_state.packed.ctypes.data
            a.sh_bounds, a.sh_regularization = self.sh_state.bounds.ctypes.data, self.sh_regularization
        l = self.lrs
        a.lr_means, a.lr_scales, a.lr_quaternions = l.means, l.scales, l.quaternions
        a.lr_opacity, a.lr_sh0, a.lr_sh_rest = l.opacities, l.sh0, l.sh_rest
        a.beta1, a.beta2, a.epsilon = 0.9, 0.999, 1e-15
        a.correction1 = float(np.float32(1.0) - np.power(b1, np.float32(self.step_index)))
        a.correction2 = float(np.float32(1.0) - np.power(b2, np.float32(self.step_index)))
        a.min_log_scale, a.max_log_scale = -float("inf"), float("inf")
        g = self.scene.gaussians()
        return float(load().oracle_train_step(C.byref(g), self.cams, len(self.cameras), C.byref(self.settings),
                                              self.outs, self.tp, self.gp, C.byref(a), self.threads, self.loss_kind,
                                              self.ssim_weight, self.photometric_weight))


# ---------------------------------------------------------------- densification (ADC-IGS)
STATS_FIELDS = ("gradient", "count", "max_screen_radius", "priority", "max_geometry_gradient", "first_view",
                "view_support", "accumulated_error")


def make_stats(n: int) -> Dict[str, np.ndarray]:
    return {k: np.zeros(n, np.int32 if k == "first_view" else np.float32) for k in STATS_FIELDS}


def _stats_abi(st: Dict[str, np.ndarray]) -> _abi.pht_densify_stats:
    s = _abi.pht_densify_stats()
    for k in STATS_FIELDS:
        setattr(s, k, st[k].ctypes.data)
    return s


def densify_accumulate(st: Dict[str, np.ndarray], observations: Sequence[dict]) -> None:
    """observations: dicts with the pht_densify_observation fields (arrays as numpy)."""
    n = len(st["gradient"])
    obs = (_abi.pht_densify_observation * len(observations))()
    keep = []
    for k, o in enumerate(observations):
        for name in ("refine_weight", "visibility", "radii", "geometry_gradient", "image_error", "normalized_size"):
            a = o.get(name)
            if a is not None:
                a = np.ascontiguousarray(a, np.int32 if name == "radii" else np.float32)
                keep.append(a)
                setattr(obs[k], name, a.ctypes.data)
        for name in ("width", "height", "use_maximum", "require_contribution_visibility", "view_index"):
            setattr(obs[k], name, int(o.get(name, -1 if name == "view_index" else 0)))
        obs[k].step_score_power = float(o.get("step_score_power", 1.0))
        obs[k].oversize_screen_threshold = float(o.get("oversize_screen_threshold", 0.0))
    s = _stats_abi(st)
    load().oracle_densify_accumulate(C.byref(s), obs, len(observations), n)


def prune_masks(model: Dict[str, np.ndarray], min_opacity: float, max_bounds: float, center):
    n = len(model["opacity_logits"])
    g = _abi.pht_gaussians()
    g.count = n
    for k in ("means", "log_scales", "quaternions", "opacity_logits", "sh"):
        setattr(g, k, model[k].ctypes.data)
    keep, hard, op = np.zeros(n, np.uint8), np.zeros(n, np.uint8), np.zeros(n, np.float32)
    c = f32(center)
    load().oracle_prune_masks(C.byref(g), int(model["sh"].size // max(n, 1)), min_opacity, max_bounds, _p(c), _p(keep),
                              _p(hard), _p(op))
    return keep, hard, op


def select_keys(values, mask=None, exclude=None, mode=0, seed=0) -> np.ndarray:
    v = f32(values)
    keys = np.zeros(v.size, np.uint64)
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    e = None if exclude is None else np.ascontiguousarray(exclude, np.uint8)
    load().oracle_select_keys(_p(v), _p(m), _p(e), v.size, mode, C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), _p(keys))
    return keys


def select_rows(values, k, mode, mask=None, exclude=None, seed=0):
    """The k eligible rows with the largest keys, by a full sort (the checker of the radix
    select): (rows ascending, eligible count, threshold key)."""
    keys = select_keys(values, mask, exclude, mode, seed)
    elig = np.flatnonzero(keys != 0)
    k = min(max(int(k), 0), len(elig))
    order = elig[np.argsort(keys[elig], kind="stable")[::-1]]
    rows = np.sort(order[:k])
    return rows, len(elig), (int(keys[order[k - 1]]) if k else 0)


def key_value(key: int, mode: int) -> float:
    hi = (key >> 32) & 0xFFFFFFFF
    o = (~hi & 0xFFFFFFFF) if mode == 1 else hi
    bits = (o ^ 0x80000000) if o & 0x80000000 else (~o & 0xFFFFFFFF)
    return float(np.array([bits], np.uint32).view(np.float32)[0])


def densify_split(model: Dict[str, np.ndarray], parents, child_base, samples, mode, min_opacity, split_at=0.0,
                  screens=None, split_k=0.0, lp: Optional[Dict[str, np.ndarray]] = None, lp_format: int = 0) -> None:
    p = _abi.pht_split_params()
    p.means, p.log_scales = model["means"].ctypes.data, model["log_scales"].ctypes.data
    p.opacity_logits, p.quaternions = model["opacity_logits"].ctypes.data, model["quaternions"].ctypes.data
    p.lp_format = lp_format
    if lp_format:
        p.log_scales_lp, p.opacity_logits_lp = lp["log_scales"].ctypes.data, lp["opacity_logits"].ctypes.data
    p.mode, p.minimum_opacity, p.split_at_screen_size, p.split_opacity_k = mode, min_opacity, split_at, split_k
    par = np.ascontiguousarray(parents, np.int32)
    smp = f32(samples)
    scr = None if screens is None else f32(screens)
    load().oracle_densify_split(C.byref(p), _p(par), len(par), child_base, _p(smp), _p(scr))


def adc_decay(model: Dict[str, np.ndarray], opacity_decay: float, log_scale_decay: float,
              lp: Optional[Dict[str, np.ndarray]] = None, lp_format: int = 0) -> None:
    n = len(model["opacity_logits"])
    load().oracle_adc_decay(_p(model["log_scales"]), _p(model["opacity_logits"]),
                            _p(lp["log_scales"]) if lp_format else None, _p(lp["opacity_logits"]) if lp_format else None,
                            lp_format, n, opacity_decay, log_scale_decay)


def igs_rows(st, keep_rows, retained, opacities, median, allow_growth, gthr, sthr, weights=True):
    pr, pos = np.zeros(retained, np.float32), np.zeros(retained, np.uint8)
    counts = np.zeros(3, np.uint32)
    rw, gw, ow = (np.zeros(retained, np.float32) for _ in range(3))
    kr = None if keep_rows is None else np.ascontiguousarray(keep_rows, np.int32)
    s = _stats_abi(st)
    load().oracle_igs_rows(C.byref(s), _p(kr), retained, _p(f32(opacities)), median, int(allow_growth), gthr, sthr,
                           _p(pr), _p(pos), _p(counts), _p(rw) if weights else None, _p(gw), _p(ow))
    return pr, pos, [int(c) for c in counts], rw, gw, ow


def select_igs_parents(rw, ow, gw, slots, desired, capacity, random):
    """select_igs_parents (densification_igs.cpp:226-277) by sorting: (parents ascending,
    replacement, oversized, growth)."""
    from photara_studio_amd import densify as D
    rw = f32(rw)
    chosen = np.zeros(len(rw), bool)
    k = min(slots, capacity)
    n_rep = n_over = n_grow = 0
    if k and len(rw):
        rows, _, _ = select_rows(rw, k, 2, seed=D.gumbel_seed(random))
        chosen[rows] = True
        n_rep = len(rows)
    remaining = capacity - n_rep
    if remaining and ow is not None:
        rows, _, _ = select_rows(ow, remaining, 3, exclude=chosen.astype(np.uint8))
        chosen[rows] = True
        n_over = len(rows)
        remaining -= n_over
    if remaining and desired and gw is not None:
        rows, _, _ = select_rows(gw, min(remaining, desired), 2, exclude=chosen.astype(np.uint8),
                                 seed=D.gumbel_seed(random))
        chosen[rows] = True
        n_grow = len(rows)
    return np.flatnonzero(chosen).astype(np.int32), n_rep, n_over, n_grow


MOMENT_NAMES = ("means_first", "means_second", "scales_first", "scales_second", "rotations_first", "rotations_second",
                "opacity_first", "opacity_second", "sh_first", "sh_second")
MODEL_NAMES = ("means", "log_scales", "quaternions", "opacity_logits", "sh")


def igs_refine(model: Dict[str, np.ndarray], stats: Dict[str, np.ndarray], iteration: int, scene_extent: float,
               center, options, random, moments: Dict[str, np.ndarray], quant: Optional[Dict[str, np.ndarray]] = None,
               lp: Optional[Dict[str, np.ndarray]] = None, lp_format: int = 0, growth_cap: int = 0):
    """IgsStrategy::refine (densification_igs.cpp:279-515) on host arrays, selection by sorting.
    Mutates the dicts (rows replaced); returns (grown, pruned, info)."""
    from photara_studio_amd import densify as D
    old = len(model["opacity_logits"])
    if not D.is_refinement_iteration(iteration, options) or old == 0:
        return 0, 0, {}
    keep0, hard, op = prune_masks(model, options.prune_opacity, 100.0 * scene_extent, center)
    recycled = (keep0 == 0) & (hard == 0)
    keep = hard == 0
    hard_count = int(hard.sum())
    frac = float(min(max(np.float32(options.dense_recycle_fraction), np.float32(0)), np.float32(1)))
    limit = int(np.ceil(float(old) * frac))
    budget = limit - hard_count if limit > hard_count else 0
    if budget:
        rows, _, _ = select_rows(op, budget, 1, mask=recycled)
        keep[rows] = False
    best_rows, _, _ = select_rows(op, 1, 0)
    best = np.zeros(old, bool)
    best[best_rows] = True
    keep |= best & (hard == 0)
    retained = int(keep.sum())
    if retained > options.densification_cap:
        rows, _, _ = select_rows(op, retained - options.densification_cap, 1, mask=keep, exclude=best)
        keep[rows] = False
        retained -= len(rows)
    keep_rows = np.flatnonzero(keep).astype(np.int32) if retained != old else None
    pruned = old - retained
    gcap = options.densification_cap if growth_cap == 0 else min(options.densification_cap, max(old, growth_cap))
    capacity = gcap - retained if gcap > retained else 0
    rp = np.float32(1.0) - np.float32(iteration) / max(np.float32(1.0), np.float32(options.iterations))
    prog = max(rp, np.float32(0.0))
    od, lsd = D.adc_decay_factors(float(np.float32(options.opacity_decay) * prog),
                                  float(np.float32(options.scale_decay) * prog))

    def remap(param_rows, adam_rows):
        def take(a, rows):
            out = np.zeros((len(rows),) + a.shape[1:], a.dtype)
            ok = rows >= 0
            out[ok] = a[rows[ok]]
            return out
        for k in list(model):
            model[k] = take(model[k], param_rows)
        if lp is not None:
            for k in list(lp):
                lp[k] = take(lp[k], param_rows)
        for k in list(moments):
            moments[k] = take(moments[k], adam_rows)
        if quant is not None:
            for k in list(quant):
                quant[k] = take(quant[k], adam_rows)

    info = {}
    if capacity == 0:
        if keep_rows is not None:
            remap(keep_rows, keep_rows)
        adc_decay(model, od, lsd, lp, lp_format)
        return 0, pruned, info
    gthr, sthr = float(options.densify_gradient_threshold), float(options.densify_screen_threshold)
    pr, pos, counts, _, _, _ = igs_rows(stats, keep_rows, retained, op, 1.0, False, gthr, sthr, weights=False)
    allow = iteration < D.grow_stop_iteration(options)
    growth_candidates = counts[1] if allow else 0
    median = 1.0
    if counts[0]:
        rows, _, thr = select_rows(pr, counts[0] // 2 + 1, 1, mask=pos)
        median = max(key_value(thr, 1), 1e-9)
    _, _, _, rw, gw, ow = igs_rows(stats, keep_rows, retained, op, median, allow, gthr, sthr)
    desired = int(D.round_half_away(growth_candidates * float(np.float32(options.densify_select_fraction)))) \
        if allow else 0
    info = dict(median=median, positive=counts[0])
    parents, _, _, _ = select_igs_parents(rw, ow if allow else None, gw if allow else None, pruned, desired, capacity,
                                          random)
    count = len(parents)
    if count == 0:
        if keep_rows is not None:
            remap(keep_rows, keep_rows)
    else:
        samples = D.normal_floats(random, 3 * count).reshape(count, 3)
        base = keep_rows if keep_rows is not None else np.arange(old, dtype=np.int32)
        param_rows = np.concatenate([base, base[parents]]).astype(np.int32)
        adam_rows = param_rows.copy()
        adam_rows[parents] = -1
        adam_rows[retained:] = -1
        remap(param_rows, adam_rows)
        densify_split(model, parents, retained, samples, 2, options.prune_opacity, lp=lp, lp_format=lp_format)
    adc_decay(model, od, lsd, lp, lp_format)
    info.update(parents=parents)
    return count, pruned, info
