"""Training-step driver: forward + photometric loss + backward + fused Adam over a set of views.

Reproduces the per-iteration call sequence of the reference trainer (Trainer::train_impl,
photara/src/splat/trainer.cpp:1922-2491: rasterizer.forward, compute_step_loss, rasterizer.backward
with the fused structure Adam) for a fixed batch of views, either launched eagerly or replayed as
one captured hipGraph (the whole step, no host synchronisation inside it).
"""
from __future__ import annotations

import collections
import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _abi
from ._abi import check
from .rasterizer import (Camera, GaussianModel, RasterizeOptions, SHAdamUpdate, ShAdamQuant, StructureAdamUpdate,
                         Workspace, default_instance_capacity, make_adam, make_camera, make_gaussians,
                         make_settings, tiles_of, validate_model)


@dataclass
class LearningRates:
    """TrainingOptions defaults (photara/include/splat/options.hpp:242-258)."""
    means: float = 1.6e-4
    scales: float = 5e-3
    opacities: float = 5e-2
    quaternions: float = 1e-3
    sh0: float = 2.5e-3
    sh_rest: float = 1.25e-4

    @staticmethod
    def uniform(lr: float) -> "LearningRates":
        return LearningRates(lr, lr, lr, lr, lr, lr)


def to_low_precision(model: GaussianModel, dtype: torch.dtype) -> GaussianModel:
    """Attach bf16/fp16 copies of SH, log-scales and opacity logits (FP32 masters stay)."""
    model.sh_lp = model.sh.to(dtype).contiguous()
    model.log_scales_lp = model.log_scales.to(dtype).contiguous()
    model.opacity_logits_lp = model.opacity_logits.to(dtype).contiguous()
    return model


_COUNT_RING = 8


class TrainStep:
    def __init__(self, model: GaussianModel, cameras: Sequence[Camera], targets: Sequence[torch.Tensor],
                 options: RasterizeOptions, lrs: Optional[LearningRates] = None,
                 instance_capacity: Optional[int] = None, loss: str = "l1", ssim_weight: float = 0.2,
                 photometric_weight: float = 1.0, sh_optimizer: str = "quant",
                 sh_regularization_weight: float = 0.0, densify: bool = False,
                 view_indices: Optional[Sequence[int]] = None, densify_screen_threshold: float = 0.5,
                 view_outputs: Optional[bool] = None):
        """loss: "l1" (loss_kernel's L1, cuda_ops.cu:886-898) or "l1_ssim" (the trainer's default
        (1 - w) L1 + w (1 - SSIM), fused_ssim.cu:473-560, w = ssim_weight).
        sh_optimizer: "quant", the trainer's SH Adam with fp16 update + 8-bit log-second state
        (trainer.cpp:1569-1577, sh_adam_quant.cuh), or "fp32" (FP32 moments, the active-prefix Adam
        of cuda_ops.cu:1506-1546). sh_regularization_weight: the L2 weight on the non-DC SH
        coefficients applied inside the quantised step (TrainingOptions::sh_regularization_weight,
        0.001 in the reference's defaults).
        densify: accumulate the ADC-IGS densification statistics inside the step (every view of the
        batch observed as the trainer observes its render, trainer.cpp:2441-2449, with view ids
        ``view_indices`` and oversize threshold ``densify_screen_threshold``); ``refine()`` then
        runs a refinement and re-binds the step to the new model.
        view_outputs: also write each view's per-Gaussian radii and contribution visibility (the
        render's radii / visibility outputs) into ``self.radii`` / ``self.visibility``; by default
        only when ``densify`` is set, whose statistics read them (otherwise nothing in the step
        does, and the forward skips those writes)."""
        if loss not in ("l1", "l1_ssim"):
            raise ValueError(f"unknown loss {loss!r}")
        if sh_optimizer not in ("quant", "fp32"):
            raise ValueError(f"unknown SH optimizer {sh_optimizer!r}")
        validate_model(model, options)
        self.model, self.cameras, self.options = model, list(cameras), options
        self.targets = [t.contiguous() for t in targets]
        self.lrs = lrs or LearningRates()
        dev = model.means.device
        n, nv = model.size(), len(self.cameras)
        self.sh_optimizer = sh_optimizer
        self.adam = StructureAdamUpdate.zeros_like(
            model, sh_moments=sh_optimizer == "fp32", means_lr=self.lrs.means, scales_lr=self.lrs.scales,
            quaternions_lr=self.lrs.quaternions, opacities_lr=self.lrs.opacities, sh_lr=self.lrs.sh0,
            sh_rest_lr=self.lrs.sh_rest)
        self.sh_state = self.sh_adam = None
        if sh_optimizer == "quant":
            self.sh_state = ShAdamQuant.zeros(n, int(model.sh.shape[1]) * 3, dev)
            self.sh_adam = SHAdamUpdate.of(self.sh_state, lr=self.lrs.sh0, rest_lr=self.lrs.sh_rest,
                                           beta1=self.adam.beta1, beta2=self.adam.beta2, epsilon=self.adam.epsilon,
                                           regularization_weight=float(sh_regularization_weight))
        self.cams = (_abi.pht_camera * nv)(*[make_camera(c, options) for c in self.cameras])
        self.settings = make_settings(model, options)
        self.gauss = make_gaussians(model, options)
        ts = int(options.tile_size)
        self.tiles = tiles_of(self.cameras, ts)
        self.pixels = sum(int(c.width) * int(c.height) for c in self.cameras)
        self.capacity = instance_capacity or default_instance_capacity(n, nv)
        self.ws = Workspace(dev, n, nv, ts, self.tiles, self.pixels, self.capacity)
        self.colors = [torch.empty((3, c.height, c.width), device=dev) for c in self.cameras]
        self.alphas = [torch.empty((c.height, c.width), device=dev) for c in self.cameras]
        self.view_outputs = bool(densify) if view_outputs is None else bool(view_outputs)
        if densify and not self.view_outputs:
            raise ValueError("the densification statistics need the per-view radii and visibility")
        self.radii = self.visibility = None
        if self.view_outputs:
            self.radii = [torch.empty((n,), dtype=torch.int32, device=dev) for _ in self.cameras]
            self.visibility = [torch.empty((n,), device=dev) for _ in self.cameras]
        self.grad_colors = [torch.empty_like(c) for c in self.colors]
        self.outs = (_abi.pht_view_outputs * nv)()
        for k in range(nv):
            self.outs[k].color, self.outs[k].alpha = self.colors[k].data_ptr(), self.alphas[k].data_ptr()
            if self.view_outputs:
                self.outs[k].radii, self.outs[k].visibility = self.radii[k].data_ptr(), self.visibility[k].data_ptr()
        self.dl = (_abi.pht_view_grads * nv)()
        for k in range(nv):
            self.dl[k].grad_color = self.grad_colors[k].data_ptr()
        self._target_ptrs = (C.c_void_p * nv)(*[t.data_ptr() for t in self.targets])
        self._grad_ptrs = (C.c_void_p * nv)(*[g.data_ptr() for g in self.grad_colors])
        self.loss_sum = torch.zeros((1,), device=dev)
        self.loss_kind = loss
        self.ssim_weight, self.photometric_weight = float(ssim_weight), float(photometric_weight)
        self.loss_cfg = _abi.pht_loss_config()
        self.loss_cfg.kind = _abi.PHT_LOSS_L1_SSIM if loss == "l1_ssim" else _abi.PHT_LOSS_L1
        self.loss_cfg.ssim_weight = self.ssim_weight
        self.loss_cfg.photometric_weight = self.photometric_weight
        self.ssim_scratch = []
        if loss == "l1_ssim":
            lib = _abi.load()
            for c in self.cameras:
                nbytes = int(lib.pht_l1_ssim_scratch_bytes(int(c.width), int(c.height)))
                if nbytes == 0:
                    raise ValueError("the L1+SSIM loss needs views wider and taller than 10 pixels")
                self.ssim_scratch.append(torch.empty((nbytes // 4,), dtype=torch.float32, device=dev))
            self._scratch_ptrs = (C.c_void_p * nv)(*[t.data_ptr() for t in self.ssim_scratch])
            self.loss_cfg.ssim_scratch = C.cast(self._scratch_ptrs, C.POINTER(C.c_void_p))
        self.hparams = torch.zeros((3,), device=dev)
        self.step_index = 0
        self.graph = None
        self.lib = _abi.load()
        # Counters of launched steps, copied back asynchronously (pinned ring, one row per step):
        # a step whose instance list overflowed the workspace left the model unchanged and is
        # re-run after the workspace grows (see _poll).
        self._counts_host = torch.zeros((_COUNT_RING, 4), dtype=torch.int32, pin_memory=True)
        self._pending = collections.deque()  # (step_index, ring row, event, "graph" | "eager")
        self._ring_next = 0
        self.overflow_recoveries = 0
        self.densify = bool(densify)
        self.dstats = None
        self.view_indices = list(view_indices) if view_indices is not None else list(range(nv))
        if len(self.view_indices) != nv:
            raise ValueError("one view index per camera")
        self.densify_screen_threshold = float(densify_screen_threshold)
        if self.densify:
            from .densify import make_densification_stats
            self.dstats = make_densification_stats(n, dev)
            self._bind_densify()

    def _bind_densify(self) -> None:
        d = _abi.pht_densify_step()
        d.stats = self.dstats.abi()
        for k, v in enumerate(self.view_indices):
            d.view_index[k] = int(v)
        d.use_maximum, d.require_contribution_visibility = 1, 1
        d.step_score_power, d.oversize_screen_threshold = 1.0, self.densify_screen_threshold
        d.geometry_from_refine = 1
        self._dstep = d

    def _backward(self, adam, stream) -> None:
        if self.densify:
            check(self.lib.pht_backward_densify(self.ws.handle, C.byref(self.gauss), self.dl, None, C.byref(adam),
                                                C.byref(self._dstep), stream))
        else:
            check(self.lib.pht_backward(self.ws.handle, C.byref(self.gauss), self.dl, None, C.byref(adam), stream))

    def refine(self, iteration: int, options, scene_extent: float, scene_center, random):
        """One ADC-IGS refinement of the model with this step's statistics and optimizer state
        (densification_igs.cpp:291-515), then the step is re-bound to the new model: parameter
        and per-view buffers, workspace capacity, statistics, and (at the next graph step) the
        captured graph. Returns the refinement counts."""
        from .densify import AdamStates, IgsStrategy, make_densification_stats
        if not self.densify:
            raise RuntimeError("refine() needs a step built with densify=True")
        self.sync()
        if self.graph is not None:
            self.lib.pht_step_graph_destroy(self.graph)
            self.graph = None
        counts, _ = IgsStrategy().refine(self.model, self.dstats, iteration, scene_extent, scene_center, options,
                                         random, AdamStates(self.adam, self.sh_state))
        if self.sh_adam is not None:
            st = self.sh_state
            self.sh_adam.first, self.sh_adam.packed, self.sh_adam.bounds = st.first, st.packed, st.bounds
        dev = self.model.means.device
        n = self.model.size()
        self.gauss = make_gaussians(self.model, self.options)
        self.radii = [torch.empty((n,), dtype=torch.int32, device=dev) for _ in self.cameras]
        self.visibility = [torch.empty((n,), device=dev) for _ in self.cameras]
        for k in range(len(self.cameras)):
            self.outs[k].radii, self.outs[k].visibility = self.radii[k].data_ptr(), self.visibility[k].data_ptr()
        self.dstats = make_densification_stats(n, dev)
        self._bind_densify()
        if not self.ws.fits(n, len(self.cameras), int(self.options.tile_size), self.tiles, self.pixels,
                            self.capacity):
            self.ws = None
            torch.cuda.empty_cache()
            self.ws = Workspace(dev, int(n * 1.25), len(self.cameras), int(self.options.tile_size), self.tiles,
                                self.pixels, self.capacity)
        self.ensure_capacity()  # the instance list grows with the model
        return counts

    def _corrections(self, step: int):
        b1, b2 = np.float32(self.adam.beta1), np.float32(self.adam.beta2)
        c1 = float(np.float32(1.0) - np.power(b1, np.float32(step)))
        c2 = float(np.float32(1.0) - np.power(b2, np.float32(step)))
        return c1, c2

    def _track(self, mode: str) -> None:
        """After a step's launches: enqueue the copy of its counters into the next ring row and
        record an event behind it."""
        while len(self._pending) >= _COUNT_RING:
            self._poll(block=True, limit=1)
        row = self._ring_next % _COUNT_RING
        self._ring_next += 1
        stream = torch.cuda.current_stream()
        check(self.lib.pht_get_counts_async(self.ws.handle, C.c_void_p(self._counts_host[row].data_ptr()),
                                            C.c_void_p(stream.cuda_stream)))
        ev = torch.cuda.Event()
        ev.record(stream)
        self._pending.append((self.step_index, row, ev, mode))

    def _poll(self, block: bool = False, limit: Optional[int] = None) -> None:
        """Reads the counters of finished steps in launch order (waiting for them when `block`).
        An overflowed step skipped its ranges, blends and Adam update, so the model is as before
        it; every later step saw the same model and the same instance count and was skipped too.
        The workspace then grows to the measured count and those steps run again, with their own
        step indices (Adam's bias corrections) and launch modes."""
        done = 0
        while self._pending and (limit is None or done < limit):
            s, row, ev, mode = self._pending[0]
            if block:
                ev.synchronize()
            elif not ev.query():
                return
            self._pending.popleft()
            done += 1
            if int(self._counts_host[row, 2]) != 0:
                modes = [mode] + [p[3] for p in self._pending]
                self._pending.clear()
                torch.cuda.synchronize()
                if self.step_index - s + 1 != len(modes):
                    raise RuntimeError("step counters out of order")
                self.step_index = s - 1
                self.overflow_recoveries += 1
                self.ensure_capacity()
                for m in modes:
                    self.step_graph() if m == "graph" else self.step_eager()
                return

    def sync(self) -> None:
        """Waits for every launched step and re-runs any that overflowed the instance capacity."""
        self._poll(block=True)

    def step_eager(self) -> None:
        """One iteration launched kernel by kernel on the current stream (no host sync)."""
        self._poll()
        self.step_index += 1
        c1, c2 = self._corrections(self.step_index)
        self.adam.correction1, self.adam.correction2 = c1, c2
        if self.sh_adam is not None:
            self.sh_adam.correction1, self.sh_adam.correction2 = c1, c2
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        nv = len(self.cameras)
        self.loss_sum.zero_()
        if self.loss_kind == "l1":
            # the L1 loss and its gradient come out of the forward's blend (pht_forward_l1)
            check(self.lib.pht_forward_l1(self.ws.handle, C.byref(self.gauss), self.cams, nv, C.byref(self.settings),
                                          self.outs, self._target_ptrs, self._grad_ptrs,
                                          C.c_void_p(self.loss_sum.data_ptr()), self.photometric_weight, 0, stream))
        else:
            check(self.lib.pht_forward(self.ws.handle, C.byref(self.gauss), self.cams, nv, C.byref(self.settings),
                                       self.outs, 0, stream))
            for k in range(nv):
                cam = self.cameras[k]
                check(self.lib.pht_l1_ssim_loss(C.c_void_p(self.colors[k].data_ptr()),
                                                C.c_void_p(self.targets[k].data_ptr()), None,
                                                C.c_void_p(self.grad_colors[k].data_ptr()), 0,
                                                C.c_void_p(self.loss_sum.data_ptr()),
                                                C.c_void_p(self.ssim_scratch[k].data_ptr()), int(cam.width),
                                                int(cam.height), self.ssim_weight, self.photometric_weight, stream))
        adam = make_adam(self.model, self.adam, self.sh_adam)
        self._backward(adam, stream)
        self._track("eager")

    def capture(self) -> None:
        """Capture forward + loss + backward + Adam as one hipGraph (hyper-parameters on device)."""
        nv = len(self.cameras)
        self._adam_graph = make_adam(self.model, self.adam, self.sh_adam)
        self._adam_graph.device_hparams = self.hparams.data_ptr()
        handle = C.c_void_p()
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(self.lib.pht_step_graph_create_densify(self.ws.handle, C.byref(self.gauss), self.cams, nv,
                                                     C.byref(self.settings), self.outs, self._target_ptrs, self._grad_ptrs,
                                                     C.c_void_p(self.loss_sum.data_ptr()), C.byref(self.loss_cfg),
                                                     C.byref(self._adam_graph),
                                                     C.byref(self._dstep) if self.densify else None, stream,
                                                     C.byref(handle)))
        self.graph = handle

    def step_graph(self) -> None:
        """One iteration as a replay of the captured step (no host sync; an instance-list overflow
        is found from the step's counters at a later call and the step re-run)."""
        self._poll()
        if self.graph is None:
            self.capture()
        self.step_index += 1
        c1, c2 = self._corrections(self.step_index)
        self.hparams.copy_(torch.tensor([c1, c2, self.adam.means_lr], dtype=torch.float32))
        check(self.lib.pht_step_graph_launch(self.graph, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        self._track("graph")

    def counts(self) -> _abi.pht_counts:
        return self.ws.counts()

    def ensure_capacity(self, headroom: float = 1.25) -> int:
        """Run forward passes until the instance list fits the workspace (growing it by the
        measured count x headroom); returns the instance count. Parameters are not updated."""
        lib = self.lib
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        while True:
            check(lib.pht_forward(self.ws.handle, C.byref(self.gauss), self.cams, len(self.cameras),
                                  C.byref(self.settings), self.outs, 0, stream))
            c = self.ws.counts()
            if not c.overflow:
                return int(c.instances)
            if self.graph is not None:
                lib.pht_step_graph_destroy(self.graph)
                self.graph = None
            self.capacity = int(c.instances * headroom) + 4096
            dev = self.model.means.device
            rows = max(self.model.size(), int(self.ws.cap.gaussians))
            self.ws = None
            torch.cuda.empty_cache()
            self.ws = Workspace(dev, rows, len(self.cameras), int(self.options.tile_size), self.tiles, self.pixels,
                                self.capacity)

    def loss(self) -> float:
        self.sync()
        return float(self.loss_sum.item())

    def __del__(self):
        try:
            if self.graph is not None:
                self.lib.pht_step_graph_destroy(self.graph)
                self.graph = None
        except Exception:
            pass
