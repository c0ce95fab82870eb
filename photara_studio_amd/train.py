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
                 view_outputs: Optional[bool] = None, view_groups: int = 1):
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
        does, and the forward skips those writes).
        view_groups: split the views into this many contiguous groups (1-4), each forwarded and
        blended on its own workspace and stream, with one per-Gaussian backward + Adam over all of
        them (pht_backward_groups): one group's sorts then run beside another group's blends. The
        step computes the same update; only the launch schedule differs."""
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
        self.view_groups = int(view_groups)
        if not 1 <= self.view_groups <= min(4, nv):
            raise ValueError("view_groups must be between 1 and min(4, views)")
        if densify and self.view_groups > 1:
            raise ValueError("the densification statistics need a single view group")
        sizes = [nv // self.view_groups + (1 if k < nv % self.view_groups else 0) for k in range(self.view_groups)]
        self.groups = [(sum(sizes[:k]), sizes[k]) for k in range(self.view_groups)]  # (first view, views)
        if self.view_groups == 1:
            self.capacities = [instance_capacity or default_instance_capacity(n, nv)]
        else:
            self.capacities = [(int(instance_capacity) * nk // nv + 4096) if instance_capacity else
                               default_instance_capacity(n, nk) for _, nk in self.groups]
        self.wss = [Workspace(dev, n, nk, ts, tiles_of(self.cameras[v0:v0 + nk], ts),
                              sum(int(c.width) * int(c.height) for c in self.cameras[v0:v0 + nk]), cap)
                    for (v0, nk), cap in zip(self.groups, self.capacities)]
        self.ws, self.capacity = self.wss[0], self.capacities[0]
        self._streams = [torch.cuda.Stream(device=dev) for _ in self.groups] if self.view_groups > 1 else []
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
        self._counts_host = torch.zeros((_COUNT_RING, self.view_groups, 4), dtype=torch.int32, pin_memory=True)
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
            self.ws, self.wss = None, []
            torch.cuda.empty_cache()
            self.ws = Workspace(dev, int(n * 1.25), len(self.cameras), int(self.options.tile_size), self.tiles,
                                self.pixels, self.capacity)
            self.wss = [self.ws]
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
        for k, ws in enumerate(self.wss):
            check(self.lib.pht_get_counts_async(ws.handle, C.c_void_p(self._counts_host[row, k].data_ptr()),
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
            if bool((self._counts_host[row, :, 2] != 0).any()):
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

    @staticmethod
    def _ptr(arr, first: int, elem):
        """Typed pointer to element `first` of a ctypes array (a view group's slice)."""
        return C.cast(C.addressof(arr) + first * C.sizeof(elem), C.POINTER(elem))

    def _forward_loss(self, ws: Workspace, v0: int, nk: int, stream) -> None:
        """Forward and loss of views v0 .. v0 + nk on `ws`, enqueued on `stream` (a c_void_p)."""
        cams = self._ptr(self.cams, v0, _abi.pht_camera)
        outs = self._ptr(self.outs, v0, _abi.pht_view_outputs)
        if self.loss_kind == "l1":
            # the L1 loss and its gradient come out of the forward's blend (pht_forward_l1)
            check(self.lib.pht_forward_l1(ws.handle, C.byref(self.gauss), cams, nk, C.byref(self.settings), outs,
                                          self._ptr(self._target_ptrs, v0, C.c_void_p),
                                          self._ptr(self._grad_ptrs, v0, C.c_void_p),
                                          C.c_void_p(self.loss_sum.data_ptr()), self.photometric_weight, 0, stream))
            return
        check(self.lib.pht_forward(ws.handle, C.byref(self.gauss), cams, nk, C.byref(self.settings), outs, 0, stream))
        for k in range(v0, v0 + nk):
            cam = self.cameras[k]
            check(self.lib.pht_l1_ssim_loss(C.c_void_p(self.colors[k].data_ptr()),
                                            C.c_void_p(self.targets[k].data_ptr()), None,
                                            C.c_void_p(self.grad_colors[k].data_ptr()), 0,
                                            C.c_void_p(self.loss_sum.data_ptr()),
                                            C.c_void_p(self.ssim_scratch[k].data_ptr()), int(cam.width),
                                            int(cam.height), self.ssim_weight, self.photometric_weight, stream))

    def step_eager(self) -> None:
        """One iteration launched kernel by kernel on the current stream (no host sync); with view
        groups, each group's forward and blend backward run on the group's stream."""
        self._poll()
        self.step_index += 1
        c1, c2 = self._corrections(self.step_index)
        self.adam.correction1, self.adam.correction2 = c1, c2
        if self.sh_adam is not None:
            self.sh_adam.correction1, self.sh_adam.correction2 = c1, c2
        main = torch.cuda.current_stream()
        stream = C.c_void_p(main.cuda_stream)
        self.loss_sum.zero_()
        adam = make_adam(self.model, self.adam, self.sh_adam)
        if self.view_groups == 1:
            self._forward_loss(self.ws, 0, len(self.cameras), stream)
            self._backward(adam, stream)
        else:
            for ws, gs, (v0, nk) in zip(self.wss, self._streams, self.groups):
                gs.wait_stream(main)
                self._forward_loss(ws, v0, nk, C.c_void_p(gs.cuda_stream))
            wsp = (C.c_void_p * self.view_groups)(*[ws.handle.value for ws in self.wss])
            gsp = (C.c_void_p * self.view_groups)(*[gs.cuda_stream for gs in self._streams])
            check(self.lib.pht_backward_groups(wsp, self.view_groups, C.byref(self.gauss), self.dl, None,
                                               C.byref(adam), gsp, stream))
        self._track("eager")

    def capture(self) -> None:
        """Capture forward + loss + backward + Adam as one hipGraph (hyper-parameters on device);
        with view groups the graph forks into one branch per group and joins before the
        per-Gaussian backward."""
        nv = len(self.cameras)
        self._adam_graph = make_adam(self.model, self.adam, self.sh_adam)
        self._adam_graph.device_hparams = self.hparams.data_ptr()
        handle = C.c_void_p()
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        if self.view_groups == 1:
            check(self.lib.pht_step_graph_create_densify(self.ws.handle, C.byref(self.gauss), self.cams, nv,
                                                         C.byref(self.settings), self.outs, self._target_ptrs,
                                                         self._grad_ptrs, C.c_void_p(self.loss_sum.data_ptr()),
                                                         C.byref(self.loss_cfg), C.byref(self._adam_graph),
                                                         C.byref(self._dstep) if self.densify else None, stream,
                                                         C.byref(handle)))
        else:
            wsp = (C.c_void_p * self.view_groups)(*[ws.handle.value for ws in self.wss])
            sizes = (C.c_int * self.view_groups)(*[nk for _, nk in self.groups])
            check(self.lib.pht_step_graph_create_groups(wsp, self.view_groups, sizes, C.byref(self.gauss), self.cams,
                                                        nv, C.byref(self.settings), self.outs, self._target_ptrs,
                                                        self._grad_ptrs, C.c_void_p(self.loss_sum.data_ptr()),
                                                        C.byref(self.loss_cfg), C.byref(self._adam_graph), stream,
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
        """The last forward's counts, summed over the view groups (overflow: any group's)."""
        if self.view_groups == 1:
            return self.ws.counts()
        out = _abi.pht_counts()
        for ws in self.wss:
            c = ws.counts()
            out.visible += c.visible
            out.instances += c.instances
            out.fp64_conics += c.fp64_conics
            out.overflow = 1 if (out.overflow or c.overflow) else 0
        return out

    def ensure_capacity(self, headroom: float = 1.25) -> int:
        """Run forward passes until every view group's instance list fits its workspace (growing
        a workspace to its measured count x headroom); returns the instance count. Parameters are
        not updated."""
        lib = self.lib
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        total = 0
        for k, (v0, nk) in enumerate(self.groups):
            while True:
                check(lib.pht_forward(self.wss[k].handle, C.byref(self.gauss), self._ptr(self.cams, v0, _abi.pht_camera),
                                      nk, C.byref(self.settings), self._ptr(self.outs, v0, _abi.pht_view_outputs), 0,
                                      stream))
                c = self.wss[k].counts()
                if not c.overflow:
                    total += int(c.instances)
                    break
                if self.graph is not None:
                    lib.pht_step_graph_destroy(self.graph)
                    self.graph = None
                self.capacities[k] = int(c.instances * headroom) + 4096
                dev = self.model.means.device
                rows = max(self.model.size(), int(self.wss[k].cap.gaussians))
                cams = self.cameras[v0:v0 + nk]
                ts = int(self.options.tile_size)
                self.wss[k] = None
                if k == 0:
                    self.ws = None
                torch.cuda.empty_cache()
                self.wss[k] = Workspace(dev, rows, nk, ts, tiles_of(cams, ts),
                                        sum(int(cm.width) * int(cm.height) for cm in cams), self.capacities[k])
            self.ws, self.capacity = self.wss[0], self.capacities[0]
        return total

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
