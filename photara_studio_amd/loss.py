"""Photometric losses of one rendered view over torch tensors, on the gfx950 kernels.

``l1_loss``       : loss_kernel's L1 part (photara/src/splat/cuda_ops.cu:886-898).
``l1_ssim_loss``  : cuda_backend::fused_l1_ssim_loss (photara/src/splat/fused_ssim.cu:473-560),
                    the trainer's default objective (1 - w) L1 + w (1 - SSIM), w = 0.2.
Both return (loss as a 1-element device tensor, gradient wrt ``color``) and never synchronise.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import torch

from . import _abi
from ._abi import check


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check_image(t: torch.Tensor, name: str) -> None:
    if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.dim() != 3 or t.shape[0] != 3:
        raise ValueError(f"{name} must be a contiguous float32 [3,H,W] device tensor")


def l1_loss(color: torch.Tensor, target: torch.Tensor, weight: float = 1.0,
            grad: Optional[torch.Tensor] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    _check_image(color, "color")
    _check_image(target, "target")
    if target.shape != color.shape:
        raise ValueError("target must have the shape of color")
    grad = torch.empty_like(color) if grad is None else grad
    loss = torch.zeros((1,), device=color.device)
    check(_abi.load().pht_l1_loss(_ptr(color), _ptr(target), _ptr(grad), _ptr(loss), color[0].numel(),
                                  float(weight), _stream()))
    return loss, grad


def l1_ssim_loss(color: torch.Tensor, target: torch.Tensor, mask: Optional[torch.Tensor] = None,
                 ssim_weight: float = 0.2, photometric_weight: float = 1.0, grad: Optional[torch.Tensor] = None,
                 accumulate: bool = False) -> Tuple[torch.Tensor, torch.Tensor]:
    """Fused L1 + SSIM of one view; with ``accumulate`` the gradient is added to ``grad``."""
    _check_image(color, "color")
    _check_image(target, "target")
    if target.shape != color.shape:
        raise ValueError("target must have the shape of color")
    _, h, w = color.shape
    if mask is not None and (mask.dtype != torch.float32 or not mask.is_contiguous() or tuple(mask.shape) != (h, w)):
        raise ValueError("mask must be a contiguous float32 [H,W] tensor")
    if grad is None:
        if accumulate:
            raise ValueError("accumulate needs a grad tensor")
        grad = torch.empty_like(color)
    lib = _abi.load()
    scratch_bytes = int(lib.pht_l1_ssim_scratch_bytes(int(w), int(h)))
    scratch = torch.empty((max(scratch_bytes // 4, 1),), dtype=torch.float32, device=color.device)
    loss = torch.zeros((1,), device=color.device)
    check(lib.pht_l1_ssim_loss(_ptr(color), _ptr(target), _ptr(mask), _ptr(grad), int(bool(accumulate)), _ptr(loss),
                               _ptr(scratch), int(w), int(h), float(ssim_weight), float(photometric_weight),
                               _stream()))
    return loss, grad
