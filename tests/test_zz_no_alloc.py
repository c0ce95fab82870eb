"""The no-allocation L1+SSIM entry point (pht_l1_ssim_loss_noalloc) against pht_l1_ssim_loss with
caller scratch, through the C ABI: identical gradients and loss sums on two calls in a row, the
second on a smaller view (the static partials array keeps the first call's larger contents), and
a view above the array's limit is refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _pair(lib, cuda, h, w, seed, masked):
    g = torch.Generator(device="cpu").manual_seed(seed)
    color = torch.rand((3, h, w), generator=g).to(cuda)
    target = torch.rand((3, h, w), generator=g).to(cuda)
    mask = (torch.rand((h, w), generator=g) > 0.3).float().to(cuda) if masked else None
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    mp = C.c_void_p(mask.data_ptr()) if mask is not None else None
    grad_a = torch.full_like(color, 0.25)
    grad_b = torch.full_like(color, 0.25)
    loss_a = torch.full((1,), 0.5, device=cuda)
    loss_b = torch.full((1,), 0.5, device=cuda)
    nbytes = int(lib.pht_l1_ssim_scratch_bytes(w, h))
    scratch = torch.empty((nbytes // 4,), dtype=torch.float32, device=cuda)
    assert lib.pht_l1_ssim_loss(C.c_void_p(color.data_ptr()), C.c_void_p(target.data_ptr()), mp,
                                C.c_void_p(grad_a.data_ptr()), 1, C.c_void_p(loss_a.data_ptr()),
                                C.c_void_p(scratch.data_ptr()), w, h, 0.2, 1.0, stream) == 0
    assert lib.pht_l1_ssim_loss_noalloc(C.c_void_p(color.data_ptr()), C.c_void_p(target.data_ptr()), mp,
                                        C.c_void_p(grad_b.data_ptr()), 1, C.c_void_p(loss_b.data_ptr()), w, h, 0.2,
                                        1.0, stream) == 0
    torch.cuda.synchronize()
    assert torch.equal(grad_a, grad_b)
    assert torch.equal(loss_a, loss_b)
    assert float(loss_a) != 0.5


def test_noalloc_matches_scratch_entry_point(cuda):
    from photara_studio_amd import _abi
    lib = _abi.load()
    print("[noalloc] 1080x1920 then 300x200 (masked)", flush=True)
    _pair(lib, cuda, 1080, 1920, 1, False)
    _pair(lib, cuda, 200, 300, 2, True)
    # above the static array's limit: refused before any launch
    color = torch.zeros((3, 4112, 4112), device=cuda)
    rc = lib.pht_l1_ssim_loss_noalloc(C.c_void_p(color.data_ptr()), C.c_void_p(color.data_ptr()), None,
                                      C.c_void_p(color.data_ptr()), 0, C.c_void_p(color.data_ptr()), 4112, 4112,
                                      0.2, 1.0, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == _abi.PHT_EINVAL
