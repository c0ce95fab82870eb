"""The torch.autograd.Function over the HIP kernels: gradients of an arbitrary torch loss match the
FP64 oracle's gradients and FP64 central finite differences (h = 1e-6) on 64 seeded parameters."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import small_scene, training_options

pytestmark = pytest.mark.gpu


def test_autograd_matches_fp64_differences(cuda, oracle):
    from photara_studio_amd import rasterize
    from photara_studio_amd.scenes import look_at
    arrays = small_scene(60, 0.6, 31)
    arrays["log_scales"] = arrays["log_scales"] + 2.0
    cam = look_at((0.3, 0.2, 3.0), width=64, height=48)
    opts = training_options()
    rng = np.random.default_rng(4)
    wc = rng.normal(size=(3, 48, 64))
    wa = rng.normal(size=(48, 64))
    params = {k: torch.from_numpy(v).to(cuda).requires_grad_(True) for k, v in arrays.items()}
    print("[autograd] forward/backward through torch.autograd", flush=True)
    color, alpha = rasterize(params["means"], params["log_scales"], params["quaternions"], params["opacity_logits"],
                             params["sh"], cam, opts)
    loss = (color * torch.from_numpy(wc).float().to(cuda)).sum() + (alpha * torch.from_numpy(wa).float().to(cuda)).sum()
    loss.backward()
    grads = {k: v.grad.cpu().numpy().astype(np.float64) for k, v in params.items()}
    # FP64 oracle gradient of the same loss
    a64 = {k: v.astype(np.float64) for k, v in arrays.items()}
    scene = oracle.Scene(arrays)
    with oracle.inputs64(a64):
        r = oracle.OracleRun(scene, [cam], opts, 64, threads=4)
        og = r.backward([wc.astype(np.float32)], [wa.astype(np.float32)])

    def loss64(a):
        with oracle.inputs64(a):
            rr = oracle.OracleRun(scene, [cam], opts, 64, threads=4)
        return float((rr.color_hp[0] * wc).sum() + (rr.alpha_hp[0] * wa).sum())

    keys = ["means", "log_scales", "quaternions", "opacity_logits", "sh"]
    bad_fd, bad_ref = [], []
    for trial in range(64):
        key = keys[trial % 5]
        idx = tuple(int(rng.integers(0, s)) for s in a64[key].shape)
        a1 = {k: v.copy() for k, v in a64.items()}
        a2 = {k: v.copy() for k, v in a64.items()}
        a1[key][idx] += 1e-6
        a2[key][idx] -= 1e-6
        fd = (loss64(a1) - loss64(a2)) / 2e-6
        got = grads[key][idx]
        scale = max(abs(fd), 1e-3 * np.abs(grads[key]).max(), 1e-6)
        if abs(got - fd) > 1e-3 * scale:
            bad_fd.append((key, idx, got, fd))
        if abs(got - og[key][idx]) > 1e-3 * max(abs(og[key][idx]), 1e-3 * np.abs(og[key]).max(), 1e-6):
            bad_ref.append((key, idx, got, og[key][idx]))
    print(f"  {64 - len(bad_fd)}/64 within 1e-3 of FP64 differences, {64 - len(bad_ref)}/64 of the FP64 oracle")
    assert not bad_fd, bad_fd
    assert not bad_ref, bad_ref
