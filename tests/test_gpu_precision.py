"""Reduced-precision storage against the FP64 oracle (a scaled-down 4-view reduced-precision
workload: 400k Gaussians, 4 x 960x540): SH, log-scales and opacity logits stored in bf16 / fp16,
blending and accumulation in FP32.

The rendered image is piecewise smooth in the inputs: a pair whose alpha sits within rounding of
the 1/255 cut-off is kept in one precision and dropped in the other, which moves that pixel by
alpha x colour (up to ~4e-3). Two FP64 references are therefore reported per format:
  * the decision-matched FP64 oracle (the float run's tile lists, skips, termination and SH
    clamps, every value in double): the arithmetic error; bar 1e-3 max-abs for images and
    1e-3 for gradients (max-abs error over max-abs value);
  * the plain FP64 oracle of the same stored inputs: the same plus threshold flips; only
    isolated pixels (<= 1e-4 of them) may exceed 1e-3.
Also printed: the count of splats routed through the FP64 2x2 conic path (equal to the float
oracle's), and the image error against the FP64 render of the FP32 master values (the storage
quantisation itself, no bar).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import small_scene, training_options

pytestmark = pytest.mark.gpu

PARAMS = ("means", "log_scales", "quaternions", "opacity_logits", "sh")


def _img_err(res, ref):
    diffs = [np.abs(x.color.cpu().numpy().astype(np.float64) - ref.color_hp[k]) for k, x in enumerate(res)]
    adiffs = [np.abs(x.alpha.cpu().numpy().astype(np.float64) - ref.alpha_hp[k]) for k, x in enumerate(res)]
    over = sum(int((d.max(axis=0) > 1e-3).sum()) for d in diffs)
    return max(float(d.max()) for d in diffs), max(float(d.max()) for d in adiffs), over


def _grad_err(grads, ref):
    out = {}
    for key in PARAMS:
        got = getattr(grads, key).cpu().numpy().astype(np.float64)
        want = np.asarray(ref[key], np.float64)
        out[key] = float(np.abs(got - want).max() / max(np.abs(want).max(), 1e-30))
    return out


@pytest.mark.parametrize("lp", ["bf16", "fp16"])
def test_reduced_precision_error_vs_fp64(cuda, oracle, lp):
    from photara_studio_amd.rasterizer import Rasterizer
    from photara_studio_amd.scenes import model_from, ring_cameras
    from photara_studio_amd.train import to_low_precision

    arrays = small_scene(400_000, 3.0, 23)
    cams = ring_cameras(4, 8.0, [(960, 540)])
    opts = training_options()
    print(f"[precision] {lp}: GPU forward/backward", flush=True)
    model = to_low_precision(model_from(arrays, cuda), torch.bfloat16 if lp == "bf16" else torch.float16)
    r = Rasterizer()
    res = r.forward_views(model, cams, opts)
    counts = res[0].context.impl.ws.counts()
    rng = np.random.default_rng(3)
    gimg = [rng.normal(0, 1e-3, size=tuple(x.color.shape)).astype(np.float32) for x in res]
    grads = r.backward_views(model, res, [torch.from_numpy(g).to(cuda) for g in gimg])
    n_pix = sum(c.width * c.height for c in cams)

    print(f"[precision] {lp}: FP64 oracles", flush=True)
    scene = oracle.Scene(arrays, lp=lp)
    matched = oracle.OracleRun(scene, cams, opts, 65)
    m_img, m_alpha, _ = _img_err(res, matched)
    m_grad = _grad_err(grads, matched.backward(gimg))
    plain = oracle.OracleRun(scene, cams, opts, 64)
    p_img, p_alpha, over = _img_err(res, plain)
    p_grad = _grad_err(grads, plain.backward(gimg))
    o32 = oracle.OracleRun(scene, cams, opts, 32)
    quant, _, _ = _img_err(res, oracle.OracleRun(oracle.Scene(arrays), cams, opts, 64))
    print(f"  {lp} vs decision-matched FP64: image {m_img:.2e}, alpha {m_alpha:.2e}; gradients "
          + ", ".join(f"{k} {v:.2e}" for k, v in m_grad.items()))
    print(f"  {lp} vs plain FP64: image {p_img:.2e} ({over} of {n_pix} pixels above 1e-3), alpha {p_alpha:.2e}; "
          "gradients " + ", ".join(f"{k} {v:.2e}" for k, v in p_grad.items()))
    print(f"  fp64-routed conics {int(counts.fp64_conics)} (float oracle {int(o32.counts.fp64_conics)}); "
          f"storage quantisation vs FP64 of the FP32 masters: image {quant:.2e}")
    assert m_img <= 1e-3 and m_alpha <= 1e-3, (m_img, m_alpha)
    assert all(v <= 1e-3 for v in m_grad.values()), m_grad
    assert over <= n_pix * 1e-4, (over, n_pix)
    assert int(counts.fp64_conics) == int(o32.counts.fp64_conics)
