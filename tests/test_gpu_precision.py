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


@pytest.mark.timeout(900)
@pytest.mark.parametrize("lp", ["bf16", "fp16"])
def test_full_c3_error_vs_fp64(cuda, oracle, lp):
    """The full reduced-precision workload: 6,000,000 Gaussians, four 1920x1080 views on a radius-8
    ring, SH / log-scales / opacity logits stored in bf16 or fp16, FP32 blending and accumulation.
    - Decision-matched FP64 (the arithmetic error): images and alpha within 1e-3 max-abs (measured
      ~2e-5), gradients within 1e-3 relative (measured <= 4e-5).
    - Order-matched FP64 (the float run's tile lists and blend order, double 1/255 and termination
      decisions): a cut-off flip moves a pixel by at most ~1/255 x colour: every pixel within
      1.2/255, at most 1e-4 of them above 1e-3 (measured ~130 of 8.3M, max 4.2e-3).
    - Plain FP64: the same plus splats whose depths lie within float rounding of each other and
      blend in the other order (the reference sorts by float32 depth bits): at most 1e-3 of the
      pixels above 1e-3 (measured ~3,300, max 8.2e-2)."""
    from photara_studio_amd.rasterizer import Rasterizer
    from photara_studio_amd.scenes import WORKLOADS, model_from, synthetic_arrays, workload_cameras
    from photara_studio_amd.train import to_low_precision
    import os
    threads = min(16, os.cpu_count() or 1)
    n, box = WORKLOADS["c3"]
    arrays = synthetic_arrays(n, box, 0)
    cams = workload_cameras("c3")
    opts = training_options()
    print(f"[c3] {lp}: GPU forward/backward of {n} gaussians x {len(cams)} views", flush=True)
    model = to_low_precision(model_from(arrays, cuda), torch.bfloat16 if lp == "bf16" else torch.float16)
    r = Rasterizer()
    res = r.forward_views(model, cams, opts)
    counts = res[0].context.impl.ws.counts()
    rng = np.random.default_rng(3)
    gimg = [rng.normal(0, 1e-3, size=tuple(x.color.shape)).astype(np.float32) for x in res]
    grads = r.backward_views(model, res, [torch.from_numpy(g).to(cuda) for g in gimg])
    n_pix = sum(c.width * c.height for c in cams)
    scene = oracle.Scene(arrays, lp=lp)
    print(f"[c3] {lp}: decision-matched FP64 oracle ({threads} threads)", flush=True)
    matched = oracle.OracleRun(scene, cams, opts, 65, threads=threads)
    m_img, m_alpha, _ = _img_err(res, matched)
    m_grad = _grad_err(grads, matched.backward(gimg))
    fp64_routed = int(matched.counts.fp64_conics)
    del matched
    print(f"[c3] {lp}: order-matched FP64 oracle", flush=True)
    o_img, _, o_over = _img_err(res, oracle.OracleRun(scene, cams, opts, 66, threads=threads))
    print(f"[c3] {lp}: plain FP64 oracle", flush=True)
    p_img, _, p_over = _img_err(res, oracle.OracleRun(scene, cams, opts, 64, threads=threads))
    print(f"[c3] {lp} vs decision-matched FP64: image {m_img:.2e}, alpha {m_alpha:.2e}; gradients "
          + ", ".join(f"{k} {v:.2e}" for k, v in m_grad.items()), flush=True)
    print(f"[c3] {lp} vs order-matched FP64: image {o_img:.2e} ({o_over} of {n_pix} pixels above 1e-3); "
          f"vs plain FP64: image {p_img:.2e} ({p_over} pixels above 1e-3); FP64-routed conics "
          f"{int(counts.fp64_conics)} (oracle {fp64_routed})", flush=True)
    assert m_img <= 1e-3 and m_alpha <= 1e-3, (m_img, m_alpha)
    assert all(v <= 1e-3 for v in m_grad.values()), m_grad
    assert o_img <= 1.2 / 255.0 and o_over <= 1e-4 * n_pix, (o_img, o_over)
    assert p_over <= 1e-3 * n_pix, p_over
    assert int(counts.fp64_conics) == fp64_routed
