"""The HIP path on the reference's own test fixtures (tests/golden/reference_fixtures.json):
closed-form alphas, the background identity, the thin-splat colour gradient and the reverse
compositing scene, each also compared with the oracle (forward bit-exact)."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from test_oracle_golden import FIX, camera_from, default_options, one_gaussian

pytestmark = pytest.mark.gpu


def _render(arrays, cams, opts, cuda, sh_degree=0, colors=None):
    from photara_studio_amd.rasterizer import Rasterizer
    from photara_studio_amd.scenes import model_from
    model = model_from(arrays, cuda, sh_degree=sh_degree)
    if colors is not None:
        opts.colors_precomp = torch.from_numpy(np.asarray(colors, np.float32)).to(cuda).contiguous()
    r = Rasterizer()
    return r, model, r.forward_views(model, cams, opts)


def test_centre_alpha_closed_forms_gpu(cuda, oracle):
    f = FIX["centre_alpha"]
    arrays = one_gaussian(f)
    cam = camera_from(f["camera"])
    cx, cy = f["centre_pixel"]
    for kernel, expect in ((0.0, f["expect_alpha_kernel0"]), (0.1, f["expect_alpha_kernel01_compensated"])):
        _, _, res = _render(arrays, [cam], default_options(kernel_size=kernel), cuda)
        a = res[0].alpha.cpu().numpy()
        assert abs(float(a[cy, cx]) - expect) < f["tol"]
        o = oracle.OracleRun(oracle.Scene(arrays, sh_degree=0), [cam], default_options(kernel_size=kernel), 32)
        np.testing.assert_array_equal(a, o.alpha[0])


def test_background_identity_gpu(cuda):
    f = FIX["background_identity"]
    arrays = one_gaussian(f)
    cam = camera_from(f["camera"])
    _, _, black = _render(arrays, [cam], default_options(), cuda)
    _, _, col = _render(arrays, [cam], default_options(background=f["background"]), cuda)
    b, c = black[0].color.cpu().numpy(), col[0].color.cpu().numpy()
    alpha = black[0].alpha.cpu().numpy()
    assert black[0].visibility.cpu().numpy().tolist() == [1.0]
    bg = np.array(f["background"], np.float32)[:, None, None]
    assert np.abs(c - b - (1.0 - alpha)[None] * bg).max() < f["tol"]


def test_thin_splats_gpu(cuda):
    f = FIX["thin_splats"]
    cam = camera_from(f["camera"])
    W, H = cam.width, cam.height
    for k in range(2):
        arrays = {
            "means": np.array([f["means"][k]], np.float32),
            "log_scales": np.log(np.array([f["scales"][k]], np.float64)).astype(np.float32),
            "quaternions": np.array([f["rotations"][k]], np.float32),
            "opacity_logits": np.array([[f["opacity_logit"]]], np.float32),
            "sh": np.zeros((1, 1, 3), np.float32),
        }
        r, model, res = _render(arrays, [cam], default_options(), cuda, colors=[f["colors_precomp"]])
        assert res[0].rendered_instances > 0
        gc = torch.zeros((3, H, W), device=cuda)
        gc[0] = 1.0
        g = r.backward(model, res[0], gc, torch.zeros((H, W), device=cuda))
        expected = float(res[0].alpha.double().sum())
        if k == 0:
            assert expected > f["min_alpha_sum_sample0"]
        assert abs(float(g.colors_precomp[0, 0]) - expected) < f["rel_tol"] * max(expected, 1.0)
        assert g.sh is None
        for t in (g.means, g.log_scales, g.quaternions, g.opacity_logits):
            assert torch.isfinite(t).all()


def test_reverse_compositing_scene_gpu(cuda, oracle):
    f = FIX["reverse_compositing"]
    n = f["count"]
    arrays = {"means": np.array([[0.0, 0.0, 2.0 + g * 0.08] for g in range(n)], np.float32),
              "log_scales": np.full((n, 3), math.log(0.25), np.float32),
              "quaternions": np.tile(np.array([1.0, 0, 0, 0], np.float32), (n, 1)),
              "opacity_logits": np.array([[math.log(o / (1 - o))] for o in
                                          [0.995 if g % 6 == 0 else (0.08 if g % 3 == 0 else 0.65) for g in range(n)]],
                                         np.float32),
              "sh": np.array([[[(0.05 + 0.13 * (g % 6) - 0.5) / f["sh0"], (0.1 + 0.11 * (g % 5) - 0.5) / f["sh0"],
                                (0.12 + 0.07 * (g % 7) - 0.5) / f["sh0"]]] for g in range(n)], np.float32)}
    cam = camera_from(f["camera"])
    opts = default_options(background=f["background"])
    _, _, res = _render(arrays, [cam], opts, cuda)
    o = oracle.OracleRun(oracle.Scene(arrays, sh_degree=0), [cam], default_options(background=f["background"]), 32)
    np.testing.assert_array_equal(res[0].color.cpu().numpy(), o.color[0])
