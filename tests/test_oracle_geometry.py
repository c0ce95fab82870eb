"""Pins of the oracle's per-Gaussian geometry (CPU), each against something the formula does not
share: an independent numpy EWA restatement in matrix form, the orthonormality of the real SH
basis on the sphere, and brute-force pixel coverage for the opacity-aware tile walk."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import training_options


def rot_from_quat(q):
    """Standard rotation matrix of a unit quaternion (w, x, y, z)."""
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def ewa_numpy(arrays, cam, kernel):
    """Independent EWA: Sigma = R diag(s^2) R^T, Sigma2D = J Rcw Sigma Rcw^T J^T + k I."""
    # the camera crosses the C ABI as float32: use exactly those values
    v = np.array(cam.world_to_camera, np.float32).astype(np.float64)
    M = v.reshape(4, 4).T  # column-major storage
    Rcw, tcw = M[:3, :3], M[:3, 3]
    out = []
    W, H = cam.width, cam.height
    fx, fy = float(np.float32(cam.fx)), float(np.float32(cam.fy))
    cx, cy = float(np.float32(cam.cx)), float(np.float32(cam.cy))
    for i in range(arrays["means"].shape[0]):
        p = arrays["means"][i].astype(np.float64)
        t = Rcw @ p + tcw
        if t[2] <= 0.2:
            out.append(None)
            continue
        limx, limy = 1.3 * W / (2 * fx), 1.3 * H / (2 * fy)
        tx = min(limx, max(-limx, t[0] / t[2])) * t[2]
        ty = min(limy, max(-limy, t[1] / t[2])) * t[2]
        J = np.array([[fx / t[2], 0, -fx * tx / t[2] ** 2], [0, fy / t[2], -fy * ty / t[2] ** 2]])
        q = arrays["quaternions"][i].astype(np.float64)
        q = q / max(np.linalg.norm(q), 1e-10)
        R = rot_from_quat(q)
        s = np.exp(arrays["log_scales"][i].astype(np.float64))
        S3 = R @ np.diag(s * s) @ R.T
        S2 = J @ Rcw @ S3 @ Rcw.T @ J.T + kernel * np.eye(2)
        conic = np.linalg.inv(S2)
        mean2d = np.array([fx * t[0] / t[2] + cx, fy * t[1] / t[2] + cy])
        opacity = 1 / (1 + math.exp(-float(arrays["opacity_logits"][i, 0])))
        out.append((mean2d, np.array([conic[0, 0], conic[0, 1], conic[1, 1]]), opacity, np.linalg.norm(t)))
    return out


def test_ewa_projection_matches_independent_numpy(oracle):
    from photara_studio_amd.scenes import look_at, synthetic_arrays
    arrays = synthetic_arrays(400, 1.0, 11)
    arrays["log_scales"] = arrays["log_scales"] + 1.5
    cam = look_at((1.0, -0.5, 3.5), width=160, height=120)
    opts = training_options()
    ref = ewa_numpy(arrays, cam, float(np.float32(0.3)))
    scene = oracle.Scene(arrays)
    rows64 = oracle.project(scene, cam, opts, 64)
    rows32 = oracle.project(scene, cam, opts, 32)
    checked = 0
    for i, r in enumerate(ref):
        if r is None or rows64[i, 14] == 0:
            continue
        mean2d, conic, op, depth = r
        np.testing.assert_allclose(rows64[i, 0:2], mean2d, rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(rows64[i, 2:5], conic, rtol=1e-9, atol=1e-12 * np.abs(conic).max())
        assert abs(rows64[i, 5] - op) <= 1e-12
        assert abs(rows64[i, 6] - depth) <= 1e-6 * depth  # depth key is a float bit pattern
        np.testing.assert_allclose(rows32[i, 0:2], mean2d, rtol=0, atol=2e-4)
        np.testing.assert_allclose(rows32[i, 2:5], conic, rtol=2e-4, atol=2e-4 * np.abs(conic).max())
        checked += 1
    assert checked > 100


def test_near_singular_conic_routed_through_fp64(oracle):
    """A thin splat seen edge-on without dilation: the float 2x2 inverse loses >8 bits, the FP64
    route restores it (count reported, conic equals the double result rounded)."""
    from photara_studio_amd.scenes import look_at
    cam = look_at((0.0, 0.0, 3.0), width=128, height=128)
    arrays = {
        "means": np.array([[0.05, 0.02, 0.0]], np.float32),
        "log_scales": np.log(np.array([[0.6, 0.0008, 0.2]], np.float32)),
        "quaternions": np.array([[0.96, 0.0, 0.0, 0.28]], np.float32),
        "opacity_logits": np.array([[1.0]], np.float32),
        "sh": np.zeros((1, 16, 3), np.float32),
    }
    opts = training_options(kernel_size=0.0)
    scene = oracle.Scene(arrays)
    r32 = oracle.project(scene, cam, opts, 32)
    r64 = oracle.project(scene, cam, opts, 64)
    assert r32[0, 13] == 1.0, "expected the FP64 route"
    # the route removes the inversion's cancellation; rounding already in the float covariance stays
    np.testing.assert_allclose(r32[0, 2:5], r64[0, 2:5], rtol=2e-2)
    opts.fp64_conic = False
    r32f = oracle.project(scene, cam, opts, 32)
    assert r32f[0, 13] == 0.0
    assert np.abs(r32f[0, 2:5] - r64[0, 2:5]).max() > np.abs(r32[0, 2:5] - r64[0, 2:5]).max()


def test_sh_basis_orthonormal_and_degree0_closed_form(oracle):
    """The 16 real SH functions (sh.cuh:8-65) are orthonormal on the sphere; degree 0 is
    C0 * c + 0.5 (the closed form splat_test.cpp:2425 relies on)."""
    nodes, weights = np.polynomial.legendre.leggauss(24)
    phis = np.linspace(0, 2 * np.pi, 48, endpoint=False)
    ct, ph = np.meshgrid(nodes, phis, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    dirs = np.stack([st * np.cos(ph), st * np.sin(ph), ct], -1).reshape(-1, 3)
    w = (weights[:, None] * np.full_like(ph, 2 * np.pi / len(phis))).reshape(-1)
    Y = np.zeros((16, dirs.shape[0]))
    for k in range(16):
        coef = np.zeros((16, 3))
        coef[k, 0] = 1.0
        Y[k] = oracle.sh_eval(3, dirs, coef)[:, 0] - 0.5
    gram = (Y * w) @ Y.T
    np.testing.assert_allclose(gram, np.eye(16), atol=1e-10)
    coef = np.zeros((16, 3))
    coef[0] = [0.5, 0.25, 0.1]
    rgb = oracle.sh_eval(0, dirs[:5], coef)
    np.testing.assert_allclose(rgb, np.tile(0.28209479177387814 * coef[0] + 0.5, (5, 1)), atol=1e-15)


@pytest.mark.parametrize("ts", [16, 8])
def test_tile_walk_covers_every_contributing_pixel(oracle, ts):
    """Every tile holding a pixel whose alpha reaches 1/255 is enumerated (no false negatives),
    every enumerated tile touches the alpha >= 1/255 ellipse, and no tile repeats."""
    rng = np.random.default_rng(4)
    W, H = 160, 112
    gx, gy = (W + ts - 1) // ts, (H + ts - 1) // ts
    ys, xs = np.mgrid[0:H, 0:W]
    pix = np.stack([xs.reshape(-1), ys.reshape(-1)], 1).astype(np.float32)
    for trial in range(60):
        s1, s2 = rng.uniform(0.5, 25.0, 2)
        th = rng.uniform(0, np.pi)
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        cov = R @ np.diag([s1 * s1, s2 * s2]) @ R.T
        conic = np.linalg.inv(cov)
        c3 = np.array([conic[0, 0], conic[0, 1], conic[1, 1]], np.float32)
        mean = rng.uniform(-20, [W + 20, H + 20]).astype(np.float32)
        op = np.float32(rng.uniform(0.01, 0.99))
        tiles = oracle.enumerate_tiles(mean, c3, op, gx, gy, ts)
        assert len(tiles) == len(set(tiles))
        d = (mean[None, :] - pix).astype(np.float32)
        p2 = oracle.pair_power2(np.tile(c3, (d.shape[0], 1)), d)
        alpha = np.minimum(np.float32(0.99), op * oracle.pexp2f_blend(np.maximum(p2, np.float32(-126.0))))
        hit = (p2 <= 0) & (alpha >= np.float32(1.0 / 255.0))
        need = set(((pix[hit, 1].astype(int) // ts) * gx + pix[hit, 0].astype(int) // ts).tolist())
        assert need <= set(tiles), (trial, sorted(need - set(tiles)))
        thr = 2.0 * math.log(float(op) * 255.0)
        for t in tiles:
            tx, ty = t % gx, t // gx
            sx = np.linspace(tx * ts - 0.5, (tx + 1) * ts + 0.5, 40)
            sy = np.linspace(ty * ts - 0.5, (ty + 1) * ts + 0.5, 40)
            qx, qy = np.meshgrid(sx - mean[0], sy - mean[1])
            q = conic[0, 0] * qx ** 2 + 2 * conic[0, 1] * qx * qy + conic[1, 1] * qy ** 2
            assert q.min() <= thr * 1.001, (trial, t)
