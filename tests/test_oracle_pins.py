"""Independent pins of the oracle functions the whole-pipeline checks leave open (CPU).

Each pin recomputes a quantity from the mathematics with numpy in float64, without the oracle's
own formula, and each test also shows that a plausible mistake in the formula would fail it:
- the eigen-radius ceil(3 sqrt(lambda)), lambda = mid + max(sqrt(0.1), (l1 - l2) / 2)
  (third_party/splat_drender/src/render_forward.cu:81-83), from numpy's eigenvalues of the
  screen covariance;
- the screen-space cap q = min(1, min(1024, W, H)^2 / (8 lambda_max)) and conic = inv(q Sigma2D)
  (third_party/splat_drender/src/device/footprint.cuh:13-37), and its backward (:39-48) through
  FP64 central differences of the FP64 forward over Gaussians the cap shrinks;
- the refine statistic sum_p |(dL_p/dmean.x W, dL_p/dmean.y H)| / max(1 - T_final, 1e-5)
  (render_backward.cu:478-482) of one Gaussian, with dL_p/dmean2d from central differences of a
  numpy render of that Gaussian;
- the L1 loss and its gradient w sign(C - C_hat) / (3 P) (photara/src/splat/cuda_ops.cu:886-898).
"""
from __future__ import annotations

import math

import numpy as np

from conftest import training_options
from test_oracle_geometry import rot_from_quat


def screen_covariances(arrays, cam, kernel):
    """Per Gaussian: (mean2d, Sigma2D + kernel I, activated opacity) in float64, or None when
    culled by the near plane. Sigma = R diag(s^2) R^T, Sigma2D = J Rcw Sigma Rcw^T J^T."""
    v = np.array(cam.world_to_camera, np.float32).astype(np.float64)
    M = v.reshape(4, 4).T
    Rcw, tcw = M[:3, :3], M[:3, 3]
    W, H = cam.width, cam.height
    fx, fy = float(np.float32(cam.fx)), float(np.float32(cam.fy))
    cx, cy = float(np.float32(cam.cx)), float(np.float32(cam.cy))
    out = []
    for i in range(arrays["means"].shape[0]):
        t = Rcw @ arrays["means"][i].astype(np.float64) + tcw
        if t[2] <= 0.2:
            out.append(None)
            continue
        limx, limy = 1.3 * W / (2 * fx), 1.3 * H / (2 * fy)
        tx = min(limx, max(-limx, t[0] / t[2])) * t[2]
        ty = min(limy, max(-limy, t[1] / t[2])) * t[2]
        J = np.array([[fx / t[2], 0, -fx * tx / t[2] ** 2], [0, fy / t[2], -fy * ty / t[2] ** 2]])
        q = arrays["quaternions"][i].astype(np.float64)
        R = rot_from_quat(q / max(np.linalg.norm(q), 1e-10))
        s = np.exp(arrays["log_scales"][i].astype(np.float64))
        S2 = J @ Rcw @ (R @ np.diag(s * s) @ R.T) @ Rcw.T @ J.T + kernel * np.eye(2)
        mean2d = np.array([fx * t[0] / t[2] + cx, fy * t[1] / t[2] + cy])
        op = 1.0 / (1.0 + math.exp(-float(arrays["opacity_logits"][i, 0])))
        out.append((mean2d, S2, op))
    return out


def cap_factor(S2, W, H):
    """q = min(1, limit^2 / (8 lambda_max)), limit = min(1024, W, H)."""
    lam = np.linalg.eigvalsh(S2).max()
    limit = min(1024.0, float(min(W, H)))
    return min(1.0, limit * limit / (8.0 * lam))


def eigen_radius(S2, sigmas=3.0, floor=0.1):
    """ceil(sigmas sqrt(mid + max(sqrt(floor), half-gap))) from numpy's eigenvalues."""
    l2, l1 = np.linalg.eigvalsh(S2)
    mid = 0.5 * (l1 + l2)
    return sigmas * math.sqrt(mid + max(math.sqrt(floor), 0.5 * (l1 - l2)))


def _near_integer(x, tol):
    return abs(x - round(x)) <= tol * max(1.0, abs(x))


def test_radius_matches_eigenvalue_closed_form(oracle):
    """Anisotropic splats, isotropic splats whose eigen-gap is below the 0.1 floor, and splats
    the cap shrinks: the oracle's radius is ceil(3 sqrt(lambda)) of numpy's eigenvalues."""
    from photara_studio_amd.scenes import look_at, synthetic_arrays
    cam = look_at((0.4, -0.3, 3.2), width=160, height=120)
    rng = np.random.default_rng(5)
    arrays = synthetic_arrays(600, 1.0, 12)
    # a third isotropic with screen sigma ~0.3-1.5 px (the floor decides their radius), a third
    # anisotropic at a few px, a few large enough for the cap
    n = arrays["means"].shape[0]
    iso = np.arange(n) % 3 == 0
    arrays["log_scales"][iso] = rng.uniform(-6.5, -4.2, (iso.sum(), 1)).astype(np.float32)
    arrays["log_scales"][~iso] += 1.8
    arrays["log_scales"][:8] = np.float32(math.log(1.6))
    opts = training_options()
    kernel = float(np.float32(0.3))
    covs = screen_covariances(arrays, cam, kernel)
    scene = oracle.Scene(arrays)
    r64 = oracle.project(scene, cam, opts, 64)
    r32 = oracle.project(scene, cam, opts, 32)
    checked = floor_decides = two_sigma_differs = capped = 0
    for i, c in enumerate(covs):
        if c is None or r64[i, 15] == 0 and r64[i, 7] == 0:
            continue
        S2 = c[1]
        q = cap_factor(S2, cam.width, cam.height)
        capped += q < 1.0
        S2c = S2 * q
        x = eigen_radius(S2c)
        if _near_integer(x, 1e-7):
            continue
        assert r64[i, 7] == math.ceil(x), (i, r64[i, 7], x)
        if not _near_integer(x, 1e-4):
            assert r32[i, 7] == math.ceil(x), (i, r32[i, 7], x)
        checked += 1
        floor_decides += math.ceil(eigen_radius(S2c, floor=0.0)) != math.ceil(x)
        two_sigma_differs += math.ceil(eigen_radius(S2c, sigmas=2.0)) != math.ceil(x)
    assert checked > 300 and capped >= 4
    # the data can tell the formula from a dropped floor and from 2 sigma
    assert floor_decides > 20 and two_sigma_differs > 100


def test_screen_cap_forward_closed_form(oracle):
    """Gaussians wider than the view: the oracle's conic is inv(q Sigma2D) with
    q = limit^2 / (8 lambda_max) < 1, to double rounding."""
    from photara_studio_amd.scenes import look_at
    cam = look_at((0.0, 0.0, 3.0), width=160, height=112)
    rng = np.random.default_rng(8)
    n = 24
    arrays = {
        "means": rng.uniform(-0.3, 0.3, (n, 3)).astype(np.float32),
        "log_scales": rng.uniform(math.log(0.4), math.log(2.0), (n, 3)).astype(np.float32),
        "quaternions": rng.normal(size=(n, 4)).astype(np.float32),
        "opacity_logits": rng.normal(size=(n, 1)).astype(np.float32),
        "sh": np.zeros((n, 16, 3), np.float32),
    }
    opts = training_options()
    covs = screen_covariances(arrays, cam, float(np.float32(0.3)))
    r64 = oracle.project(oracle.Scene(arrays), cam, opts, 64)
    capped = 0
    for i, (mean2d, S2, op) in enumerate(covs):
        q = cap_factor(S2, cam.width, cam.height)
        if q >= 1.0:
            continue
        capped += 1
        conic = np.linalg.inv(q * S2)
        ref = np.array([conic[0, 0], conic[0, 1], conic[1, 1]])
        np.testing.assert_allclose(r64[i, 2:5], ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
        # the uncapped inverse is far off: the check sees the cap
        raw = np.linalg.inv(S2)
        assert abs(raw[0, 0] - ref[0]) > 1e-3 * abs(ref[0])
    assert capped >= 12


def test_screen_cap_backward_central_differences(oracle):
    """FP64 gradients of Gaussians the cap shrinks (their conic depends on lambda_max) equal FP64
    central differences (h = 1e-6) of the FP64 forward, on means, log-scales and quaternions."""
    from photara_studio_amd.scenes import look_at
    cam = look_at((0.2, 0.1, 3.0), width=56, height=44)
    rng = np.random.default_rng(3)
    n = 6
    arrays = {
        "means": rng.uniform(-0.4, 0.4, (n, 3)).astype(np.float32),
        "log_scales": rng.uniform(math.log(1.2), math.log(2.5), (n, 3)).astype(np.float32),
        "quaternions": rng.normal(size=(n, 4)).astype(np.float32),
        "opacity_logits": rng.uniform(-3.0, -1.0, (n, 1)).astype(np.float32),
        "sh": rng.normal(0, 0.3, (n, 16, 3)).astype(np.float32),
    }
    opts = training_options()
    covs = screen_covariances(arrays, cam, float(np.float32(0.3)))
    assert all(c is not None and cap_factor(c[1], cam.width, cam.height) < 0.5 for c in covs)
    a64 = {k: v.astype(np.float64) for k, v in arrays.items()}
    gc = rng.normal(size=(3, cam.height, cam.width))
    ga = rng.normal(size=(cam.height, cam.width))
    scene = oracle.Scene(arrays)

    def loss(a):
        with oracle.inputs64(a):
            r = oracle.OracleRun(scene, [cam], opts, 64, threads=2)
        return float((r.color_hp[0] * gc).sum() + (r.alpha_hp[0] * ga).sum())

    with oracle.inputs64(a64):
        r = oracle.OracleRun(scene, [cam], opts, 64, threads=2)
        g = r.backward([gc.astype(np.float32)], [ga.astype(np.float32)])
    h = 1e-6
    bad = []
    for key, width in (("means", 3), ("log_scales", 3), ("quaternions", 4)):
        for i in range(n):
            for j in range(width):
                a1 = {k: v.copy() for k, v in a64.items()}
                a2 = {k: v.copy() for k, v in a64.items()}
                a1[key][i, j] += h
                a2[key][i, j] -= h
                fd = (loss(a1) - loss(a2)) / (2 * h)
                an = float(g[key][i, j])
                if abs(an - fd) > 1e-5 * max(abs(fd), 1e-3):
                    bad.append((key, i, j, an, fd))
    assert not bad, bad


def _render_one(mean2d, conic, op, rgb, bg, W, H):
    """Colour [3,H,W] and alpha [H,W] of one Gaussian over the background, float64, with the
    blend's rules: power > 0 or alpha < 1/255 skipped, alpha clipped at 0.99."""
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    dx, dy = mean2d[0] - xs, mean2d[1] - ys
    power = -0.5 * (conic[0] * dx * dx + conic[2] * dy * dy) - conic[1] * dx * dy
    alpha = np.minimum(0.99, op * np.exp(power))
    alpha = np.where((power > 0) | (alpha < 1.0 / 255.0), 0.0, alpha)
    color = alpha[None] * np.asarray(rgb)[:, None, None] + (1.0 - alpha)[None] * np.asarray(bg)[:, None, None]
    return color, alpha


def test_refine_weight_single_gaussian_closed_form(oracle):
    """One Gaussian: refine = sum_p |(dL_p/dmx W, dL_p/dmy H)| / max(1 - T_final, 1e-5), with the
    per-pixel dL_p/dmean2d from central differences of a numpy render of the Gaussian (mean2d,
    conic and opacity from the numpy EWA, colour from the degree-0 SH closed form)."""
    from photara_studio_amd.scenes import look_at
    cam = look_at((0.0, 0.0, 3.0), width=64, height=48)
    W, H = cam.width, cam.height
    arrays = {
        "means": np.array([[0.06, -0.04, 0.1]], np.float32),
        "log_scales": np.log(np.array([[0.05, 0.09, 0.07]], np.float32)),
        "quaternions": np.array([[0.9, 0.2, -0.3, 0.25]], np.float32),
        "opacity_logits": np.array([[0.4]], np.float32),
        "sh": np.zeros((1, 16, 3), np.float32),
    }
    arrays["sh"][0, 0] = [0.7, -0.2, 0.4]
    opts = training_options()
    rng = np.random.default_rng(17)
    gc = rng.normal(size=(3, H, W))
    ga = rng.normal(size=(H, W))
    a64 = {k: v.astype(np.float64) for k, v in arrays.items()}
    scene = oracle.Scene(arrays)
    with oracle.inputs64(a64):
        r = oracle.OracleRun(scene, [cam], opts, 64, threads=1)
        g = r.backward([gc.astype(np.float32)], [ga.astype(np.float32)])
    mean2d, S2, op = screen_covariances(arrays, cam, float(np.float32(0.3)))[0]
    assert cap_factor(S2, W, H) == 1.0
    ci = np.linalg.inv(S2)
    conic = (ci[0, 0], ci[0, 1], ci[1, 1])
    rgb = np.maximum(0.28209479177387814 * a64["sh"][0, 0] + 0.5, 0.0)
    bg = np.zeros(3)
    color, alpha = _render_one(mean2d, conic, op, rgb, bg, W, H)
    np.testing.assert_allclose(r.color_hp[0], color, atol=1e-9)
    assert 0.05 < alpha.max() < 0.99

    def pixel_loss(m):
        c, a = _render_one(m, conic, op, rgb, bg, W, H)
        return (c * gc).sum(0) + a * ga, a

    h = 1e-6
    grads, stable = [], np.ones((H, W), bool)
    for axis in range(2):
        e = np.zeros(2)
        e[axis] = h
        lp, ap = pixel_loss(mean2d + e)
        lm, am = pixel_loss(mean2d - e)
        grads.append((lp - lm) / (2 * h))
        stable &= (ap > 0) == (am > 0)  # no pixel crosses the 1/255 cut inside the difference
    assert stable.sum() >= stable.size - 4
    dmx, dmy = grads[0] * stable, grads[1] * stable
    den = np.maximum(alpha, 1e-5)  # 1 - T_final of a single Gaussian is its alpha
    ref = float((np.sqrt((dmx * W) ** 2 + (dmy * H) ** 2) / den)[alpha > 0].sum())
    got = float(g["refine_weight"][0])
    assert abs(got - ref) <= 1e-6 * ref, (got, ref)
    # a dropped denominator or unscaled gradient would be far off
    assert abs(float(np.sqrt((dmx * W) ** 2 + (dmy * H) ** 2).sum()) - ref) > 0.5 * ref
    assert abs(float((np.sqrt(dmx ** 2 + dmy ** 2) / den)[alpha > 0].sum()) - ref) > 0.5 * ref


def test_l1_loss_matches_numpy_closed_form(oracle):
    """loss = w sum|C - C_hat| / (3 P); grad = ((w / P) / 3) sign(C - C_hat) in float, sign(0) = 0."""
    rng = np.random.default_rng(23)
    H, W = 37, 53
    P = H * W
    color = rng.uniform(0, 1, (3, H, W)).astype(np.float32)
    target = rng.uniform(0, 1, (3, H, W)).astype(np.float32)
    target[:, :5, :7] = color[:, :5, :7]  # exact ties: zero gradient there
    w = 0.8
    loss, grad = oracle.l1_loss(color, target, w)
    diff = (color - target).astype(np.float64)  # the float difference, summed in double
    ref_loss = float(np.float32(w)) * np.abs(diff).sum() / (3 * P)  # the weight crosses as a float
    assert abs(loss - ref_loss) <= 1e-12 * ref_loss
    step = np.float32(np.float32(np.float32(w) * np.float32(1.0 / np.float32(P))) / np.float32(3.0))
    ref_grad = (step * np.sign(diff)).astype(np.float32)
    np.testing.assert_array_equal(grad, ref_grad)
    assert (grad[:, :5, :7] == 0).all()
    # normalised by P instead of 3P would be 3x off
    assert abs(3 * ref_loss - loss) > ref_loss
