"""Pins of the oracle's fused L1 + SSIM loss (CPU).

- the 11-tap window is the normalised Gaussian with sigma 1.5 (to within 1 ulp);
- identical images give SSIM = 1 everywhere: zero loss and zero gradient;
- the loss equals an independent restatement with scipy's correlate (zero padding, the same
  window, the standard SSIM formula) to double precision, with and without a mask;
- the FP64 gradient equals FP64 central differences of the FP64 loss (h = 1e-3 on the float32
  pixel inputs, 2e-5 relative);
- the float oracle's gradient is within float rounding of the FP64 one.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage


def _images(h, w, seed, mask=False):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, size=(3, h, w)).astype(np.float32)
    y = np.clip(x + rng.normal(0, 0.1, size=x.shape), 0, 1).astype(np.float32)
    m = (rng.uniform(size=(h, w)) > 0.2).astype(np.float32) if mask else None
    return x, y, m


def _window():
    i = np.arange(11) - 5
    g = np.exp(-(i ** 2) / (2 * 1.5 ** 2))
    return g / g.sum()


# the window table of the implementation (fused_ssim.cu:21-24), checked against _window() below
WINDOW = np.array([0.001028380123898387, 0.0075987582094967365, 0.036000773310661316, 0.10936068743467331,
                   0.21300552785396576, 0.26601171493530273, 0.21300552785396576, 0.10936068743467331,
                   0.036000773310661316, 0.0075987582094967365, 0.001028380123898387], np.float32)


def _scipy_loss(x, y, m, ssim_w, phot):
    ssim_w, phot = float(np.float32(ssim_w)), float(np.float32(phot))  # the weights are float arguments
    x = x.astype(np.float64)
    y = y.astype(np.float64)
    if m is not None:
        x, y = x * m, y * m
    w1 = WINDOW.astype(np.float64)
    k = np.outer(w1, w1)
    # the constants are float products, 0.01f * 0.01f and 0.03f * 0.03f (fused_ssim.cu:533-534)
    c1 = np.float64(np.float32(0.01) * np.float32(0.01))
    c2 = np.float64(np.float32(0.03) * np.float32(0.03))
    _, h, w = x.shape
    total = 0.0
    for ch in range(3):
        f = lambda a: ndimage.correlate(a, k, mode="constant", cval=0.0)
        mu1, mu2 = f(x[ch]), f(y[ch])
        s11, s22, s12 = f(x[ch] * x[ch]) - mu1 ** 2, f(y[ch] * y[ch]) - mu2 ** 2, f(x[ch] * y[ch]) - mu1 * mu2
        ssim = ((2 * mu1 * mu2 + c1) * (2 * s12 + c2)) / ((mu1 ** 2 + mu2 ** 2 + c1) * (s11 + s22 + c2))
        v = ssim_w * (1 - ssim) + (1 - ssim_w) * np.abs(x[ch] - y[ch])
        total += v[5:h - 5, 5:w - 5].sum()
    return total * float(np.float32(phot) / np.float32(3 * (w - 10) * (h - 10)))


def test_window_is_normalised_gaussian():
    lit = WINDOW
    # the table was evaluated in float32: within 1 ulp of the correctly rounded window
    np.testing.assert_allclose(lit.astype(np.float64), _window(), rtol=1.2e-7, atol=0)
    assert abs(float(lit.astype(np.float64).sum()) - 1.0) < 1e-6


def test_identical_images_zero_loss(oracle):
    x, _, _ = _images(30, 40, 0)
    loss, grad = oracle.l1_ssim(x, x, precision=64)
    assert abs(loss) < 1e-12
    assert np.abs(grad).max() < 1e-12


@pytest.mark.parametrize("masked", [False, True])
def test_loss_matches_scipy_restatement(oracle, masked):
    x, y, m = _images(37, 53, 1, masked)
    loss, _ = oracle.l1_ssim(x, y, m, 0.2, 0.8, precision=64)
    want = _scipy_loss(x, y, m, 0.2, 0.8)
    assert abs(loss - want) <= 1e-12 * max(1.0, abs(want)), (loss, want)


@pytest.mark.parametrize("masked", [False, True])
def test_fp64_gradient_matches_central_differences(oracle, masked):
    x, y, m = _images(24, 20, 2, masked)
    _, grad = oracle.l1_ssim(x, y, m, 0.2, 1.0, precision=64)
    rng = np.random.default_rng(3)
    h = 1e-3  # pixel values are float32 inputs: perturb on a float32-representable step
    for _ in range(40):
        c, i, j = int(rng.integers(0, 3)), int(rng.integers(0, 24)), int(rng.integers(0, 20))
        xp, xm = x.copy(), x.copy()
        xp[c, i, j] += np.float32(h)
        xm[c, i, j] -= np.float32(h)
        if xp[c, i, j] == y[c, i, j] or xm[c, i, j] == y[c, i, j] or \
                np.sign(xp[c, i, j] - y[c, i, j]) != np.sign(xm[c, i, j] - y[c, i, j]):
            continue  # the L1 kink lies inside the difference interval
        step = float(xp[c, i, j]) - float(xm[c, i, j])
        fd = (oracle.l1_ssim(xp, y, m, 0.2, 1.0, precision=64)[0] -
              oracle.l1_ssim(xm, y, m, 0.2, 1.0, precision=64)[0]) / step
        # central differences carry an O(h^2) truncation error of ~1e-6 relative at h = 1e-3
        assert abs(fd - grad[c, i, j]) <= 2e-5 * max(abs(fd), 1e-4), (c, i, j, fd, grad[c, i, j])


def test_float_gradient_close_to_double(oracle):
    x, y, m = _images(64, 80, 4, True)
    l32, g32 = oracle.l1_ssim(x, y, m, 0.2, 1.0, precision=32)
    l64, g64 = oracle.l1_ssim(x, y, m, 0.2, 1.0, precision=64)
    assert abs(l32 - l64) <= 1e-5 * abs(l64)
    assert np.abs(g32 - g64).max() <= 1e-4 * np.abs(g64).max()
