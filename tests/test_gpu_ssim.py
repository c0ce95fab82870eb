"""The fused L1 + SSIM kernel against the float oracle (same operations in the same order: the
gradient is expected bit for bit, checked to 1e-6 relative; the loss is summed in a different
order, 1e-5 relative), with masks, accumulation and the size limit."""
from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _images(h, w, seed, mask=False):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, size=(3, h, w)).astype(np.float32)
    y = np.clip(x + rng.normal(0, 0.1, size=x.shape), 0, 1).astype(np.float32)
    m = (rng.uniform(size=(h, w)) > 0.2).astype(np.float32) if mask else None
    return x, y, m


@pytest.mark.parametrize("h,w,masked", [(37, 53, False), (270, 480, True), (1080, 1920, False), (11, 11, False)])
def test_l1_ssim_matches_oracle(cuda, oracle, h, w, masked):
    from photara_studio_amd.loss import l1_ssim_loss
    x, y, m = _images(h, w, h + w, masked)
    print(f"[ssim] {h}x{w} mask={masked}", flush=True)
    tm = None if m is None else torch.from_numpy(m).to(cuda)
    loss, grad = l1_ssim_loss(torch.from_numpy(x).to(cuda), torch.from_numpy(y).to(cuda), tm, 0.2, 0.9)
    want_loss, want_grad = oracle.l1_ssim(x, y, m, 0.2, 0.9, precision=32)
    got = grad.cpu().numpy()
    exact = bool(np.array_equal(got, want_grad))
    rel = float(np.abs(got - want_grad).max() / max(np.abs(want_grad).max(), 1e-30))
    print(f"  gradient rel err {rel:.2e} (bit-exact: {exact}); loss {float(loss):.8f} vs {want_loss:.8f}")
    assert rel <= 1e-6
    assert abs(float(loss) - want_loss) <= 1e-5 * abs(want_loss)


def test_l1_ssim_accumulates_and_validates(cuda, oracle):
    from photara_studio_amd.loss import l1_ssim_loss
    x, y, _ = _images(40, 60, 7)
    tx, ty = torch.from_numpy(x).to(cuda), torch.from_numpy(y).to(cuda)
    base = torch.full_like(tx, 0.25)
    _, g = l1_ssim_loss(tx, ty, grad=base.clone(), accumulate=True)
    _, g1 = l1_ssim_loss(tx, ty)
    np.testing.assert_allclose(g.cpu().numpy(), 0.25 + g1.cpu().numpy(), rtol=0, atol=1e-7)
    with pytest.raises(ValueError):
        l1_ssim_loss(tx[:, :10, :].contiguous(), ty[:, :10, :].contiguous())
