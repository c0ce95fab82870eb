"""End-to-end training through the autograd path (torch.optim.Adam over the HIP forward and
backward): PSNR against the ground-truth render rises."""
from __future__ import annotations

import pytest
import torch

pytestmark = pytest.mark.gpu


def test_autograd_fit_improves_psnr(cuda):
    from photara_studio_amd.fit import fit
    from photara_studio_amd.rasterizer import RasterizeOptions, Rasterizer
    from photara_studio_amd.scenes import look_at, model_from, synthetic_torch

    cams = [look_at((0.0, 0.0, 4.0), width=256, height=256)]
    opts = RasterizeOptions(active_sh_degree=3, kernel_size=0.3)
    opts.rasterization.compensate_2d = False
    gt = synthetic_torch(50_000, 1.0, 1, cuda)
    targets = [res.color.clone() for res in Rasterizer().forward_views(model_from(gt, cuda), cams, opts)]
    params = {k: v.clone() for k, v in synthetic_torch(50_000, 1.0, 2, cuda).items()}
    print("[fit] 300 autograd steps", flush=True)
    torch.cuda.synchronize()
    mem = []
    curve = fit(params, cams, targets, opts, steps=300, every=100,
                log=lambda m: (print(m, flush=True), mem.append(torch.cuda.memory_allocated())))
    # one workspace serves every step: device memory does not grow with the step count
    assert max(mem[1:]) - mem[1] <= 64 << 20, mem
    assert [s for s, _ in curve] == [0, 100, 200, 300]
    assert all(torch.isfinite(v).all() for v in params.values())
    assert curve[-1][1] > curve[0][1] + 2.0, curve
