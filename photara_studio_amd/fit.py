"""Training through the autograd path: torch.optim.Adam over ``rasterize`` (the HIP forward and
the recomputed-transmittance HIP backward behind a torch.autograd.Function), L1 loss against
targets rendered from a seeded ground-truth Gaussian set, PSNR recorded against the step.

    python -m photara_studio_amd.fit --gaussians 500000 --size 1024 --steps 2000

runs the 500k-Gaussian 1024x1024 training workload and prints one PSNR line every `--every`
steps plus a final JSON summary. The per-group learning rates are the reference trainer's
defaults (photara/include/splat/options.hpp:242-258).
"""
from __future__ import annotations

import argparse
import json
import math
import time
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from .autograd import rasterize
from .rasterizer import Camera, RasterizeOptions, Rasterizer
from .train import LearningRates

PARAMS = ("means", "log_scales", "quaternions", "opacity_logits", "sh")


def psnr(img: torch.Tensor, ref: torch.Tensor) -> float:
    mse = float(torch.mean((img.clamp(0.0, 1.0) - ref.clamp(0.0, 1.0)) ** 2))
    return float("inf") if mse == 0.0 else -10.0 * math.log10(mse)


def fit(params: Dict[str, torch.Tensor], cameras: Sequence[Camera], targets: Sequence[torch.Tensor],
        options: RasterizeOptions, steps: int, lrs: Optional[LearningRates] = None, every: int = 100,
        log=None) -> List[Tuple[int, float]]:
    """Trains `params` (leaf tensors, updated in place) for `steps` steps, one view per step in
    round-robin order; returns [(step, mean PSNR over the views)] at step 0, every `every` steps
    and at the end."""
    lrs = lrs or LearningRates()
    sh = params["sh"]
    sh0 = sh[:, :1].detach().clone().requires_grad_(True)
    rest = sh[:, 1:].detach().clone().requires_grad_(True)
    leaves = {k: params[k].detach().clone().requires_grad_(True) for k in PARAMS if k != "sh"}
    opt = torch.optim.Adam([
        {"params": [leaves["means"]], "lr": lrs.means},
        {"params": [leaves["log_scales"]], "lr": lrs.scales},
        {"params": [leaves["quaternions"]], "lr": lrs.quaternions},
        {"params": [leaves["opacity_logits"]], "lr": lrs.opacities},
        {"params": [sh0], "lr": lrs.sh0},
        {"params": [rest], "lr": lrs.sh_rest},
    ], betas=(0.9, 0.999), eps=1e-15)
    rast = Rasterizer()

    def render(k):
        full = torch.cat([sh0, rest], dim=1)
        return rasterize(leaves["means"], leaves["log_scales"], leaves["quaternions"], leaves["opacity_logits"],
                         full, cameras[k], options, rasterizer=rast)

    def evaluate() -> float:
        with torch.no_grad():
            return sum(psnr(render(k)[0], targets[k]) for k in range(len(cameras))) / len(cameras)

    curve = [(0, evaluate())]
    if log:
        log(f"[fit] step {0:5d}  psnr {curve[-1][1]:.3f} dB")
    for step in range(1, steps + 1):
        k = (step - 1) % len(cameras)
        color, _ = render(k)
        loss = torch.mean(torch.abs(color - targets[k]))
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        if step % every == 0 or step == steps:
            curve.append((step, evaluate()))
            if log:
                log(f"[fit] step {step:5d}  psnr {curve[-1][1]:.3f} dB  loss {float(loss):.5f}")
    with torch.no_grad():
        for key in PARAMS:
            if key == "sh":
                params[key].copy_(torch.cat([sh0, rest], dim=1))
            else:
                params[key].copy_(leaves[key])
    return curve


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gaussians", type=int, default=500_000)
    p.add_argument("--size", type=int, default=1024)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--every", type=int, default=100)
    args = p.parse_args()
    from .scenes import look_at, synthetic_torch
    dev = torch.device("cuda:0")
    cams = [look_at((0.0, 0.0, 4.0), width=args.size, height=args.size)]
    opts = RasterizeOptions(active_sh_degree=3, kernel_size=0.3)
    opts.rasterization.compensate_2d = False
    gt = synthetic_torch(args.gaussians, 1.0, 1, dev)
    from .scenes import model_from
    r = Rasterizer()
    targets = [res.color.clone() for res in r.forward_views(model_from(gt, dev), cams, opts)]
    params = {k: v.clone() for k, v in synthetic_torch(args.gaussians, 1.0, 2, dev).items()}
    t0 = time.perf_counter()
    curve = fit(params, cams, targets, opts, args.steps, every=args.every, log=lambda m: print(m, flush=True))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"workload": f"autograd fit {args.gaussians} gaussians {args.size}x{args.size}",
                      "steps": args.steps, "seconds": round(el, 2), "psnr_curve": curve}), flush=True)


if __name__ == "__main__":
    main()
