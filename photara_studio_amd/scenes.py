"""Seeded synthetic Gaussian sets and cameras for the training-step workloads.

Distributions (the 10k / 1M / 3M / 6M / 500k workloads): means uniform in a cube, log-scales
N(-4, 0.5), rotations uniform unit quaternions (Shoemake's method), opacity logits N(0, 1),
SH coefficients N(0, 0.3) at degree 3 (16 x 3 per Gaussian), all float32. Cameras are pinhole,
y-down, looking at the origin, with a 60 degree horizontal field of view.
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np
import torch

from .rasterizer import Camera, GaussianModel


def look_at(eye: Sequence[float], target: Sequence[float] = (0.0, 0.0, 0.0), width: int = 256,
            height: int = 256, fov_x_deg: float = 60.0, up: Sequence[float] = (0.0, 1.0, 0.0)) -> Camera:
    """Pinhole camera at ``eye`` looking at ``target``; x right, y down, z forward.

    world_to_camera is stored column-major (reference Camera::world_to_camera, types.hpp:17-34).
    """
    eye = np.asarray(eye, dtype=np.float64)
    z = np.asarray(target, dtype=np.float64) - eye
    z /= np.linalg.norm(z)
    down = -np.asarray(up, dtype=np.float64)
    y = down - np.dot(down, z) * z
    if np.linalg.norm(y) < 1e-9:
        y = np.array([0.0, 0.0, 1.0]) - z[2] * z
    y /= np.linalg.norm(y)
    x = np.cross(y, z)
    R = np.stack([x, y, z])  # rows: camera axes in world coordinates
    t = -R @ eye
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = t
    w2c = [float(M[r, c]) for c in range(4) for r in range(4)]  # column-major
    f = (width / 2.0) / math.tan(math.radians(fov_x_deg) / 2.0)
    return Camera(world_to_camera=w2c, position=[float(v) for v in eye], fx=f, fy=f, cx=width / 2.0,
                  cy=height / 2.0, width=int(width), height=int(height))


def ring_cameras(n: int, radius: float, resolutions: Sequence[Tuple[int, int]], height: float = 0.0) -> List[Camera]:
    """n cameras on a ring in the XZ plane looking at the origin."""
    cams = []
    for k in range(n):
        a = 2.0 * math.pi * k / n
        eye = (radius * math.sin(a), height, radius * math.cos(a))
        w, h = resolutions[k % len(resolutions)]
        cams.append(look_at(eye, width=w, height=h))
    return cams


def random_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform unit quaternions, scalar first (Shoemake, Graphics Gems III)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([b * np.cos(2 * np.pi * u3), a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3)], axis=1)
    return q.astype(np.float32)


def synthetic_arrays(n: int, box: float, seed: int, sh_bases: int = 16) -> dict:
    """Seeded numpy arrays of one Gaussian set (float32)."""
    rng = np.random.default_rng(seed)
    return {
        "means": rng.uniform(-box, box, size=(n, 3)).astype(np.float32),
        "log_scales": rng.normal(-4.0, 0.5, size=(n, 3)).astype(np.float32),
        "quaternions": random_quaternions(rng, n),
        "opacity_logits": rng.normal(0.0, 1.0, size=(n, 1)).astype(np.float32),
        "sh": rng.normal(0.0, 0.3, size=(n, sh_bases, 3)).astype(np.float32),
    }


def synthetic_torch(n: int, box: float, seed: int, device, sh_bases: int = 16) -> dict:
    """Same distributions drawn with torch's generator on the device (for large sets)."""
    gen = torch.Generator(device=device).manual_seed(seed)
    dev = torch.device(device)
    u = lambda *s: torch.rand(*s, generator=gen, device=dev)
    nrm = lambda *s: torch.randn(*s, generator=gen, device=dev)
    u1, u2, u3 = u(n), u(n), u(n)
    a, b = torch.sqrt(1.0 - u1), torch.sqrt(u1)
    q = torch.stack([b * torch.cos(2 * math.pi * u3), a * torch.sin(2 * math.pi * u2),
                     a * torch.cos(2 * math.pi * u2), b * torch.sin(2 * math.pi * u3)], dim=1)
    return {
        "means": (u(n, 3) * 2.0 - 1.0) * box,
        "log_scales": nrm(n, 3) * 0.5 - 4.0,
        "quaternions": q.contiguous(),
        "opacity_logits": nrm(n, 1),
        "sh": nrm(n, sh_bases, 3) * 0.3,
    }


def model_from(arrays: dict, device, sh_degree: int = 3) -> GaussianModel:
    t = {k: (torch.as_tensor(v) if isinstance(v, np.ndarray) else v).to(device=device, dtype=torch.float32)
         .contiguous() for k, v in arrays.items()}
    return GaussianModel(means=t["means"], log_scales=t["log_scales"], quaternions=t["quaternions"],
                         opacity_logits=t["opacity_logits"], sh=t["sh"], sh_degree=sh_degree)


# Named workloads: (gaussians, box half-width, cameras)
def workload_cameras(name: str) -> List[Camera]:
    if name == "c0":
        return [look_at((0.0, 0.0, 4.0), width=256, height=256)]
    if name == "c1":
        return [look_at((0.0, 0.0, 8.0), width=1920, height=1080)]
    if name == "c2":
        res = [(1920, 1080)] * 2 + [(1280, 720)] * 3 + [(960, 540)] * 3
        return ring_cameras(8, 8.0, res)
    if name == "c3":
        return ring_cameras(4, 8.0, [(1920, 1080)])
    if name == "c4":
        return [look_at((0.0, 0.0, 4.0), width=1024, height=1024)]
    if name.startswith("bench"):
        return ring_cameras(int(name[5:] or 1), 8.0, [(1920, 1080)])
    raise ValueError(name)


WORKLOADS = {
    "c0": (10_000, 1.0),
    "c1": (1_000_000, 3.0),
    "c2": (3_000_000, 3.0),
    "c3": (6_000_000, 3.0),
    "c4": (500_000, 1.0),
}
