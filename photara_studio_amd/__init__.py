"""MI355X (gfx950) 3D Gaussian Splatting training step.

Hand-written HIP kernels behind a C ABI (``include/pht_splat.h``), with a Python mirror of the
reference rasterizer API and a torch.autograd.Function. The kernels live in
``_lib/libpht_splat.so`` (built in-tree by ``photara_studio_amd.build``); nothing here falls back
to a CPU implementation.
"""
from .rasterizer import (Camera, GaussianModel, ModelGradients, RasterizationPolicy, RasterizeOptions,
                         Rasterizer, RenderResult, SHAdamUpdate, ShAdamQuant, StructureAdamUpdate,
                         configure_rasterization)
from .autograd import rasterize
from . import checkpoint, densify

__all__ = [
    "Camera", "GaussianModel", "ModelGradients", "RasterizationPolicy", "RasterizeOptions", "Rasterizer",
    "RenderResult", "SHAdamUpdate", "ShAdamQuant", "StructureAdamUpdate", "checkpoint", "configure_rasterization",
    "densify", "rasterize",
]
