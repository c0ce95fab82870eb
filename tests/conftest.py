"""Shared pytest setup: the ``gpu`` marker, import paths and common fixtures."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.load()
    return pyoracle


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from photara_studio_amd import _abi
    _abi.load()
    return torch.device("cuda:0")


def training_options(**kw):
    """Raster settings of the training configs: SH degree 3, classic 0.3 px filter, no compensation."""
    from photara_studio_amd.rasterizer import RasterizeOptions
    o = RasterizeOptions(active_sh_degree=3, kernel_size=0.3)
    o.rasterization.compensate_2d = False
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def small_scene(n=2000, box=1.0, seed=0):
    from photara_studio_amd.scenes import synthetic_arrays
    return synthetic_arrays(n, box, seed)
