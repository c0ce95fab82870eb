"""The quantised SH Adam on the GPU: the standalone kernel against the oracle bit for bit (same
operations, same inputs: parameters, fp16 updates, codes and bounds), and the SH step applied
by Rasterizer.backward without a structure update (the reference's unfused step)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import small_scene, training_options
from test_oracle_shq import corrections, schedule

pytestmark = pytest.mark.gpu


def _gpu_state(st, cuda):
    from photara_studio_amd.rasterizer import ShAdamQuant
    return ShAdamQuant(torch.from_numpy(st.first.view(np.float16).copy()).to(cuda),
                       torch.from_numpy(st.packed.copy()).to(cuda), torch.from_numpy(st.bounds.copy()).to(cuda),
                       st.stride)


def _assert_same(gp, gs, p, st, label):
    assert np.array_equal(gp.cpu().numpy(), p), (label, "parameter")
    assert np.array_equal(gs.first.cpu().numpy().view(np.uint16), st.first), (label, "first")
    assert np.array_equal(gs.packed.cpu().numpy(), st.packed), (label, "packed")
    assert np.array_equal(gs.bounds.cpu().numpy(), st.bounds), (label, "bounds")


@pytest.mark.parametrize("stride", [3, 12, 27, 48])
def test_standalone_matches_oracle_over_reference_schedule(cuda, oracle, stride):
    from photara_studio_amd.rasterizer import sh_adam_quant_step
    p0, steps = schedule(stride)
    p = p0.copy()
    st = oracle.ShQuantState(*p.shape)
    gp = torch.from_numpy(p0.copy()).to(cuda)
    gs = _gpu_state(st, cuda)
    for step, active, g, reg in steps:
        print(f"[shq] stride {stride} step {step}", flush=True)
        c1, c2 = corrections(step)
        oracle.sh_adam_quant_step(p, g, st, active, 0.0025, 0.000125, 0.9, 0.999, c1, c2, 1e-15, reg)
        sh_adam_quant_step(gp, torch.from_numpy(g).to(cuda), gs, active, 0.0025, 0.000125, 0.9, 0.999, c1, c2,
                           1e-15, reg)
        torch.cuda.synchronize()
        _assert_same(gp, gs, p, st, (stride, step))


def test_standalone_many_blocks(cuda, oracle):
    from photara_studio_amd.rasterizer import sh_adam_quant_step
    rng = np.random.default_rng(9)
    rows, stride = 3000, 48   # 12 blocks, the last one partial
    p = rng.normal(0, 0.1, (rows, stride)).astype(np.float32)
    st = oracle.ShQuantState(rows, stride)
    gp, gs = torch.from_numpy(p.copy()).to(cuda), _gpu_state(st, cuda)
    for step in range(1, 6):
        print(f"[shq] 3000x48 step {step}", flush=True)
        g = (rng.normal(0, 1, p.shape) * 10.0 ** rng.uniform(-8, -1, p.shape)).astype(np.float32)
        c1, c2 = corrections(step)
        active = 27 if step == 3 else stride
        oracle.sh_adam_quant_step(p, g, st, active, 0.0025, 0.000125, 0.9, 0.999, c1, c2, 1e-15, 1e-4)
        sh_adam_quant_step(gp, torch.from_numpy(g).to(cuda), gs, active, 0.0025, 0.000125, 0.9, 0.999, c1, c2,
                           1e-15, 1e-4)
        torch.cuda.synchronize()
        _assert_same(gp, gs, p, st, step)


def test_backward_applies_sh_step_without_structure_update(cuda, oracle):
    from photara_studio_amd.rasterizer import Rasterizer, SHAdamUpdate, ShAdamQuant, sh_adam_quant_step
    from photara_studio_amd.scenes import model_from, ring_cameras
    arrays = small_scene(3000, 1.0, 4)
    cam = ring_cameras(1, 3.5, [(160, 120)])[0]
    opts = training_options()
    r = Rasterizer()
    grad = torch.from_numpy(np.random.default_rng(1).normal(0, 1e-3, (3, 120, 160)).astype(np.float32)).to(cuda)
    # reference: plain gradients, then the standalone step
    m_ref = model_from(arrays, cuda)
    out = r.forward(m_ref, cam, opts)
    g = r.backward(m_ref, out, grad)
    st_ref = ShAdamQuant.zeros(m_ref.size(), 48, cuda)
    sh_adam_quant_step(m_ref.sh, g.sh, st_ref, 48, 0.0025, 0.000125, 0.9, 0.999, 0.1, 0.001, 1e-15, 0.0)
    # the backward's own SH step
    m = model_from(arrays, cuda)
    out = r.forward(m, cam, opts)
    st = ShAdamQuant.zeros(m.size(), 48, cuda)
    upd = SHAdamUpdate.of(st, lr=0.0025, rest_lr=0.000125, correction1=0.1, correction2=0.001, epsilon=1e-15)
    g2 = r.backward(m, out, grad, sh_adam=upd)
    torch.cuda.synchronize()
    assert g2.sh is None and g2.means is not None
    # the gradients of two backward runs agree to float rounding (atomics order), so the updates
    # agree where the gradient is well above that noise
    strong = g.sh.abs() > 1e-3 * g.sh.abs().max()
    assert strong.sum() > 1000
    assert float((m.sh - m_ref.sh).abs()[strong].max()) <= 1e-6
    moved = (m.sh != torch.from_numpy(arrays["sh"]).to(cuda))
    assert bool(moved[strong].all())
