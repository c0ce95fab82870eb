"""The whole training step: the fused forward + loss + backward + Adam of TrainStep against the
oracle's training step, eager against captured-graph replays, and the L1+SSIM objective.

Adam's first updates are nearly lr * sign(g), so parameters are compared where the gradient is
well above float noise (|m| > 1e-3 max|m|, m = the first moment); the moments themselves are
compared as gradients (1e-4 relative to the tensor's max)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import small_scene, training_options

pytestmark = pytest.mark.gpu

KEYS = ("means", "log_scales", "quaternions", "opacity_logits", "sh")
MOMENTS = {"means": "means", "log_scales": "scales", "quaternions": "rotations", "opacity_logits": "opacity",
           "sh": "sh"}


def _setup(cuda, oracle, n=3000, seed=11, loss="l1", sh_optimizer="fp32"):
    from photara_studio_amd.scenes import model_from, ring_cameras
    from photara_studio_amd.train import LearningRates, TrainStep
    arrays = small_scene(n, 1.0, seed)
    cams = ring_cameras(2, 3.5, [(160, 120)])
    opts = training_options()
    gt = oracle.OracleRun(oracle.Scene(small_scene(n, 1.0, seed + 1)), cams, opts, 32)
    targets = [c.copy() for c in gt.color]
    lrs = LearningRates.uniform(1e-3)
    step = TrainStep(model_from(arrays, cuda), cams, [torch.from_numpy(t).to(cuda) for t in targets], opts, lrs,
                     loss=loss, sh_optimizer=sh_optimizer)
    return arrays, cams, opts, targets, lrs, step


def _compare_to_oracle(step, trainer, label):
    for k in KEYS:
        m_gpu = getattr(step.adam, MOMENTS[k] + "_first").cpu().numpy()
        m_orc = trainer.m[k]
        scale = max(float(np.abs(m_orc).max()), 1e-30)
        rel = float(np.abs(m_gpu - m_orc).max()) / scale
        assert rel <= 1e-4, (label, k, "first moment", rel)
        p_gpu = getattr(step.model, k).cpu().numpy()
        p_orc = trainer.scene.a[k]
        strong = np.abs(m_orc) > 1e-3 * scale
        err = float(np.abs(p_gpu - p_orc)[strong].max()) if strong.any() else 0.0
        assert err <= 1e-6, (label, k, "parameters", err)


@pytest.mark.parametrize("loss", ["l1", "l1_ssim"])
def test_fused_step_matches_oracle_train_step(cuda, oracle, loss):
    arrays, cams, opts, targets, lrs, step = _setup(cuda, oracle, loss=loss)
    trainer = oracle.OracleTrainer(oracle.Scene(arrays), cams, opts, targets, lrs, loss=loss)
    for k in range(2):
        print(f"[train] {loss} step {k + 1}: GPU eager vs oracle", flush=True)
        step.step_eager()
        want_loss = trainer.step()
        torch.cuda.synchronize()
        assert abs(step.loss() - want_loss) <= 1e-5 * abs(want_loss), (step.loss(), want_loss)
        _compare_to_oracle(step, trainer, f"{loss} step {k + 1}")


def test_fused_quantised_sh_step_matches_oracle(cuda, oracle):
    """The trainer's SH optimizer (fp16 update + 8-bit log-second codes) fused into the backward,
    against the oracle's training step with the same state. The SH gradients of the two agree to
    float rounding (1e-4 relative), so the comparison is on coefficients whose gradient is well
    above that noise: their fp16 updates agree to one fp16 step on the first step, and their
    parameters to a small fraction of the learning rate on both steps (the row bounds also see the
    noise-level columns, which moves codes by at most one step)."""
    arrays, cams, opts, targets, lrs, step = _setup(cuda, oracle, sh_optimizer="quant")
    trainer = oracle.OracleTrainer(oracle.Scene(arrays), cams, opts, targets, lrs, sh_optimizer="quant")
    # its own copy of the parameters: a Scene updates the arrays it is given in place
    ref32 = oracle.OracleTrainer(oracle.Scene({k: v.copy() for k, v in arrays.items()}), cams, opts, targets, lrs)
    for k in range(2):
        print(f"[train] quantised SH step {k + 1}", flush=True)
        step.step_eager()
        want = trainer.step()
        if k == 0:
            ref32.step()
        torch.cuda.synchronize()
        assert abs(step.loss() - want) <= 1e-5 * abs(want)
        for key in KEYS[:-1]:   # structure parameters: FP32 moments, as in the L1 test
            m_gpu = getattr(step.adam, MOMENTS[key] + "_first").cpu().numpy()
            scale = max(float(np.abs(trainer.m[key]).max()), 1e-30)
            assert float(np.abs(m_gpu - trainer.m[key]).max()) / scale <= 1e-4, (k, key)
        # SH: strong where the first step's FP32 moment is well above float noise
        m1 = ref32.m["sh"].reshape(step.sh_state.first.shape)
        strong = np.abs(m1) > 1e-3 * np.abs(m1).max()
        assert strong.sum() > 1000
        sh_gpu = step.model.sh.cpu().numpy().reshape(strong.shape)
        sh_orc = trainer.scene.a["sh"].reshape(strong.shape)
        err = np.abs(sh_gpu - sh_orc)[strong].max()
        assert err <= 0.05 * 1e-3, (k, float(err))
        if k == 0:
            u_gpu = step.sh_state.first.cpu().numpy().astype(np.float32)
            u_orc = trainer.sh_state.first_f32()
            assert np.abs(u_gpu - u_orc)[strong].max() <= 2.0 ** -10 * np.abs(u_orc)[strong].max()


def test_forward_l1_matches_forward_then_l1_loss(cuda, oracle):
    """pht_forward_l1 (the L1 loss fused into the blend's epilogue) against pht_forward followed by
    pht_l1_loss per view, through the C ABI: the same colours and the same gradient bits; the
    loss sums agree to float rounding (they are added in another order)."""
    import ctypes as C
    from photara_studio_amd import _abi
    _, cams, _, _, _, step = _setup(cuda, oracle)
    lib, nv, weight = step.lib, len(step.cameras), 0.8
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    loss_fused = torch.zeros((1,), device=cuda)
    print("[train] pht_forward_l1 vs pht_forward + pht_l1_loss", flush=True)
    assert lib.pht_forward_l1(step.ws.handle, C.byref(step.gauss), step.cams, nv, C.byref(step.settings), step.outs,
                              step._target_ptrs, step._grad_ptrs, C.c_void_p(loss_fused.data_ptr()), weight, 1,
                              stream) == _abi.PHT_OK
    torch.cuda.synchronize()
    colors_fused = [c.clone() for c in step.colors]
    grads_fused = [g.clone() for g in step.grad_colors]
    assert lib.pht_forward(step.ws.handle, C.byref(step.gauss), step.cams, nv, C.byref(step.settings), step.outs, 1,
                           stream) == _abi.PHT_OK
    loss_sep = torch.zeros((1,), device=cuda)
    grads_sep = [torch.full_like(g, 7.0) for g in step.grad_colors]
    for k in range(nv):
        assert lib.pht_l1_loss(C.c_void_p(step.colors[k].data_ptr()), C.c_void_p(step.targets[k].data_ptr()),
                               C.c_void_p(grads_sep[k].data_ptr()), C.c_void_p(loss_sep.data_ptr()),
                               step.colors[k][0].numel(), weight, stream) == _abi.PHT_OK
    torch.cuda.synchronize()
    for k in range(nv):
        assert torch.equal(step.colors[k], colors_fused[k]), k
        assert torch.equal(grads_sep[k], grads_fused[k]), k
    assert float(loss_sep) > 0.0
    assert abs(float(loss_fused) - float(loss_sep)) <= 1e-6 * float(loss_sep), (float(loss_fused), float(loss_sep))
    # a missing target is refused before any launch
    bad = (C.c_void_p * nv)(*([None] * nv))
    assert lib.pht_forward_l1(step.ws.handle, C.byref(step.gauss), step.cams, nv, C.byref(step.settings), step.outs,
                              bad, step._grad_ptrs, None, weight, 0, stream) == _abi.PHT_EINVAL


@pytest.mark.parametrize("sh_optimizer", ["fp32", "quant"])
def test_graph_replays_match_eager_steps(cuda, oracle, sh_optimizer):
    _, _, _, _, _, eager = _setup(cuda, oracle, sh_optimizer=sh_optimizer)
    _, _, _, _, _, graph = _setup(cuda, oracle, sh_optimizer=sh_optimizer)
    graph.capture()
    for k in range(4):
        print(f"[train] replay {k + 1}", flush=True)
        eager.step_eager()
        graph.step_graph()
        # unrelated eager work and a synchronisation between replays
        scratch = torch.randn(1 << 16, device=cuda)
        scratch.mul_(2.0).sum().item()
        torch.cuda.synchronize()
        if k == 0:  # identical parameters: the first forward must agree bit for bit
            for a, b in zip(eager.colors, graph.colors):
                assert torch.equal(a, b)
        assert abs(eager.loss() - graph.loss()) <= 1e-5 * abs(eager.loss())
        for key in KEYS:
            if key == "sh" and sh_optimizer == "quant":
                # the stored update is ~sign(g) at noise-level gradients (which the two runs' atomics
                # orders decide differently): a replay with wrong hyper-parameters would move every
                # coefficient, so require the runs to agree on all but 1% and stay within the
                # per-step bound everywhere
                d = (eager.model.sh - graph.model.sh).abs()
                assert float((d <= 1e-5).float().mean()) >= 0.99, (k, "sh")
                assert float(d.max()) <= 2 * 3e-3 * (k + 1), (k, "sh")
                continue
            me = getattr(eager.adam, MOMENTS[key] + "_first")
            mg = getattr(graph.adam, MOMENTS[key] + "_first")
            rel = float((me - mg).abs().max() / me.abs().max().clamp_min(1e-30))
            assert rel <= 1e-3, (k, key, rel)
            strong = me.abs() > 1e-2 * me.abs().max()
            pe, pg = getattr(eager.model, key), getattr(graph.model, key)
            assert float((pe - pg).abs()[strong].max()) <= 1e-5, (k, key)


def test_ssim_step_gradient_is_the_loss_gradient(cuda, oracle):
    from photara_studio_amd.loss import l1_ssim_loss
    _, _, _, _, _, step = _setup(cuda, oracle, loss="l1_ssim")
    step.step_eager()
    torch.cuda.synchronize()
    total = 0.0
    for c, t, g in zip(step.colors, step.targets, step.grad_colors):
        loss, grad = l1_ssim_loss(c, t, None, 0.2, 1.0)
        assert torch.equal(grad, g)
        total += float(loss)
    assert abs(total - step.loss()) <= 1e-6 * abs(total)


@pytest.mark.parametrize("mode", ["graph", "eager"])
def test_overflow_mid_training_grows_and_reruns(cuda, oracle, mode):
    """The instance list outgrows the workspace between steps (all log-scales grow by 0.4 after
    two steps). The overflowed step leaves the model unchanged; the driver finds it from the step's
    counters, grows the workspace, re-captures and re-runs it with its own step index, so the run
    equals one with ample capacity step for step."""
    _, _, _, _, _, ample = _setup(cuda, oracle)
    _, _, _, _, _, tight = _setup(cuda, oracle)
    need = tight.ensure_capacity()
    from photara_studio_amd.train import TrainStep
    tight = TrainStep(tight.model, tight.cameras, tight.targets, tight.options, tight.lrs,
                      instance_capacity=need + 64, sh_optimizer="fp32")
    cap0 = tight.capacity
    run = tight.step_graph if mode == "graph" else tight.step_eager
    for k in range(4):
        print(f"[train] overflow {mode} step {k + 1}", flush=True)
        if k == 2:
            with torch.no_grad():
                ample.model.log_scales.add_(0.4)
                tight.model.log_scales.add_(0.4)
        ample.step_eager()
        run()
        assert abs(ample.loss() - tight.loss()) <= 1e-5 * abs(ample.loss()), k
        assert tight.step_index == ample.step_index == k + 1
        for key in KEYS:
            me = getattr(ample.adam, MOMENTS[key] + "_first")
            mt = getattr(tight.adam, MOMENTS[key] + "_first")
            rel = float((me - mt).abs().max() / me.abs().max().clamp_min(1e-30))
            assert rel <= 1e-3, (k, key, rel)
            strong = me.abs() > 1e-2 * me.abs().max()
            pe, pt = getattr(ample.model, key), getattr(tight.model, key)
            assert float((pe - pt).abs()[strong].max()) <= 1e-5, (k, key)
    assert tight.overflow_recoveries == 1 and tight.capacity > cap0
    assert not tight.counts().overflow
