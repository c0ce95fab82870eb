"""View groups (TrainStep(view_groups=k), pht_backward_groups, pht_step_graph_create_groups): the
step's views split over several workspaces and streams, one per-Gaussian backward + Adam over all
of them. The update is the single-group step's: the same images bit for bit, and moments and
parameters equal up to the float order of the blend backward's atomics (the tolerance of the
graph-against-eager test). Also: an instance list that outgrows one group's workspace mid-run."""
from __future__ import annotations

import pytest
import torch

from conftest import small_scene, training_options

pytestmark = pytest.mark.gpu

KEYS = ("means", "log_scales", "quaternions", "opacity_logits", "sh")
MOMENTS = {"means": "means", "log_scales": "scales", "quaternions": "rotations", "opacity_logits": "opacity",
           "sh": "sh"}


def _pair(cuda, groups, loss="l1", n=4000, seed=21, capacity=None):
    from photara_studio_amd.scenes import model_from, ring_cameras
    from photara_studio_amd.train import LearningRates, TrainStep
    arrays = small_scene(n, 1.0, seed)
    cams = ring_cameras(5, 3.5, [(160, 120), (128, 96), (200, 150)])
    g = torch.Generator(device="cpu").manual_seed(seed)
    targets = [torch.rand((3, c.height, c.width), generator=g).to(cuda) for c in cams]
    opts = training_options()
    lrs = LearningRates.uniform(1e-3)
    one = TrainStep(model_from(arrays, cuda), cams, targets, opts, lrs, loss=loss, sh_optimizer="fp32")
    grouped = TrainStep(model_from(arrays, cuda), cams, targets, opts, lrs, loss=loss, sh_optimizer="fp32",
                        view_groups=groups, instance_capacity=capacity)
    assert [nk for _, nk in grouped.groups] == [len(cams) // groups + (1 if k < len(cams) % groups else 0)
                                                for k in range(groups)]
    return one, grouped


def _compare(one, grouped, label):
    assert abs(one.loss() - grouped.loss()) <= 1e-5 * abs(one.loss()), label
    assert one.step_index == grouped.step_index, label
    for key in KEYS:
        m1 = getattr(one.adam, MOMENTS[key] + "_first")
        mg = getattr(grouped.adam, MOMENTS[key] + "_first")
        rel = float((m1 - mg).abs().max() / m1.abs().max().clamp_min(1e-30))
        assert rel <= 1e-3, (label, key, rel)
        strong = m1.abs() > 1e-2 * m1.abs().max()
        p1, pg = getattr(one.model, key), getattr(grouped.model, key)
        assert float((p1 - pg).abs()[strong].max()) <= 1e-5, (label, key)


@pytest.mark.parametrize("groups,loss", [(2, "l1"), (3, "l1"), (2, "l1_ssim")])
def test_view_groups_equal_one_group(cuda, groups, loss):
    one, grouped = _pair(cuda, groups, loss)
    for k in range(2):
        print(f"[groups] {groups} groups, {loss}: eager step {k + 1}", flush=True)
        one.step_eager()
        grouped.step_eager()
        torch.cuda.synchronize()
        if k == 0:  # identical parameters: the forward agrees bit for bit
            for a, b in zip(one.colors, grouped.colors):
                assert torch.equal(a, b)
        _compare(one, grouped, f"eager {k + 1}")
    for k in range(3):
        print(f"[groups] {groups} groups, {loss}: graph step {k + 1}", flush=True)
        one.step_graph()
        grouped.step_graph()
        # unrelated eager work and a synchronisation between replays
        scratch = torch.randn(1 << 16, device=cuda)
        scratch.mul_(2.0).sum().item()
        torch.cuda.synchronize()
        _compare(one, grouped, f"graph {k + 1}")
    c1, cg = one.counts(), grouped.counts()
    assert (c1.visible, c1.instances, c1.overflow) == (cg.visible, cg.instances, cg.overflow)


@pytest.mark.parametrize("mode", ["graph", "eager"])
def test_view_group_overflow_grows_and_reruns(cuda, mode):
    """Every log-scale grows by 0.4 after two steps, so the instance lists outgrow the tight
    workspaces; the grouped driver finds the overflow from the step's counters, grows the groups'
    workspaces, re-captures and re-runs the step, and stays equal to the single-group run."""
    one, probe = _pair(cuda, 2)
    need = probe.ensure_capacity()
    del probe
    one, grouped = _pair(cuda, 2, capacity=need)
    caps0 = list(grouped.capacities)
    run = grouped.step_graph if mode == "graph" else grouped.step_eager
    for k in range(4):
        print(f"[groups] overflow {mode} step {k + 1}", flush=True)
        if k == 2:
            with torch.no_grad():
                one.model.log_scales.add_(0.4)
                grouped.model.log_scales.add_(0.4)
        one.step_eager()
        run()
        _compare(one, grouped, f"overflow {mode} {k + 1}")
    assert grouped.overflow_recoveries == 1
    assert any(c > c0 for c, c0 in zip(grouped.capacities, caps0))
    assert not grouped.counts().overflow


def test_view_groups_argument_checks(cuda):
    from photara_studio_amd.scenes import model_from, ring_cameras
    from photara_studio_amd.train import LearningRates, TrainStep
    arrays = small_scene(500, 1.0, 3)
    cams = ring_cameras(2, 3.5, [(64, 48)])
    targets = [torch.zeros((3, 48, 64), device=cuda) for _ in cams]
    with pytest.raises(ValueError):
        TrainStep(model_from(arrays, cuda), cams, targets, training_options(), LearningRates(), view_groups=3)
    with pytest.raises(ValueError):
        TrainStep(model_from(arrays, cuda), cams, targets, training_options(), LearningRates(), view_groups=2,
                  densify=True)
