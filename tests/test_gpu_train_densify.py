"""Densification inside the batched training step: the statistics the captured / eager step
accumulates equal one accumulate_densification_stats observation per view (its radii, visibility
and refine weight), and a refinement re-binds the step (sizes, workspace, graph) so training goes
on with the remapped optimizer state."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from conftest import small_scene, training_options

pytestmark = pytest.mark.gpu


def _step(cuda, densify=True, n=6000, seed=4):
    from photara_studio_amd.scenes import model_from, ring_cameras
    from photara_studio_amd.train import LearningRates, TrainStep, to_low_precision
    a = small_scene(n, 1.0, seed)
    cams = ring_cameras(3, 3.5, [(160, 120), (128, 96), (200, 150)])
    rng = np.random.default_rng(seed)
    targets = [torch.from_numpy(rng.random((3, c.height, c.width)).astype(np.float32)).to(cuda) for c in cams]
    model = to_low_precision(model_from(a, cuda), torch.bfloat16)
    st = TrainStep(model, cams, targets, training_options(), LearningRates(), densify=densify, view_indices=[4, 9, 2])
    st.ensure_capacity()
    return st


def _per_view_refine(st):
    from photara_studio_amd import _abi
    n, nv = st.model.size(), len(st.cameras)
    out = torch.empty((n, nv, 10), dtype=torch.float32, device=st.model.means.device)
    _abi.check(st.lib.pht_debug_splat_grads(st.ws.handle, C.c_void_p(out.data_ptr()),
                                            C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return [out[:, v, 9].contiguous() for v in range(nv)]


def test_step_statistics_equal_per_view_observations(cuda):
    from photara_studio_amd import densify as D
    st = _step(cuda)
    for k in range(2):
        print(f"[densify-step] eager step {k}", flush=True)
        before = {f: getattr(st.dstats, f).clone() for f in ("gradient", "count", "max_screen_radius", "priority",
                                                              "max_geometry_gradient", "first_view", "view_support",
                                                              "accumulated_error")}
        st.step_eager()
        torch.cuda.synchronize()
        expect = D.DensificationStats(**{f: t.clone() for f, t in before.items()})
        refine = _per_view_refine(st)
        obs = [D.Observation(refine[v], st.visibility[v], st.radii[v], c.width, c.height, True, True, 1.0, 0.5,
                             refine[v], st.view_indices[v]) for v, c in enumerate(st.cameras)]
        D.accumulate_views(expect, obs)
        torch.cuda.synchronize()
        for f in before:
            assert torch.equal(getattr(st.dstats, f), getattr(expect, f)), (k, f)
    assert float(st.dstats.count.max()) >= 2 and int((st.dstats.view_support == 2).sum()) > 0
    assert set(torch.unique(st.dstats.first_view).tolist()) <= {0, 5, 10, 3}


def test_graph_step_accumulates_like_eager(cuda):
    eager, graph = _step(cuda), _step(cuda)
    for k in range(3):
        print(f"[densify-step] step {k}", flush=True)
        eager.step_eager()
        graph.step_graph()
    torch.cuda.synchronize()
    assert torch.equal(eager.dstats.count, graph.dstats.count)
    assert torch.equal(eager.dstats.first_view, graph.dstats.first_view)
    # the refine statistic sums float atomics: equal to rounding
    for f in ("gradient", "priority", "max_geometry_gradient", "max_screen_radius"):
        a, b = getattr(eager.dstats, f), getattr(graph.dstats, f)
        assert torch.allclose(a, b, rtol=1e-4, atol=1e-9), f


def test_refine_rebinds_and_training_continues(cuda):
    from photara_studio_amd import densify as D
    st = _step(cuda, n=8000)
    for _ in range(3):
        st.step_graph()
    n0 = st.model.size()
    o = D.apply_strategy_defaults(D.DensifyOptions(iterations=1000, densification_cap=12000))
    o.densify_gradient_threshold = 0.0
    counts = st.refine(200, o, 1.0, [0.0, 0.0, 0.0], D.Mt19937(7))
    torch.cuda.synchronize()
    print(f"[densify-step] refine {counts}", flush=True)
    n1 = st.model.size()
    assert counts.grown > 0 and n1 == n0 - counts.pruned + counts.grown
    assert st.dstats.size() == n1 and float(st.dstats.count.abs().sum()) == 0.0
    assert st.adam.means_first.shape[0] == n1 and st.sh_state.first.shape[0] == n1
    assert st.sh_adam.first is st.sh_state.first
    assert st.model.sh_lp.shape[0] == n1 and st.radii[0].shape[0] == n1
    for _ in range(3):
        st.step_graph()
    st.step_eager()
    torch.cuda.synchronize()
    assert np.isfinite(st.loss())
    assert bool(torch.isfinite(st.model.means).all()) and float(st.dstats.count.max()) >= 4
    assert not st.counts().overflow
