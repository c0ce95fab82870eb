"""Pins of the oracle's quantised SH Adam (the reference trainer's SH optimizer,
third_party/splat_drender/include/splat_drender/sh_adam_quant.cuh), CPU only.

- log1p / expm1 (this build's own, from log and exp) within 4 ulp of float64;
- the reference's known answer: one step from a zero state with g = 0.25, lr 0.01 moves 1 to 0.99
  (photara/tests/splat_backend_dispatch_test.cpp:103-109);
- the parameter trajectory follows exact (float64, unquantised) Adam over the reference test's
  24-step schedule (photara/tests/splat_test.cpp:1454-1512: strides 3/12/27/48, gradients over
  eight orders of magnitude, active-prefix changes, an all-zero step, NaN/inf injection, L2 from
  step 19), within the codec's resolution;
- columns outside the active prefix keep their parameters bit for bit and their first moment
  (recovered from the renormalised update) to fp16 precision;
- each group's smallest and largest log-second land on codes 0 and 255;
- regularisation r equals adding r * p to the non-DC gradients, bit for bit;
- non-finite gradients zero the state and leave the parameter.
"""
from __future__ import annotations

import numpy as np
import pytest


def ulps(got, want):
    sp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / sp


def test_log1p_expm1_ulp(oracle):
    rng = np.random.default_rng(0)
    x = np.concatenate([np.float32(10) ** rng.uniform(-40, 38.5, 400000).astype(np.float32),
                        rng.uniform(0, 1, 200000), rng.uniform(-0.999, 0, 100000)]).astype(np.float32)
    r, e = oracle.plog1pf(x), np.log1p(x.astype(np.float64))
    m = e != 0
    assert ulps(r[m], e[m]).max() <= 4.0
    assert np.array_equal(oracle.plog1pf(np.float32([0.0, -0.0])), np.float32([0.0, -0.0]))
    y = np.concatenate([rng.uniform(-100, 88.7, 400000), np.float32(10) ** rng.uniform(-40, 0, 100000),
                        -np.float32(10) ** rng.uniform(-40, 0, 100000)]).astype(np.float32)
    r, e = oracle.pexpm1f(y), np.expm1(y.astype(np.float64))
    m = (e != 0) & (np.abs(e) < 3e38)
    assert np.isfinite(r[m]).all() and ulps(r[m], e[m]).max() <= 4.0


def test_reference_known_answer(oracle):
    st = oracle.ShQuantState(2, 3)
    p = np.ones((2, 3), np.float32)
    oracle.sh_adam_quant_step(p, np.full((2, 3), 0.25, np.float32), st, 3, 0.01, 0.01, 0.9, 0.999, 0.1, 0.001, 1e-8)
    assert np.abs(p - 0.99).max() < 1e-4


def schedule(stride, rows=17):
    """The reference test's gradient and active-prefix schedule (splat_test.cpp:1454-1512)."""
    count = rows * stride
    values = (0.05 * np.sin(0.17 * np.arange(count, dtype=np.float32))).astype(np.float32)
    steps = []
    for step in range(1, 25):
        active = 3 if (step <= 3 or step == 15) else min(stride, 12) if step <= 6 else \
            min(stride, 27) if step <= 9 else stride
        i = np.arange(count)
        g = (np.float32(10.0) ** (-2.0 - (i % 9)).astype(np.float32) *
             np.cos(0.11 * (i + step * 7)).astype(np.float32)).astype(np.float32)
        g[i >= count - stride] = 0.0
        if step == 13:
            g[:] = 0.0
        if step == 17:
            g[0] = np.nan
            g[stride + 1] = np.inf
        reg = 0.003 if step > 18 else 0.0
        steps.append((step, active, g.reshape(rows, stride), reg))
    return values.reshape(rows, stride), steps


def corrections(step):
    c1 = float(np.float32(1) - np.float32(0.9) ** np.float32(step))
    c2 = float(np.float32(1) - np.float32(0.999) ** np.float32(step))
    return c1, c2


@pytest.mark.parametrize("stride", [3, 12, 27, 48])
def test_trajectory_follows_exact_adam(oracle, stride):
    lr, rest_lr, eps = 0.0025, 0.000125, 1e-15
    p0, steps = schedule(stride)
    p = p0.copy()
    st = oracle.ShQuantState(*p.shape)
    pe = p0.astype(np.float64)
    m = np.zeros_like(pe)
    v = np.zeros_like(pe)
    lrs = np.where(np.arange(stride) >= 3, rest_lr, lr)[None, :]
    for step, active, g, reg in steps:
        c1, c2 = corrections(step)
        before = p.copy()
        oracle.sh_adam_quant_step(p, g, st, active, lr, rest_lr, 0.9, 0.999, c1, c2, eps, reg)
        # inactive columns: parameters untouched
        assert np.array_equal(p[:, active:], before[:, active:]), step
        # exact Adam on the active columns, same non-finite rules
        ge = g.astype(np.float64).copy()
        if reg > 0:
            ge[:, 3:] += reg * pe[:, 3:]
        act = np.zeros_like(ge, bool)
        act[:, :active] = True
        bad = act & ~np.isfinite(ge)
        ok = act & ~bad
        m[bad] = 0.0
        v[bad] = 0.0
        m[ok] = 0.9 * m[ok] + 0.1 * ge[ok]
        v[ok] = 0.999 * v[ok] + 0.001 * ge[ok] ** 2
        upd = (m / c1) / (np.sqrt(v / c2) + eps)
        pe[ok] -= (lrs * upd)[ok]
        # the codec keeps sqrt(v) to about exp(range / 510) (8-bit codes over a log range up to
        # ~25 here: <= 5%), and u to fp16; the error per step stays a small fraction of lr
        # (measured: at most 2% of lr * step over this schedule)
        tol = 0.04 * np.where(np.arange(stride) >= 3, rest_lr, lr) * step
        assert np.all(np.abs(p - pe) <= tol[None, :] + 1e-7), (step, float(np.abs(p - pe).max()))
    # a plausible codec slip (lr misapplied, corrections dropped) would exceed the bound
    assert np.abs(p - p0).max() > 10 * 0.04 * rest_lr


def test_inactive_columns_keep_first_moment(oracle):
    rng = np.random.default_rng(3)
    rows, stride = 9, 48
    p = rng.normal(0, 0.1, (rows, stride)).astype(np.float32)
    st = oracle.ShQuantState(rows, stride)
    for step in range(1, 4):
        g = rng.normal(0, 1e-3, (rows, stride)).astype(np.float32)
        c1, c2 = corrections(step)
        oracle.sh_adam_quant_step(p, g, st, stride, 0.01, 0.001, 0.9, 0.999, c1, c2, 1e-15)

    def first_moment(step):
        c1, c2 = corrections(step)
        codes = st.packed.astype(np.float32)
        lo = np.where(np.arange(stride) < 3, st.bounds[:, 0:1], st.bounds[:, 2:3])
        hi = np.where(np.arange(stride) < 3, st.bounds[:, 1:2], st.bounds[:, 3:4])
        log_s = lo + (hi - lo) * codes / 255.0
        v = (1e-15 * np.expm1(log_s.astype(np.float64))) ** 2
        return st.first_f32() * (np.sqrt(v / c2) + 1e-15) * c1

    m3 = first_moment(3)
    p_before = p.copy()
    c1, c2 = corrections(4)
    oracle.sh_adam_quant_step(p, np.zeros_like(p), st, 3, 0.01, 0.001, 0.9, 0.999, c1, c2, 1e-15)
    assert np.array_equal(p[:, 3:], p_before[:, 3:])
    m4 = first_moment(4)
    # bounds of the non-DC group are re-taken from the same codes: the moment survives to fp16
    # precision of u and the codec's re-rounding of log_s (one code step at most)
    rel = np.abs(m4[:, 3:] - m3[:, 3:]) / np.maximum(np.abs(m3[:, 3:]), 1e-30)
    assert np.median(rel) < 2e-3 and rel.max() < 0.1


def test_group_extremes_take_end_codes(oracle):
    rng = np.random.default_rng(5)
    rows, stride = 16, 48
    p = rng.normal(0, 0.1, (rows, stride)).astype(np.float32)
    st = oracle.ShQuantState(rows, stride)
    g = (rng.normal(0, 1, (rows, stride)) * 10.0 ** rng.uniform(-6, -1, (rows, stride))).astype(np.float32)
    oracle.sh_adam_quant_step(p, g, st, stride, 0.01, 0.001, 0.9, 0.999, 0.1, 0.001, 1e-15)
    for r in range(rows):
        for grp in (slice(0, 3), slice(3, stride)):
            codes = st.packed[r, grp]
            assert codes.min() == 0 and codes.max() == 255
    # the log-seconds are log1p(sqrt(v) / 1e-15) of the first step's v = 0.001 g^2
    want = np.log1p(np.sqrt(0.001 * g.astype(np.float64) ** 2) / 1e-15)
    np.testing.assert_allclose(st.bounds[:, 0], want[:, :3].min(1), rtol=1e-6)
    np.testing.assert_allclose(st.bounds[:, 1], want[:, :3].max(1), rtol=1e-6)
    np.testing.assert_allclose(st.bounds[:, 2], want[:, 3:].min(1), rtol=1e-6)
    np.testing.assert_allclose(st.bounds[:, 3], want[:, 3:].max(1), rtol=1e-6)


def test_regularisation_is_added_gradient(oracle):
    rng = np.random.default_rng(7)
    rows, stride, r = 5, 48, np.float32(0.003)
    p0 = rng.normal(0, 0.1, (rows, stride)).astype(np.float32)
    g = rng.normal(0, 1e-3, (rows, stride)).astype(np.float32)
    a, b = p0.copy(), p0.copy()
    sa, sb = oracle.ShQuantState(rows, stride), oracle.ShQuantState(rows, stride)
    oracle.sh_adam_quant_step(a, g, sa, stride, 0.01, 0.001, 0.9, 0.999, 0.1, 0.001, 1e-15, float(r))
    g2 = g.copy()
    g2[:, 3:] = g[:, 3:] + r * p0[:, 3:]
    oracle.sh_adam_quant_step(b, g2, sb, stride, 0.01, 0.001, 0.9, 0.999, 0.1, 0.001, 1e-15, 0.0)
    assert np.array_equal(a, b)
    assert np.array_equal(sa.first, sb.first) and np.array_equal(sa.packed, sb.packed)
    assert np.array_equal(sa.bounds, sb.bounds)


def test_non_finite_gradient_resets_state(oracle):
    p = np.full((2, 12), 0.5, np.float32)
    st = oracle.ShQuantState(2, 12)
    g = np.full((2, 12), 1e-3, np.float32)
    oracle.sh_adam_quant_step(p, g, st, 12, 0.01, 0.001, 0.9, 0.999, 0.1, 0.001, 1e-15)
    before = p.copy()
    g[0, 0], g[1, 5] = np.nan, np.inf
    c1, c2 = corrections(2)
    oracle.sh_adam_quant_step(p, g, st, 12, 0.01, 0.001, 0.9, 0.999, c1, c2, 1e-15)
    assert p[0, 0] == before[0, 0] and p[1, 5] == before[1, 5]
    assert st.first[0, 0] == 0 and st.first[1, 5] == 0
    assert np.all(p[0, 1:] != before[0, 1:])


def test_zero_state_stays_zero(oracle):
    p = np.zeros((3, 3), np.float32)
    st = oracle.ShQuantState(3, 3)
    oracle.sh_adam_quant_step(p, np.zeros_like(p), st, 3, 0.01, 0.001, 0.9, 0.999, 0.1, 0.001, 1e-15)
    assert not p.any() and not st.first.any() and not st.packed.any() and not st.bounds.any()
