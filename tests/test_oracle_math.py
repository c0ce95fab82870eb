"""Pins of the oracle's scalar primitives (CPU): exp/log written from Cody-Waite reduction and
series, the fixed-order footprint power, and the bf16/fp16 storage conversions."""
from __future__ import annotations

import numpy as np
import torch


def ulp_err(got: np.ndarray, want64: np.ndarray) -> np.ndarray:
    """Error in units of the float32 ulp at the exact value."""
    want32 = want64.astype(np.float32)
    ulp = np.spacing(np.abs(want32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want64) / ulp


def test_pexpf_error_bound(oracle):
    rng = np.random.default_rng(0)
    x = np.concatenate([np.linspace(-87.0, 88.5, 400001, dtype=np.float32),
                        rng.uniform(-15.0, 0.0, 200000).astype(np.float32),
                        np.array([0.0, -0.0, 1e-30, -1e-30, 1.0, -1.0], np.float32)])
    y = oracle.pexpf(x)
    err = ulp_err(y, np.exp(x.astype(np.float64)))
    assert err.max() <= 2.0, err.max()
    assert oracle.pexpf(np.array([0.0], np.float32))[0] == 1.0
    special = oracle.pexpf(np.array([-np.inf, np.nan, 100.0, -200.0], np.float32))
    assert special[0] == 0.0 and np.isnan(special[1]) and special[2] == np.inf and special[3] == 0.0
    # subnormal results stay accurate (denormals are not flushed)
    xs = np.linspace(-103.0, -88.0, 20001, dtype=np.float32)
    ys = oracle.pexpf(xs)
    want = np.exp(xs.astype(np.float64))
    assert np.abs(ys.astype(np.float64) - want).max() <= 2 * 1.4e-45


def test_pexp2f_blend_error_bound(oracle):
    """The blend's 2^x on the exponents that reach it ([-8, 0]): every 64th float32 bit pattern
    of the interval (16M values) plus the ends and the integers, within 2.1 ulp of float64 2^x (the
    degree-5 fit's 2.03) and 2.5e-7 relative; 2^integer exact; NaN propagates."""
    lo_bits = np.float32(-8.0).view(np.uint32)  # negative floats: larger bits = more negative
    neg = np.arange(0x80000000, lo_bits + 1, 1 << 6, dtype=np.uint64).astype(np.uint32).view(np.float32)
    ints = -np.arange(0, 9, dtype=np.float32)
    x = np.concatenate([neg, ints, np.array([0.0, -0.0], np.float32)])
    y = oracle.pexp2f_blend(x).astype(np.float64)
    want = np.exp2(x.astype(np.float64))
    ulp = np.abs(y - want) / np.spacing(want.astype(np.float32)).astype(np.float64)
    assert ulp.max() <= 2.1, float(ulp.max())
    assert float((np.abs(y - want) / want).max()) <= 2.5e-7
    np.testing.assert_array_equal(oracle.pexp2f_blend(ints), np.exp2(ints))
    assert np.isnan(oracle.pexp2f_blend(np.array([np.nan], np.float32))[0])


def test_pair_power2_is_log2e_times_the_power(oracle):
    """The base-2 exponent of a pair equals log2(e) * (-q/2), q = a dx^2 + 2 b dx dy + c dy^2,
    evaluated exactly in float64 from the same float32 inputs, to float rounding."""
    rng = np.random.default_rng(3)
    n = 200000
    c3 = np.stack([rng.uniform(0.001, 4.0, n), rng.uniform(-0.5, 0.5, n), rng.uniform(0.001, 4.0, n)], 1)
    c3[:, 1] *= np.sqrt(c3[:, 0] * c3[:, 2])  # positive definite
    c3 = c3.astype(np.float32)
    d = rng.uniform(-12, 12, (n, 2)).astype(np.float32)
    got = oracle.pair_power2(c3, d).astype(np.float64)
    a, b, c = (c3[:, k].astype(np.float64) for k in range(3))
    dx, dy = d[:, 0].astype(np.float64), d[:, 1].astype(np.float64)
    q = a * dx * dx + 2 * b * dx * dy + c * dy * dy
    want = -0.5 * np.log2(np.e) * q
    scale = 0.5 * np.log2(np.e) * (a * dx * dx + 2 * np.abs(b * dx * dy) + c * dy * dy)
    assert np.all(np.abs(got - want) <= 4 * 2.0 ** -24 * scale + 1e-30)


def test_plogf_error_bound(oracle):
    rng = np.random.default_rng(1)
    x = np.concatenate([np.exp(rng.uniform(-87, 88, 400000)).astype(np.float32),
                        np.linspace(1.0, 255.0, 200001, dtype=np.float32),
                        np.array([1.0, 2.0, 0.5, 255.0, 1e-40, 1.17549435e-38], np.float32)])
    y = oracle.plogf(x)
    want = np.log(x.astype(np.float64))
    err = np.abs(y.astype(np.float64) - want) / np.spacing(np.maximum(np.abs(want), 1e-30).astype(np.float32))
    # near x = 1 the result is tiny; bound it absolutely there
    assert err[np.abs(want) > 1e-3].max() <= 2.0
    assert np.abs(y.astype(np.float64) - want)[np.abs(want) <= 1e-3].max() <= 1e-10
    assert oracle.plogf(np.array([1.0], np.float32))[0] == 0.0
    sp = oracle.plogf(np.array([0.0, -1.0, np.inf], np.float32))
    assert sp[0] == -np.inf and np.isnan(sp[1]) and sp[2] == np.inf


def test_reduced_precision_conversions_match_torch(oracle):
    from photara_studio_amd import _abi
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0, 3, 100000), rng.normal(0, 1e-6, 10000), rng.normal(0, 3e4, 10000),
                        [0.0, -0.0, 65504.0, 65520.0, 1e9, -1e9, 6e-8, 1e-7, np.inf, -np.inf]]).astype(np.float32)
    for fmt, dt in ((_abi.PHT_LP_BF16, torch.bfloat16), (_abi.PHT_LP_FP16, torch.float16)):
        bits, back = oracle.lp_roundtrip(x, fmt)
        ref = torch.from_numpy(x).to(dt)
        np.testing.assert_array_equal(bits, ref.view(torch.int16).numpy().view(np.uint16))
        np.testing.assert_array_equal(back, ref.float().numpy())
