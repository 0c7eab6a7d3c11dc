"""The CPU restatement of the delta network (oracle/mlp_ref.py), pinned by
finite differences and by an independent bf16 rounding rule."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import mlp_ref as R


def _bf16_nearest_even(x: float) -> float:
    """Round a finite fp32 value to bf16 (8 significant bits) by exact
    rational arithmetic: nearest, ties to the even neighbour."""
    if x == 0.0 or not np.isfinite(x):
        return x
    f = Fraction(float(np.float32(x)))
    sign = -1 if f < 0 else 1
    f = abs(f)
    e = 0
    while f >= 2:
        f /= 2
        e += 1
    while f < 1:
        f *= 2
        e -= 1
    scaled = f * 128  # 7 explicit mantissa bits
    lo = scaled.numerator // scaled.denominator
    rem = scaled - lo
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and lo % 2 == 1):
        lo += 1
    return float(sign * Fraction(lo, 128) * Fraction(2) ** e)


def test_bf16_rounding_matches_exact_rule():
    rng = np.random.default_rng(5)
    special = [1.0, 1.00390625, 1.01171875, -1.00390625, 3.0e-3, -7.5, 65504.0, 1e-20,
               np.float32(1.0) + np.float32(2.0 ** -8) * 3, 0.0]
    xs = np.concatenate([np.array(special, np.float32),
                         rng.normal(0, 1, 2000).astype(np.float32),
                         (rng.normal(0, 1, 500) * 1e6).astype(np.float32)])
    got = R.bf16_round(xs)
    want = np.array([_bf16_nearest_even(float(x)) for x in xs], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # ties go to the even neighbour
    assert R.bf16_round(np.float32(1.00390625)) == np.float32(1.0)
    assert R.bf16_round(np.float32(1.01171875)) == np.float32(1.015625)


def _problem(seed, n=37):
    rng = np.random.default_rng(seed)
    flat = rng.normal(0, 0.3, R.N_PARAMS)
    mean = rng.normal(0, 1, 32)
    inv = rng.uniform(0.5, 2.0, 32)
    x = rng.normal(0, 1, (n, 32)) / inv + mean
    y = rng.normal(0, 1, (n, R.N_OUT))
    w = rng.uniform(0.5, 2.0, n)
    return flat, mean, inv, x, y, w


@pytest.mark.parametrize("seed", [1, 2])
def test_gradient_matches_finite_differences(seed):
    flat, mean, inv, x, y, w = _problem(seed)
    L, g = R.loss_grad(flat, mean, inv, x, y, w, round_weights=False)
    assert L == pytest.approx(R.loss(flat, mean, inv, x, y, w, round_weights=False))
    rng = np.random.default_rng(seed + 10)
    idx = np.concatenate([rng.choice(2048, 40, replace=False), 2048 + rng.choice(64, 10, replace=False),
                          2112 + rng.choice(64 * R.N_OUT, 12, replace=False),
                          2112 + 64 * R.N_OUT + np.arange(R.N_OUT)])
    eps = 1e-6
    for i in idx:
        p, m = flat.copy(), flat.copy()
        p[i] += eps
        m[i] -= eps
        fd = (R.loss(p, mean, inv, x, y, w, False) - R.loss(m, mean, inv, x, y, w, False)) / (2 * eps)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-8), i


def test_rounded_gradient_is_the_gradient_at_the_rounded_weights():
    flat, mean, inv, x, y, w = _problem(3)
    rounded = flat.copy()
    rounded[:2048] = R.bf16_round(flat[:2048].astype(np.float32)).astype(np.float64)
    w2 = slice(2112, 2112 + 64 * R.N_OUT)
    rounded[w2] = R.bf16_round(flat[w2].astype(np.float32)).astype(np.float64)
    La, ga = R.loss_grad(flat, mean, inv, x, y, w, round_weights=True)
    Lb, gb = R.loss_grad(rounded, mean, inv, x, y, w, round_weights=False)
    assert La == Lb
    assert np.array_equal(ga, gb)


def test_sample_weights_scale_the_loss():
    flat, mean, inv, x, y, _ = _problem(4)
    L1, g1 = R.loss_grad(flat, mean, inv, x, y, None)
    L2, g2 = R.loss_grad(flat, mean, inv, x, y, np.full(len(x), 2.0))
    assert L2 == pytest.approx(2 * L1)
    assert np.allclose(g2, 2 * g1)
