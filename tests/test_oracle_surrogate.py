"""Pins the float64 surrogate oracle: analytic gradient against central
differences of the loss, sensitivity of that check to planted mistakes, the
bf16 rounding helper against its definition, and the Adam recurrence against
a closed form."""
import math

import numpy as np
import pytest

from oracle import surrogate as S


def _setup(seed=0, n=48, max_range=100.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(S.NPARAM)
    for (w, b, o, i) in S.LAYERS:
        p[w:w + o * i] = rng.standard_normal(o * i) * math.sqrt(2.0 / i)
        p[b:b + o] = rng.standard_normal(o) * 0.1
    p[S.B4] = 0.3
    p[S.B4 + 1] = 0.5
    freq = rng.standard_normal((S.NF, 4)) * np.array([12.0, 12.0, 2.0, 2.0])
    rays = np.stack([rng.uniform(0, 64, n), rng.uniform(0, 64, n), *(lambda h: (np.sin(h), np.cos(h)))(
        rng.uniform(0, 2 * math.pi, n))], axis=1)
    x = S.features(rays, freq, 64, 64)
    _, o = S.forward(p, x)
    # targets around the prediction: both Huber regimes and both pinball sides
    target = max_range * o[:, 0] + rng.choice([-5.0, -0.4, 0.3, 3.0, 12.0], n)
    return p, x, target, max_range


def _directional_mismatch(g, p, x, target, R, seed=1, eps=1e-6, n_dirs=6):
    _, _, o = S.loss_and_grad(p, x, target, R)
    ae0 = np.abs(R * o[:, 0] - target)
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(n_dirs):
        v = rng.standard_normal(S.NPARAM)
        v /= np.linalg.norm(v)
        fd = (S.loss(p + eps * v, x, target, R, ae_fixed=ae0) - S.loss(p - eps * v, x, target, R, ae_fixed=ae0)) / (2 * eps)
        worst = max(worst, abs(fd - g @ v) / max(abs(g @ v), 1e-12))
    return worst


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gradient_matches_finite_differences(seed):
    p, x, t, R = _setup(seed)
    assert S.finite_difference_check(p, x, t, R, seed=seed) < 1e-5


def test_finite_difference_check_catches_planted_mistakes():
    p, x, t, R = _setup(3)
    _, g, _ = S.loss_and_grad(p, x, t, R)
    assert _directional_mismatch(g, p, x, t, R) < 1e-5
    planted = {}
    g1 = g.copy(); g1[S.B4 + 1] *= 10.0; planted["pinball weight dropped (b4)"] = g1
    g2 = g.copy(); w = g2[S.W2:S.W2 + S.H * S.H].reshape(S.H, S.H); g2[S.W2:S.W2 + S.H * S.H] = w.T.reshape(-1)
    planted["W2 gradient transposed"] = g2
    g3 = g.copy(); g3[S.B1:S.B1 + S.H] *= -1.0; planted["b1 sign"] = g3
    g4 = g.copy(); g4[S.W4:S.W4 + S.H] *= 2.0; planted["range head scaled"] = g4
    for name, gp in planted.items():
        assert _directional_mismatch(gp, p, x, t, R, n_dirs=12) > 1e-3, name


def test_clamped_bound_head_has_no_gradient():
    p, x, t, R = _setup(4, n=8)
    p = p.copy()
    p[S.B4 + 1] = 40.0  # exp bound saturates at the clamp for every ray
    _, g, o = S.loss_and_grad(p, x, t, R)
    assert np.all(o[:, 1] > S.CLAMP)
    assert g[S.B4 + 1] == 0.0 and not np.any(g[S.W4 + S.H:S.W4 + 2 * S.H])


def test_bf16_rounding_definition():
    # exact bf16 values are fixed points; halfway cases go to even; NaN stays NaN
    vals = np.concatenate([np.array([1.0, -2.5, 0.0, -0.0], np.float32),
                           np.array([0x7F7F0000, 0x00010000], np.uint32).view(np.float32)])  # max, subnormal
    assert np.array_equal(S.bf16_to_f64(S.bf16_bits(vals)), vals.astype(np.float64))
    one = np.uint32(0x3F800000)
    half_up_even = np.array([one + 0x8000, one + 0x18000, one + 0x8001], np.uint32).view(np.float32)
    assert list(S.bf16_bits(half_up_even)) == [0x3F80, 0x3F82, 0x3F81]
    assert S.bf16_bits(np.array([np.nan], np.float32))[0] == 0x7FC0
    rng = np.random.default_rng(0)
    x = rng.standard_normal(10000).astype(np.float32)
    r = S.bf16_to_f64(S.bf16_bits(x))
    half_ulp = np.exp2(np.floor(np.log2(np.abs(x.astype(np.float64)))) - 8)  # 8 significant bits
    assert np.all(np.abs(r - x) <= half_ulp)


def test_emulated_bf16_forward_is_close_to_float64():
    p, x, t, R = _setup(5)
    _, o64 = S.forward(p, x)
    _, o16 = S.forward(p, x, emulate_bf16=True)
    rel = np.linalg.norm(o16 - o64) / np.linalg.norm(o64)
    assert 1e-5 < rel < 3e-2


def test_adam_first_steps_closed_form():
    rng = np.random.default_rng(6)
    p0 = rng.standard_normal(100)
    g = rng.standard_normal(100)
    p1, m1, m2 = S.adam_step(p0, np.zeros(100), np.zeros(100), g, 1, 0.01)
    # first step of Adam moves every coordinate by lr * sign(g) (up to eps)
    assert np.allclose(p1, p0 - 0.01 * np.sign(g), rtol=0, atol=1e-7)
    p2, _, _ = S.adam_step(p1, m1, m2, g, 2, 0.01)
    assert np.allclose(p2, p1 - 0.01 * np.sign(g), rtol=0, atol=1e-7)  # constant gradient: same again
