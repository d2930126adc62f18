"""CPU float64 restatement of the learned range surrogate (test infrastructure).

The surrogate has no counterpart in the reference; this module is the checker
for the library's hand-written forward, backward and optimiser kernels.  Only
tests, the smoke check and the benchmark's CPU leg use it.

Network (parameters flat, fp32, row-major [out][in]):
    v   = (x0 / W, y0 / H, sin h, cos h)
    phi = [sin(2 pi B v), cos(2 pi B v)]          B: 32 x 4  -> 64 features
    a1  = relu(W1 phi + b1), a2 = relu(W2 a1 + b2), a3 = relu(W3 a2 + b3)
    o   = W4 a3 + b4
Loss (mean over rays), with e = max_range * o0 - target and b = exp(clamp(o1, -20, 20)):
    huber(e; delta 1) + 0.1 * pinball_0.95(|e| - b)
where |e| is held constant in the pinball term (its gradient flows to o1 only),
and the pinball gradient is zero where the clamp is active.

``emulate_bf16=True`` rounds where the GPU rounds (weights, ReLU outputs, the
output gradient and each back-propagated activation gradient to bf16, round to
nearest even from the fp32 value); sums stay in float64.  Without it every
step is float64.

Pinning: ``finite_difference_check`` compares the analytic gradient with
central differences of ``loss`` (float64); ``tests/test_oracle.py`` runs it.
"""
from __future__ import annotations

import math

import numpy as np

F, NF, H, OUT = 64, 32, 256, 2
TAU, DELTA, PIN_W, CLAMP = 0.95, 1.0, 0.1, 20.0

# flat parameter layout
W1 = 0
B1 = W1 + H * F
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H * H
W4 = B3 + H
B4 = W4 + OUT * H
NPARAM = B4 + OUT

LAYERS = ((W1, B1, H, F), (W2, B2, H, H), (W3, B3, H, H), (W4, B4, OUT, H))


def bf16_bits(x) -> np.ndarray:
    """Round to bf16 (nearest even) from float32; returns the uint16 bit patterns."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = 0x7FC0
    return r


def bf16_to_f64(bits) -> np.ndarray:
    b = np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16
    return b.view(np.float32).astype(np.float64)


def bf16_round(x) -> np.ndarray:
    """float64 -> float32 -> bf16 -> float64 (the GPU rounds fp32 results)."""
    return bf16_to_f64(bf16_bits(np.asarray(x, dtype=np.float64).astype(np.float32)))


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    return [(p[w:w + o * i].reshape(o, i), p[b:b + o]) for (w, b, o, i) in LAYERS]


def features(rays4, freq, width: int, height: int) -> np.ndarray:
    """Fourier features (n x 64) in float64."""
    r = np.asarray(rays4, dtype=np.float64).reshape(-1, 4)
    v = np.stack([r[:, 0] / width, r[:, 1] / height, r[:, 2], r[:, 3]], axis=1)
    ph = 2.0 * math.pi * (v @ np.asarray(freq, dtype=np.float64).reshape(NF, 4).T)
    return np.concatenate([np.sin(ph), np.cos(ph)], axis=1)


def forward(params, x, emulate_bf16: bool = False):
    """Activations [x, a1, a2, a3] and outputs o (n x 2)."""
    layers = unpack(params)
    rnd = bf16_round if emulate_bf16 else (lambda a: a)
    acts = [np.asarray(x, dtype=np.float64)]
    h = acts[0]
    for k, (Wm, b) in enumerate(layers):
        z = h @ rnd(Wm).T + b
        if k < 3:
            h = rnd(np.maximum(z, 0.0))
            acts.append(h)
        else:
            return acts, z
    raise AssertionError


def _loss_terms(o, target, max_range, ae_fixed=None):
    e = max_range * o[:, 0] - np.asarray(target, dtype=np.float64)
    ae = np.abs(e)
    huber = np.where(ae <= DELTA, 0.5 * e * e, DELTA * (ae - 0.5 * DELTA))
    lb = np.clip(o[:, 1], -CLAMP, CLAMP)
    b = np.exp(lb)
    aq = ae if ae_fixed is None else ae_fixed
    pin = np.where(aq > b, TAU * (aq - b), (1.0 - TAU) * (b - aq))
    return e, ae, b, lb, huber + PIN_W * pin


def loss(params, x, target, max_range, ae_fixed=None, emulate_bf16=False) -> float:
    _, o = forward(params, x, emulate_bf16)
    return float(np.mean(_loss_terms(o, target, max_range, ae_fixed)[4]))


def loss_and_grad(params, x, target, max_range, emulate_bf16: bool = False):
    """(loss, flat gradient float64, outputs)."""
    acts, o = forward(params, x, emulate_bf16)
    n = o.shape[0]
    e, ae, b, lb, terms = _loss_terms(o, target, max_range)
    ge = np.where(ae <= DELTA, e, np.sign(e) * DELTA)
    gb = np.where(ae > b, -TAU, 1.0 - TAU) * b
    gb = np.where(lb == o[:, 1], gb, 0.0)
    dz = np.stack([ge * max_range / n, PIN_W * gb / n], axis=1)
    rnd = bf16_round if emulate_bf16 else (lambda a: a)
    if emulate_bf16:
        dz = bf16_round(dz)
    layers = unpack(params)
    grad = np.zeros(NPARAM)
    for k in range(3, -1, -1):
        w_off, b_off, fo, fi = LAYERS[k]
        a_in = acts[k]
        grad[w_off:w_off + fo * fi] = (dz.T @ a_in).reshape(-1)
        grad[b_off:b_off + fo] = dz.sum(axis=0)
        if k == 0:
            break
        dprev = dz @ rnd(layers[k][0])
        dprev = np.where(a_in > 0.0, dprev, 0.0)
        dz = rnd(dprev)
    return float(np.mean(terms)), grad, o


def finite_difference_check(params, x, target, max_range, n_dirs: int = 6, eps: float = 1e-6,
                            seed: int = 0) -> float:
    """Largest relative disagreement between the analytic directional derivative
    and a central difference of ``loss`` along random directions (float64,
    |e| held at its value at ``params`` as the loss defines)."""
    p0 = np.asarray(params, dtype=np.float64)
    _, g, o = loss_and_grad(p0, x, target, max_range)
    ae0 = np.abs(max_range * o[:, 0] - np.asarray(target, dtype=np.float64))
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(n_dirs):
        v = rng.standard_normal(NPARAM)
        v /= np.linalg.norm(v)
        lp = loss(p0 + eps * v, x, target, max_range, ae_fixed=ae0)
        lm = loss(p0 - eps * v, x, target, max_range, ae_fixed=ae0)
        fd = (lp - lm) / (2.0 * eps)
        an = float(g @ v)
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-12))
    return worst


def adam_step(p, m1, m2, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (float64), t >= 1."""
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    p = p - lr * (m1 / (1 - b1 ** t)) / (np.sqrt(m2 / (1 - b2 ** t)) + eps)
    return p, m1, m2
