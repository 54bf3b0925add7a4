"""Pins of the learned correction's CPU oracle (oracle/mlp_oracle.py).

BF16 rounding against torch's round-to-nearest-even cast and known bit
patterns; the backward pass against central finite differences of the loss
(float64, no rounding), so a dropped term, a wrong transpose or a missing
SiLU derivative fails; the rounded gradients against the exact ones."""
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, str(ROOT / "oracle"))
import mlp_oracle as mo  # noqa: E402


def test_bf16_known_bit_patterns():
    bits = lambda x: int(mo.bf16(np.float32(x)).view(np.uint32).item()) >> 16  # noqa: E731
    assert bits(1.0) == 0x3F80
    assert bits(1.0 + 2.0 ** -8) == 0x3F80          # tie -> even
    assert bits(1.0 + 3 * 2.0 ** -8) == 0x3F82      # tie -> even (up)
    assert bits(1.0 + 2.0 ** -8 + 2.0 ** -20) == 0x3F81
    assert bits(-2.5) == 0xC020
    assert bits(np.inf) == 0x7F80
    assert bits(np.nan) == 0x7FC0
    assert bits(3.3895e38) == 0x7F7F                 # largest finite BF16
    assert bits(3.4e38) == 0x7F80                    # beyond it: infinity


def test_bf16_matches_torch_cast():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-30, 1000),
                        rng.normal(0, 1e30, 1000)]).astype(np.float32)
    ours = mo.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))


def _problem(seed, n=40):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed)
    p[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.1, mo.H)
    p[mo.OFF_B2:mo.OFF_W3] = rng.normal(0, 0.1, mo.H)
    p[mo.OFF_B3:mo.OFF_B3 + 3] = rng.normal(0, 0.1, 3)
    feats = rng.normal(0, 1, (n, mo.F)).astype(np.float32)
    feats[:, 10:] = 0.0
    mean = rng.normal(0, 0.1, mo.F).astype(np.float32)
    std = rng.uniform(0.5, 2.0, mo.F).astype(np.float32)
    targets = rng.normal(0, 1, (n, 3)).astype(np.float32)
    return p, mean, std, feats, targets


def test_backward_matches_finite_differences():
    p, mean, std, feats, targets = _problem(3)
    g, L = mo.gradients(p, mean, std, feats, targets, rounding=False)
    assert L == pytest.approx(mo.loss(p, mean, std, feats, targets, rounding=False), rel=1e-12)
    rng = np.random.default_rng(7)
    used = [k for k in range(mo.N_PARAMS)
            if not (mo.OFF_W3 + 3 * mo.H <= k < mo.OFF_B3 or k >= mo.OFF_B3 + 3)]
    # every block of parameters is probed
    picks = []
    for lo, hi in ((mo.OFF_W1, mo.OFF_B1), (mo.OFF_B1, mo.OFF_W2), (mo.OFF_W2, mo.OFF_B2),
                   (mo.OFF_B2, mo.OFF_W3), (mo.OFF_W3, mo.OFF_W3 + 3 * mo.H),
                   (mo.OFF_B3, mo.OFF_B3 + 3)):
        cand = [k for k in used if lo <= k < hi]
        picks += list(rng.choice(cand, size=min(12, len(cand)), replace=False))
    pd = p.astype(np.float64)
    for k in picks:
        h = 1e-6
        a, b = pd.copy(), pd.copy()
        a[k] += h
        b[k] -= h
        fd = (_loss64(a, mean, std, feats, targets) - _loss64(b, mean, std, feats, targets)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def _loss64(p64, mean, std, feats, targets):
    """Float64 loss with no rounding, parameters kept in float64."""
    P = {"W1": p64[mo.OFF_W1:mo.OFF_B1].reshape(mo.H, mo.F), "b1": p64[mo.OFF_B1:mo.OFF_W2],
         "W2": p64[mo.OFF_W2:mo.OFF_B2].reshape(mo.H, mo.H), "b2": p64[mo.OFF_B2:mo.OFF_W3],
         "W3": p64[mo.OFF_W3:mo.OFF_B3].reshape(16, mo.H), "b3": p64[mo.OFF_B3:]}
    x = ((feats.astype(np.float64) - mean) * (np.float32(1) / std.astype(np.float32))).T
    silu = lambda z: z / (1 + np.exp(-z))  # noqa: E731
    h1 = silu(P["W1"] @ x + P["b1"][:, None])
    h2 = silu(P["W2"] @ h1 + P["b2"][:, None])
    y = (P["W3"] @ h2 + P["b3"][:, None])[:3].T
    return float(((y - targets) ** 2).sum(1).mean())


def test_rounded_gradients_track_exact_gradients():
    p, mean, std, feats, targets = _problem(5, n=256)
    g_r, L_r = mo.gradients(p, mean, std, feats, targets, rounding=True)
    g_x, L_x = mo.gradients(p, mean, std, feats, targets, rounding=False)
    assert L_r == pytest.approx(L_x, rel=2e-2)
    rel = np.linalg.norm(g_r - g_x) / np.linalg.norm(g_x)
    assert rel < 3e-2
    # padding rows of W3 / b3 get no gradient
    assert not g_r[mo.OFF_W3 + 3 * mo.H:mo.OFF_B3].any() and not g_r[mo.OFF_B3 + 3:].any()


def test_forward_zero_weights_gives_bias():
    p = np.zeros(mo.N_PARAMS, np.float32)
    p[mo.OFF_B3:mo.OFF_B3 + 3] = [0.5, -1.0, 2.0]
    feats = np.random.default_rng(0).normal(0, 1, (10, mo.F)).astype(np.float32)
    y, _ = mo.forward(p, np.zeros(mo.F), np.ones(mo.F), feats)
    assert np.array_equal(y, np.tile(np.float32([0.5, -1.0, 2.0]), (10, 1)))
