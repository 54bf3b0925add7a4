"""CPU oracle for the learned Barnes-Hut correction's network.

TEST INFRASTRUCTURE ONLY: the checker the HIP MLP kernels are compared
against. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import it; nothing in the product imports or calls it.

The reference has no learned component (its walk only counts accepted cells,
src/c/barnes_hut/barnes_hut.c:611), so this restates the network that
include/nbx_learned.h specifies, with the same rounding points as the HIP
kernels:

    xb  = bf16((f - mean) * inv_std),  inv_std = float32(1 / std)
    z1  = W1b xb + b1            h1b = bf16(silu(z1))
    z2  = W2b h1b + b2           h2b = bf16(silu(z2))
    y   = W3b h2b + b3           (outputs 0..2 used)
    L   = mean_p sum_o (y - t)^2
    dy  = 2 (y - t) / B          dyb = bf16(dy)
    gW3 = dyb h2b^T              gb3 = sum_p dy
    dz2 = (W3b^T dyb) silu'(z2)  dz2b = bf16(dz2)
    gW2 = dz2b h1b^T             gb2 = sum_p dz2
    dz1 = (W2b^T dz2b) silu'(z1) dz1b = bf16(dz1)
    gW1 = dz1b xb^T              gb1 = sum_p dz1

BF16 products are exact in FP32, so the matrix products here are accumulated
in float64 and rounded to float32 once; the device accumulates in FP32 in a
different order, so the two agree to FP32 rounding except where a BF16
rounding point falls on the other side of a tie.

Parity pins: `bf16` against torch's round-to-nearest-even BF16 cast and known
bit patterns; `gradients(rounding=False)` against central finite differences
of `loss(rounding=False)` in float64; the device MFMA lane layout against a
float64 product of BF16-rounded matrices.
"""
import numpy as np

F, H, O_PAD, O = 16, 64, 16, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + O_PAD * H
N_PARAMS = OFF_B3 + O_PAD


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, np.uint64(0x7FC00000), r).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def unpack(params):
    p = np.asarray(params, dtype=np.float32)
    assert p.shape == (N_PARAMS,)
    return {
        "W1": p[OFF_W1:OFF_B1].reshape(H, F), "b1": p[OFF_B1:OFF_W2],
        "W2": p[OFF_W2:OFF_B2].reshape(H, H), "b2": p[OFF_B2:OFF_W3],
        "W3": p[OFF_W3:OFF_B3].reshape(O_PAD, H), "b3": p[OFF_B3:],
    }


def pack(W1, b1, W2, b2, W3, b3):
    out = np.zeros(N_PARAMS, dtype=np.float32)
    out[OFF_W1:OFF_B1] = np.asarray(W1).ravel()
    out[OFF_B1:OFF_W2] = b1
    out[OFF_W2:OFF_B2] = np.asarray(W2).ravel()
    out[OFF_B2:OFF_W3] = b2
    out[OFF_W3:OFF_B3] = np.asarray(W3).ravel()
    out[OFF_B3:] = b3
    return out


def init_params(seed):
    """He-style initial weights; unused output rows stay zero."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2.0 / F), (H, F))
    W2 = rng.normal(0, np.sqrt(2.0 / H), (H, H))
    W3 = np.zeros((O_PAD, H))
    W3[:O] = rng.normal(0, np.sqrt(1.0 / H), (O, H))
    return pack(W1, np.zeros(H), W2, np.zeros(H), W3, np.zeros(O_PAD))


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s + z * s * (1.0 - s)


def _mm(a, b, rounding):
    r = np.asarray(a, np.float64) @ np.asarray(b, np.float64)
    return r.astype(np.float32) if rounding else r


def _round(x, rounding):
    return bf16(x) if rounding else np.asarray(x, np.float64)


def forward(params, mean, std, feats, rounding=True):
    """Returns y (n x 3) and the intermediates; columns are samples."""
    P = unpack(params)
    dt = np.float32 if rounding else np.float64
    f = np.asarray(feats, dt)
    with np.errstate(divide="ignore"):
        inv_std = (np.float32(1.0) / np.asarray(std, np.float32)).astype(dt)   # +inf -> 0
    xs = ((f - np.asarray(mean, dt)) * inv_std).T                             # F x n
    xb = _round(xs, rounding)
    W1, W2, W3 = (_round(P[k], rounding) for k in ("W1", "W2", "W3"))
    z1 = _mm(W1, xb, rounding) + P["b1"][:, None].astype(dt)
    h1 = _round(_silu(z1.astype(np.float64)).astype(dt), rounding)
    z2 = _mm(W2, h1, rounding) + P["b2"][:, None].astype(dt)
    h2 = _round(_silu(z2.astype(np.float64)).astype(dt), rounding)
    y = _mm(W3, h2, rounding) + P["b3"][:, None].astype(dt)
    cache = {"xb": xb, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "W2": W2, "W3": W3}
    return y[:O].T, cache


def loss(params, mean, std, feats, targets, rounding=True):
    y, _ = forward(params, mean, std, feats, rounding)
    d = np.asarray(y, np.float64) - np.asarray(targets, np.float64)
    return float((d * d).sum(axis=1).mean())


def gradients(params, mean, std, feats, targets, rounding=True):
    """dL/dparams (N_PARAMS) and L, with the kernels' rounding points."""
    n = len(feats)
    y, c = forward(params, mean, std, feats, rounding)
    dt = np.float32 if rounding else np.float64
    diff = (y.astype(dt) - np.asarray(targets, dt)).T                # 3 x n
    L = float((diff.astype(np.float64) ** 2).sum() / n)
    dy = np.zeros((O_PAD, n), dt)
    dy[:O] = 2.0 * diff * dt(1.0 / n)
    dyb = _round(dy, rounding)
    gW3 = _mm(dyb, c["h2"].T, rounding)
    gb3 = dy.astype(np.float64).sum(axis=1)
    dz2 = (_mm(c["W3"].T, dyb, rounding) * _dsilu(c["z2"].astype(np.float64))).astype(dt)
    dz2b = _round(dz2, rounding)
    gW2 = _mm(dz2b, c["h1"].T, rounding)
    gb2 = dz2.astype(np.float64).sum(axis=1)
    dz1 = (_mm(c["W2"].T, dz2b, rounding) * _dsilu(c["z1"].astype(np.float64))).astype(dt)
    dz1b = _round(dz1, rounding)
    gW1 = _mm(dz1b, c["xb"].T, rounding)
    gb1 = dz1.astype(np.float64).sum(axis=1)
    g = np.zeros(N_PARAMS, np.float64)
    g[OFF_W1:OFF_B1] = gW1.ravel()
    g[OFF_B1:OFF_W2] = gb1
    g[OFF_W2:OFF_B2] = gW2.ravel()
    g[OFF_B2:OFF_W3] = gb2
    g[OFF_W3:OFF_B3] = gW3.ravel()
    g[OFF_B3:] = gb3
    return g, L
