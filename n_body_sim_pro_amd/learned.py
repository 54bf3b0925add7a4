"""Learned Barnes-Hut correction: host side of include/nbx_learned.h.

A theta = 1.0 FP32 walk, its quadrupole force q, and a small MLP (BF16 MFMA
kernels in csrc/nbx_mlp.hip) that predicts what q leaves of the walk's
residual against the FP64 direct sum, in each particle's local frame. Without
a model attached, the hip_bh_learned force function is plain hip_barnes_hut.

Training data are (BH walk, FP64 direct sum) pairs of N = 65,536 Plummer and
exponential-disk snapshots; both sides are computed on the GPU by libnbx.so.

    python -m n_body_sim_pro_amd.learned --out n_body_sim_pro_amd/weights/bh_learned_theta1.npz
"""
import argparse
import ctypes as C
import json
import math
import pathlib
import sys
import time

import numpy as np

from .abi import Error
from .engine import (Engine, NBXError, _check, _learned_lib, disk_galaxy, plummer)

N_FEATURES = 16
N_HIDDEN = 64
# Inputs the model uses (include/nbx_learned.h numbering); the others get
# std = +inf and standardise to 0. (Measured on MI355X against the FP64 direct
# sum, median error of the corrected walk over the quadrupole-corrected walk
# without the network, 10M / 2M two-galaxy: inputs 0, 1, 7, 8, 9: 0.975 /
# 0.939; with the density 2 as well: 0.992 / 0.940; 7, 8, 9 alone: 0.989 /
# 0.936.)
FEATURES = (0, 1, 7, 8, 9)
WEIGHTS = pathlib.Path(__file__).resolve().parent / "weights" / "bh_learned_theta1.npz"
_F = C.POINTER(C.c_float)


def _fp(a):
    return a.ctypes.data_as(_F)


class Mlp:
    """A model on one device (nbx_mlp_*): parameters, Adam state, training set."""

    def __init__(self, device=0):
        self.lib = _learned_lib()
        err = Error()
        self.h = self.lib.nbx_mlp_create(int(device), C.byref(err))
        if not self.h:
            raise NBXError(err.status, err.message.decode(errors="replace"))
        self.n_params = int(self.lib.nbx_mlp_param_count())
        self.device = int(device)

    def close(self):
        if getattr(self, "h", None):
            self.lib.nbx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_params(self, params, mean, std, out_scale):
        p = np.ascontiguousarray(params, dtype=np.float32)
        mu = np.ascontiguousarray(mean, dtype=np.float32)
        sd = np.ascontiguousarray(std, dtype=np.float32)
        if p.shape != (self.n_params,) or mu.shape != (N_FEATURES,) or sd.shape != (N_FEATURES,):
            raise ValueError("parameter, mean or std shape does not match the model")
        err = Error()
        _check(self.lib.nbx_mlp_set_params(self.h, _fp(p), _fp(mu), _fp(sd), float(out_scale),
                                           C.byref(err)), err)

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        mu = np.zeros(N_FEATURES, np.float32)
        sd = np.zeros(N_FEATURES, np.float32)
        s = C.c_float(0.0)
        err = Error()
        _check(self.lib.nbx_mlp_get_params(self.h, _fp(p), _fp(mu), _fp(sd), C.byref(s),
                                           C.byref(err)), err)
        return {"params": p, "mean": mu, "std": sd, "out_scale": s.value}

    def forward(self, feats):
        x = np.ascontiguousarray(feats, dtype=np.float32)
        out = np.zeros((len(x), 3), np.float32)
        err = Error()
        _check(self.lib.nbx_mlp_forward(self.h, _fp(x), len(x), _fp(out), C.byref(err)), err)
        return out

    def gradients(self, feats, targets):
        x = np.ascontiguousarray(feats, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float(0.0)
        err = Error()
        _check(self.lib.nbx_mlp_gradients(self.h, _fp(x), _fp(t), len(x), _fp(g), C.byref(loss),
                                          C.byref(err)), err)
        return g, loss.value

    def load_dataset(self, feats, targets):
        x = np.ascontiguousarray(feats, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        err = Error()
        _check(self.lib.nbx_mlp_load_dataset(self.h, _fp(x), _fp(t), len(x), C.byref(err)), err)

    def train(self, epochs, batch, lr, seed):
        hist = np.zeros(max(int(epochs), 1), np.float32)
        err = Error()
        _check(self.lib.nbx_mlp_train(self.h, int(epochs), int(batch), float(lr), int(seed),
                                      _fp(hist), C.byref(err)), err)
        return hist[:epochs]

    def save(self, path, **meta):
        d = self.get_params()
        path = pathlib.Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, params=d["params"], mean=d["mean"], std=d["std"],
                 out_scale=np.float32(d["out_scale"]), meta=np.array(json.dumps(meta)))

    @classmethod
    def load(cls, path=WEIGHTS, device=0):
        with np.load(path, allow_pickle=False) as z:
            m = cls(device)
            m.set_params(z["params"], z["mean"], z["std"], float(z["out_scale"]))
            m.meta = json.loads(str(z["meta"])) if "meta" in z.files else {}
        return m


def probe_mfma(a, b, device=0):
    """One v_mfma_f32_16x16x16_bf16 on the device: a @ b for 16x16 inputs
    rounded to BF16 (checks the lane layout the kernels assume)."""
    lib = _learned_lib()
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    if a.shape != (16, 16) or b.shape != (16, 16):
        raise ValueError("probe_mfma takes two 16x16 matrices")
    d = np.zeros((16, 16), np.float32)
    err = Error()
    _check(lib.nbx_mlp_probe_mfma(int(device), _fp(a), _fp(b), _fp(d), C.byref(err)), err)
    return d


def init_params(seed, features=FEATURES):
    """He-style initial weights in the nbx_learned.h layout; the columns of
    unused features and the padding rows are zero."""
    rng = np.random.default_rng(seed)
    F, H = N_FEATURES, N_HIDDEN
    W1 = np.zeros((H, F))
    W1[:, list(features)] = rng.normal(0.0, math.sqrt(2.0 / len(features)), (H, len(features)))
    W2 = rng.normal(0.0, math.sqrt(2.0 / H), (H, H))
    W3 = np.zeros((16, H))
    W3[:3] = rng.normal(0.0, math.sqrt(1.0 / H), (3, H)) * 0.1
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), W3.ravel(),
                           np.zeros(16)]).astype(np.float32)


def quad_norm(acc_bh, feats, G=1.0):
    """|q|, the walk's quadrupole force (features 5, 6 are its components
    along and across e1 in units of |a_bh|)."""
    a = np.linalg.norm(acc_bh, axis=1)
    f = np.asarray(feats, dtype=np.float64)
    return a * np.hypot(f[:, 5], f[:, 6])


def residual_in_frame(acc_bh, acc_ref, frames, qn):
    """(a_ref - a_bh) / |q| in the local frames (rows e1, e2, e3); zero where
    the walk accepted no cell (q = 0: no correction)."""
    r = (np.asarray(acc_ref) - np.asarray(acc_bh)) / np.where(qn > 0, qn, np.inf)[:, None]
    return np.einsum("nij,nj->ni", frames.astype(np.float64), r)


def quad_in_frame(feats):
    """q / |q| in the local frame: (cos, sin, 0) of its angle to a (features 7, 8)."""
    f = np.asarray(feats, dtype=np.float64)
    return np.stack([f[:, 7], f[:, 8], np.zeros(len(f))], axis=1)


def quad_force(acc_bh, feats, frames):
    """The walk's quadrupole force q (G applied as in acc_bh) from features 5, 6
    and the frames: |a| (f5 e1 + f6 e2)."""
    a = np.linalg.norm(acc_bh, axis=1)
    f = np.asarray(feats, dtype=np.float64)
    fr = np.asarray(frames, dtype=np.float64)
    return a[:, None] * (f[:, 5:6] * fr[:, 0] + f[:, 6:7] * fr[:, 1])


def target_in_frame(acc_bh, acc_ref, feats, frames):
    """The network's target: (a_ref - a_bh - q) / |q| in the local frame."""
    qn = quad_norm(acc_bh, feats)
    return residual_in_frame(acc_bh, acc_ref, frames, qn) - quad_in_frame(feats)


def snapshot_pairs(engine, pos, mass, theta=1.0, G=1.0, eps=0.01):
    """Features, frames, walk accelerations and FP64 direct accelerations."""
    engine.upload(pos, mass)
    engine.theta = theta
    feats, frames = engine.bh_features(G, eps, frames=True)
    a_bh = engine.accelerations()
    engine.compute_force("hip_direct_fp64", G, eps)
    a_ref = engine.accelerations()
    return feats, frames, a_bh, a_ref


def training_set(n=65536, seeds=(1, 2, 3, 4), theta=1.0, G=1.0, eps=0.01, device=0, log=None):
    """Plummer and disk snapshots -> (features, targets in frame: what the
    quadrupole force leaves of the residual, in units of |q|)."""
    X, Y = [], []
    with Engine(n, device=device, theta=theta) as eng:
        for kind in ("plummer", "disk"):
            for seed in seeds:
                d = plummer(n, seed=seed) if kind == "plummer" else disk_galaxy(n, seed=seed)
                f, fr, a_bh, a_ref = snapshot_pairs(eng, d["pos"], d["m"], theta, G, eps)
                qn = quad_norm(a_bh, f)
                r = target_in_frame(a_bh, a_ref, f, fr)
                keep = qn > 0
                X.append(f[keep])
                Y.append(r[keep])
                if log:
                    err = np.linalg.norm(a_bh - a_ref, axis=1) / np.linalg.norm(a_ref, axis=1)
                    log(f"snapshot {kind} seed {seed}: median BH error {np.median(err):.3e}")
    return np.concatenate(X), np.concatenate(Y)


def fit_scaling(feats, resid, features=FEATURES, clip=8.0):
    """Feature standardisation, robust output scale and clipped scaled targets."""
    mean = feats.mean(axis=0).astype(np.float32)
    std = feats.std(axis=0).astype(np.float32)
    unused = np.ones(N_FEATURES, bool)
    unused[list(features)] = False
    unused |= std < 1e-6
    mean[unused] = 0.0
    std[unused] = np.inf
    mad = np.median(np.abs(resid))
    out_scale = float(1.0 / (1.4826 * mad)) if mad > 0 else 1.0
    targets = np.clip(resid * out_scale, -clip, clip).astype(np.float32)
    return mean, std, out_scale, targets


def train(feats, resid, epochs=60, batch=1024, lr=2e-3, seed=0, device=0, features=FEATURES,
          log=None):
    """Trains a model on the device; returns it and a report."""
    mean, std, out_scale, targets = fit_scaling(feats, resid, features)
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(feats))
    n_val = max(len(feats) // 10, 1)
    val, tr = order[:n_val], order[n_val:]
    mlp = Mlp(device)
    mlp.set_params(init_params(seed, features), mean, std, out_scale)
    mlp.load_dataset(feats[tr], targets[tr])
    history = []
    t0 = time.time()
    for ep in range(epochs):
        lr_ep = lr * 0.5 * (1.0 + math.cos(math.pi * ep / epochs))
        loss = float(mlp.train(1, batch, lr_ep, seed * 1000003 + ep)[0])
        if ep % 5 == 4 or ep == epochs - 1:
            pv = mlp.forward(feats[val])
            vloss = float(((pv - targets[val]) ** 2).sum(1).mean())
            history.append((ep, loss, vloss))
            if log:
                log(f"epoch {ep + 1}/{epochs} train {loss:.4f} val {vloss:.4f} "
                    f"({time.time() - t0:.1f}s)")
    base = float((targets[val] ** 2).sum(1).mean())
    return mlp, {"history": history, "val_loss_zero_model": base, "out_scale": out_scale,
                 "train_samples": int(len(tr)), "val_samples": int(n_val)}


def attach(engine, path=WEIGHTS):
    """Loads the trained model and attaches it to `engine`; returns the model."""
    mlp = Mlp.load(path, device=getattr(engine, "device", 0))
    engine.set_mlp(mlp)
    return mlp


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--seeds", default="1,2,3,4")
    ap.add_argument("--theta", type=float, default=1.0)
    ap.add_argument("--eps", type=float, default=0.01)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--features", default=",".join(map(str, FEATURES)))
    a = ap.parse_args(argv)
    features = tuple(int(k) for k in a.features.split(","))

    def log(msg):
        print(msg, flush=True)

    seeds = tuple(int(s) for s in a.seeds.split(","))
    X, Y = training_set(a.n, seeds, a.theta, 1.0, a.eps, log=log)
    log(f"training set {X.shape[0]} samples")
    mlp, rep = train(X, Y, a.epochs, a.batch, a.lr, a.seed, features=features, log=log)
    meta = {"theta": a.theta, "eps": a.eps, "n": a.n, "seeds": list(seeds),
            "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "seed": a.seed,
            "features": list(features),
            "val_loss": rep["history"][-1][2] if rep["history"] else None,
            "val_loss_zero_model": rep["val_loss_zero_model"]}
    mlp.save(a.out, **meta)
    log(json.dumps(meta))
    return 0


if __name__ == "__main__":
    sys.exit(main())
