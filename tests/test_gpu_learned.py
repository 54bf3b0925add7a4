"""Learned Barnes-Hut correction on the GPU (include/nbx_learned.h).

- the v_mfma_f32_16x16x16_bf16 lane layout the kernels assume;
- MLP inference and gradients against the CPU oracle (oracle/mlp_oracle.py);
- training lowers the loss; the feature walk against a numpy restatement of
  the walk over the exported tree; the correction kernel against the host
  formula; without a model (or with a zero model) hip_bh_learned is plain BH;
- the committed model lowers the theta = 1.0 error on held-out snapshots.
"""
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, str(ROOT / "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def learned(nbx):
    from n_body_sim_pro_amd import learned as L
    return L


def _random_model(seed):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed)
    p[mo.OFF_W3:mo.OFF_W3 + 3 * mo.H] *= 10.0
    p[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.1, mo.H)
    p[mo.OFF_B2:mo.OFF_W3] = rng.normal(0, 0.1, mo.H)
    p[mo.OFF_B3:mo.OFF_B3 + 3] = rng.normal(0, 0.1, 3)
    mean = rng.normal(0, 0.2, mo.F).astype(np.float32)
    std = rng.uniform(0.5, 2.0, mo.F).astype(np.float32)
    return p, mean, std


def _feats(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.normal(0, 1.5, (n, mo.F)).astype(np.float32)
    f[:, 8:] = 0.0
    return f


def test_mfma_bf16_lane_layout(nbx, learned):
    rng = np.random.default_rng(0)
    a = rng.normal(0, 1, (16, 16)).astype(np.float32)
    b = rng.normal(0, 1, (16, 16)).astype(np.float32)
    d = learned.probe_mfma(a, b)
    ref = mo.bf16(a).astype(np.float64) @ mo.bf16(b).astype(np.float64)
    assert np.allclose(d, ref, rtol=1e-6, atol=1e-5 * np.abs(ref).max())
    # a transposed operand would not pass
    assert not np.allclose(d, mo.bf16(a).astype(np.float64) @ mo.bf16(b).T.astype(np.float64),
                           rtol=1e-3)


@pytest.mark.parametrize("n", [1, 17, 1000, 4096])
def test_forward_matches_oracle(learned, n):
    p, mean, std = _random_model(1)
    f = _feats(n, 2)
    with learned.Mlp() as m:
        m.set_params(p, mean, std, 1.0)
        y = m.forward(f)
    yo, _ = mo.forward(p, mean, std, f)
    diff = np.abs(y - yo)
    scale = np.abs(yo).max()
    print(f"n={n} forward max |diff| {diff.max():.3e} median {np.median(diff):.3e} scale {scale:.3e}")
    assert np.median(diff) <= 1e-5 * scale + 1e-7
    assert diff.max() <= 2e-2 * scale


@pytest.mark.parametrize("n", [1, 17, 1000, 5000])
def test_gradients_match_oracle(learned, n):
    p, mean, std = _random_model(3)
    f = _feats(n, 4)
    t = np.random.default_rng(5).normal(0, 1, (n, 3)).astype(np.float32)
    with learned.Mlp() as m:
        m.set_params(p, mean, std, 1.0)
        g, L = m.gradients(f, t)
    go, Lo = mo.gradients(p, mean, std, f, t)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    print(f"n={n} loss {L:.6f} oracle {Lo:.6f} grad rel norm diff {rel:.3e}")
    assert L == pytest.approx(Lo, rel=1e-3)
    assert rel < 1e-2
    for lo, hi in ((mo.OFF_W1, mo.OFF_B1), (mo.OFF_W2, mo.OFF_B2), (mo.OFF_W3, mo.OFF_B3)):
        blk = np.linalg.norm(g[lo:hi] - go[lo:hi]) / np.linalg.norm(go[lo:hi])
        assert blk < 2e-2, (lo, blk)
    assert not g[mo.OFF_W3 + 3 * mo.H:mo.OFF_B3].any() and not g[mo.OFF_B3 + 3:].any()


def test_training_lowers_loss(learned):
    n = 32768
    f = _feats(n, 6)
    t = np.stack([np.tanh(f[:, 0] + 0.5 * f[:, 1]), 0.3 * f[:, 2] * f[:, 3],
                  0.2 * np.sin(f[:, 4])], 1).astype(np.float32)
    with learned.Mlp() as m:
        m.set_params(learned.init_params(0), np.zeros(16), np.ones(16), 1.0)
        _, l0 = m.gradients(f, t)
        m.load_dataset(f, t)
        hist = m.train(15, 512, 3e-3, 1)
        _, l1 = m.gradients(f, t)
    print("loss history", np.round(hist, 4), "start", l0, "end", l1)
    assert l1 < 0.15 * l0
    assert hist[-1] < hist[0]


# ---- feature walk: numpy restatement over the exported tree ----

def _walk_features_numpy(rows, perm, pos32, m32, root_h, theta, eps2):
    """Per-particle walk statistics over pre-order node rows, in FP64. Each
    cell's quadrupole Q = sum m (3 x x^T - |x|^2 I) about its centre of mass is
    summed directly over the particles of its subtree."""
    nodes = len(rows)
    depth = rows[:, 0].astype(int)
    skip = np.empty(nodes, dtype=int)
    stack = []
    for k in range(nodes):
        while stack and depth[stack[-1]] >= depth[k]:
            skip[stack.pop()] = k
        stack.append(k)
    for k in stack:
        skip[k] = nodes
    n = len(pos32)
    q = pos32.astype(np.float64)
    leaf_particle = np.where(rows[:, 10] >= 0, perm[np.maximum(rows[:, 10], 0).astype(int)], -1)
    qm = m32.astype(np.float64)
    quad = np.zeros((nodes, 3, 3))
    for k in range(nodes):
        if leaf_particle[k] >= 0:
            continue
        sub = leaf_particle[k:skip[k]]
        sub = sub[sub >= 0]
        x = q[sub] - rows[k, 7:10]
        m2 = np.einsum("p,pi,pj->ij", qm[sub], x, x)
        quad[k] = 3 * m2 - np.trace(m2) * np.eye(3)
    skip_until = np.zeros(n, dtype=int)
    tot = np.zeros((n, 3))
    qf = np.zeros((n, 3))
    n_cells = np.zeros(n)
    n_leaves = np.zeros(n)
    q2_sum = np.zeros(n)
    q2_min = np.full(n, np.inf)
    leaf_depth = np.zeros(n, dtype=int)
    idx = np.arange(n)
    theta2 = theta * theta
    for k in range(nodes):
        active = skip_until <= k
        com = rows[k, 7:10]
        d = com[None, :] - q
        d2 = (d * d).sum(1)
        r2 = d2 + eps2
        s = rows[k, 6] / (r2 * np.sqrt(r2))
        if leaf_particle[k] >= 0:
            leaf_depth[leaf_particle[k]] = depth[k]
            use = active & (idx != leaf_particle[k])
            tot[use] += s[use, None] * d[use]
            n_leaves[use] += 1
        else:
            size2 = 4.0 * rows[k, 4] ** 2
            acc = active & (d2 > 0) & (size2 < theta2 * d2)
            tot[acc] += s[acc, None] * d[acc]
            da = d[acc]
            Qd = da @ quad[k]
            dQd = (da * Qd).sum(1)
            ra = r2[acc]
            qf[acc] += (2.5 * dQd / ra)[:, None] * da / ra[:, None] ** 2.5 - Qd / ra[:, None] ** 2.5
            n_cells[acc] += 1
            q2 = d2[acc] / size2
            q2_sum[acc] += q2
            q2_min[acc] = np.minimum(q2_min[acc], q2)
            skip_until[acc] = skip[k]
    a = np.linalg.norm(tot, axis=1)
    e1 = tot / a[:, None]
    q_par = (qf * e1).sum(1)
    perp = qf - q_par[:, None] * e1
    q_perp = np.linalg.norm(perp, axis=1)
    e2 = perp / np.where(q_perp > 0, q_perp, 1)[:, None]
    qn = np.hypot(q_par, q_perp)
    has = n_cells > 0
    f = np.zeros((n, 16))
    f[:, 0] = np.where(has, np.sqrt(q2_sum / np.maximum(n_cells, 1)), 0)
    f[:, 1] = np.where(has, np.sqrt(np.where(has, q2_min, 0)), 0)
    f[:, 2] = np.log2(m32) - 3 * (np.log2(2 * root_h) - leaf_depth)
    f[:, 3] = np.log2(1 + n_cells)
    f[:, 4] = np.log2(1 + n_leaves)
    f[:, 5] = q_par / a
    f[:, 6] = q_perp / a
    with np.errstate(invalid="ignore", divide="ignore"):
        f[:, 7] = np.where(qn > 0, q_par / qn, 0)
        f[:, 8] = np.where(qn > 0, q_perp / qn, 0)
    f[:, 9] = np.log2(qn / a + 2.0 ** -40)
    return f, e1, e2, q_perp / a, n_cells, n_leaves


def test_feature_walk_matches_numpy_restatement(nbx):
    d = nbx.plummer(3000, seed=11)
    theta, eps = 1.0, 0.01
    with nbx.Engine(3000, theta=theta, reference_lists=True) as e:
        e.upload(d["pos"], d["m"])
        feats, frames = e.bh_features(1.0, eps, frames=True)
        acc_feat = e.accelerations()
        stats = e.stats()
        rows = e.bh_export_nodes()
        root, _, perm = e.bh_morton()
        e.compute_force("hip_barnes_hut", 1.0, eps)
        acc_plain = e.accelerations()
    # the feature walk leaves exactly the plain walk's accelerations
    assert np.array_equal(acc_feat, acc_plain)
    # per-particle counts add up to the walk's own totals
    cells = np.rint(np.exp2(feats[:, 3].astype(np.float64)) - 1)
    leaves = np.rint(np.exp2(feats[:, 4].astype(np.float64)) - 1)
    assert cells.sum() == stats["accepted_approximations"]
    assert leaves.sum() == stats["exact_interactions"]
    # frames are orthonormal and right-handed
    g = np.einsum("nij,nkj->nik", frames.astype(np.float64), frames.astype(np.float64))
    assert np.abs(g - np.eye(3)).max() < 1e-5
    assert np.abs(np.linalg.det(frames.astype(np.float64)) - 1).max() < 1e-5
    pos32 = d["pos"].astype(np.float32)
    m32 = d["m"].astype(np.float32)
    f, e1, e2, qperp, nc, nl = _walk_features_numpy(rows, perm, pos32, m32, root[3], theta,
                                                    eps * eps)
    same = (nc == cells) & (nl == leaves)
    print(f"particles with identical interaction lists: {same.mean():.4f}")
    assert same.mean() > 0.99
    fs, fg = f[same], feats[same].astype(np.float64)
    assert np.allclose(fs[:, 0:2], fg[:, 0:2], rtol=1e-5, atol=1e-6)
    assert np.abs(fs[:, 2] - fg[:, 2]).max() < 1e-4
    assert np.array_equal(np.rint(np.exp2(fs[:, 3:5]) - 1), np.rint(np.exp2(fg[:, 3:5]) - 1))
    # quadrupole force: FP32 walk vs FP64 moments summed over each subtree
    assert np.abs(fs[:, 5:7] - fg[:, 5:7]).max() < 1e-4
    okq = np.hypot(fs[:, 5], fs[:, 6]) > 1e-4
    assert okq.mean() > 0.9
    assert np.abs(fs[okq, 7:9] - fg[okq, 7:9]).max() < 1e-3
    assert np.abs(fs[okq, 9] - fg[okq, 9]).max() < 1e-3
    assert not fg[:, 10:].any()
    assert np.abs(frames[same, 0].astype(np.float64) - e1[same]).max() < 1e-5
    ok = same & (qperp > 1e-3)
    assert ok.mean() > 0.3
    assert np.abs(frames[ok, 1].astype(np.float64) - e2[ok]).max() < 1e-3


def test_group_walk_features_are_consistent(nbx):
    """Default engine: the feature walk is the group walk with the member
    test. Its accelerations are the plain group walk's with that test up to
    FP32 summation order (it flushes its staged lists in smaller batches), the
    per-particle counts add up to the walk's totals (every member of a group
    shares the list length), frames are orthonormal and the quadrupole
    features are consistent."""
    d = nbx.two_galaxy(50_000, seed=8)
    with nbx.Engine(50_000, theta=1.0, member_test=True) as e:
        e.upload(d["pos"], d["m"])
        feats, frames = e.bh_features(1.0, 0.01, frames=True)
        acc_feat = e.accelerations()
        stats = e.stats()
        e.compute_force("hip_barnes_hut", 1.0, 0.01)
        plain = e.accelerations()
        plain_stats = e.stats()
    a = np.linalg.norm(plain, axis=1)
    assert (np.linalg.norm(acc_feat - plain, axis=1) / a).max() < 1e-5
    assert (stats["accepted_approximations"], stats["exact_interactions"]) == (
        plain_stats["accepted_approximations"], plain_stats["exact_interactions"])
    cells = np.rint(np.exp2(feats[:, 3].astype(np.float64)) - 1)
    leaves = np.rint(np.exp2(feats[:, 4].astype(np.float64)) - 1)
    assert cells.sum() == stats["accepted_approximations"]
    assert leaves.sum() == stats["exact_interactions"]
    g = np.einsum("nij,nkj->nik", frames.astype(np.float64), frames.astype(np.float64))
    assert np.abs(g - np.eye(3)).max() < 1e-5
    f = feats.astype(np.float64)
    has = cells > 0
    assert np.all(f[has, 1] >= 1.0 - 1e-6)        # accepted cells: d / edge > 1 / theta = 1
    assert np.all(f[has, 0] >= f[has, 1] - 1e-5)  # rms >= minimum
    qn = np.hypot(f[:, 5], f[:, 6])
    okq = has & (qn > 0)
    assert np.all(np.abs(np.hypot(f[okq, 7], f[okq, 8]) - 1) < 1e-4)   # cos^2 + sin^2 = 1
    assert np.allclose(np.exp2(f[okq, 9]), qn[okq], rtol=1e-3)
    # the quadrupole force is the bulk of the theta = 1 monopole error here
    assert 1e-3 < np.median(qn[okq]) < 1e-1
    assert not f[:, 10:].any()


def test_feature_walk_fallback_is_the_reference_list_walk(nbx):
    """With the split off (NBX_FLAG_BH_NO_SPLIT) the feature walk leaves the
    particles of non-compact groups to the per-query walk that runs after
    it: their raw features are exactly the reference-list engine's, every
    other particle's exactly the default feature walk's."""
    n = 200_000
    d = nbx.two_galaxy(n, seed=42)
    out = {}
    for key, kw in (("default", {}), ("no_split", {"split": False}),
                    ("reference", {"reference_lists": True})):
        with nbx.Engine(n, theta=1.0, **kw) as e:
            e.upload(d["pos"], d["m"])
            out[key] = e.bh_features(1.0, 0.01)
    assert all(np.isfinite(f).all() for f in out.values())
    same_group = np.all(out["no_split"] == out["default"], axis=1)
    same_ref = np.all(out["no_split"] == out["reference"], axis=1)
    assert np.all(same_group | same_ref)
    frac = 1.0 - same_group.mean()
    print(f"particles on the per-query feature walk: {frac:.3%}")
    assert 0.02 < frac < 0.3


def test_correction_kernel_matches_host_formula(nbx, learned):
    d = nbx.plummer(20000, seed=12)
    p, mean, std = _random_model(7)
    out_scale = 1.0
    with nbx.Engine(20000, theta=1.0) as e, learned.Mlp() as m:
        m.set_params(p, mean, std, out_scale)
        e.upload(d["pos"], d["m"])
        feats, frames = e.bh_features(1.0, 0.01, frames=True)
        a_bh = e.accelerations()
        y = m.forward(feats)
        e.set_mlp(m)
        e.compute_force("hip_bh_learned", 1.0, 0.01)
        a_l = e.accelerations()
    a = np.linalg.norm(a_bh, axis=1)
    qn = learned.quad_norm(a_bh, feats)
    q = learned.quad_force(a_bh, feats, frames)
    want = a_bh + q + (qn / out_scale)[:, None] * np.einsum("ni,nij->nj", y.astype(np.float64),
                                                             frames.astype(np.float64))
    corr = np.linalg.norm(want - a_bh, axis=1)
    assert np.median(corr / a) > 1e-3                    # the correction is not trivial
    err = np.linalg.norm(a_l - want, axis=1) / a
    print(f"apply vs host formula: max {err.max():.3e}")
    assert err.max() < 1e-5


def test_without_model_hip_bh_learned_is_plain_bh(nbx, learned):
    """No model: hip_bh_learned is hip_barnes_hut, bit for bit (default group
    walk and reference lists alike). A zero model gives the corrected path's
    base, here the per-particle walk with reference lists, plus its quadrupole
    force and nothing else."""
    d = nbx.two_galaxy(30000, seed=3)
    for ref_lists in (False, True):
        with nbx.Engine(30000, theta=1.0, reference_lists=ref_lists) as e:
            e.upload(d["pos"], d["m"])
            e.compute_force("hip_barnes_hut", 1.0, 0.01)
            plain = e.accelerations()
            e.compute_force("hip_bh_learned", 1.0, 0.01)
            assert np.array_equal(e.accelerations(), plain)
            if ref_lists:
                feats, frames = e.bh_features(1.0, 0.01, frames=True)
                assert np.array_equal(e.accelerations(), plain)
                q = learned.quad_force(plain, feats, frames)
                with learned.Mlp() as m:
                    m.set_params(np.zeros(m.n_params), np.zeros(16), np.ones(16), 1.0)
                    e.set_mlp(m)
                    e.compute_force("hip_bh_learned", 1.0, 0.01)
                    dq = e.accelerations() - plain - q
                    qn = np.linalg.norm(q, axis=1)
                    assert np.all(np.linalg.norm(dq, axis=1) <= 1e-5 * qn + 1e-12)
                    e.set_mlp(None)
            e.compute_force("hip_bh_learned", 1.0, 0.01)
            assert np.array_equal(e.accelerations(), plain)


def _errors(nbx, learned, d, mlp, eps=0.01):
    """The corrected path's base walk (the group walk with the member test)
    vs the corrected walk, both against the FP64 direct sum."""
    n = len(d["m"])
    with nbx.Engine(n, theta=1.0, member_test=True) as e:
        e.upload(d["pos"], d["m"])
        e.compute_force("hip_direct_fp64", 1.0, eps)
        ref = e.accelerations()
        e.compute_force("hip_barnes_hut", 1.0, eps)
        plain = e.accelerations()
        e.set_mlp(mlp)
        e.compute_force("hip_bh_learned", 1.0, eps)
        corr = e.accelerations()
        e.set_mlp(None)
    r = np.linalg.norm(ref, axis=1)
    return (np.linalg.norm(plain - ref, axis=1) / r, np.linalg.norm(corr - ref, axis=1) / r)


@pytest.mark.parametrize("kind,seed,ratio", [("plummer", 101, 0.8), ("disk", 102, 0.25),
                                             ("two_galaxy", 103, 0.25)])
def test_committed_model_lowers_theta1_error(nbx, learned, kind, seed, ratio):
    """Held-out snapshots (seeds outside the training set; the two-galaxy
    system is not in it at all): median error of theta = 1.0 + model over
    plain theta = 1.0 must stay below `ratio`."""
    assert learned.WEIGHTS.exists(), f"{learned.WEIGHTS} is missing"
    gen = {"plummer": nbx.plummer, "disk": nbx.disk_galaxy, "two_galaxy": nbx.two_galaxy}[kind]
    d = gen(65536, seed=seed)
    mlp = learned.Mlp.load(learned.WEIGHTS)
    plain, corr = _errors(nbx, learned, d, mlp)
    mlp.close()
    print(f"{kind}: median error plain {np.median(plain):.4e} learned {np.median(corr):.4e} "
          f"p90 plain {np.quantile(plain, 0.9):.4e} learned {np.quantile(corr, 0.9):.4e}")
    assert np.median(corr) < ratio * np.median(plain)


def test_committed_model_beats_theta05(nbx, learned):
    """The accuracy half of the learned target at a size the FP64 direct sum
    covers in full: theta = 1.0 + model has a lower median error than plain
    theta = 0.5 on a 400k two-galaxy collision (the model was trained on 65,536
    particle Plummer and disk snapshots only)."""
    n = 400_000
    d = nbx.two_galaxy(n, seed=44)
    idx = np.sort(np.random.default_rng(5).choice(n, 4096, replace=False))
    mlp = learned.Mlp.load(learned.WEIGHTS)
    with nbx.Engine(n, theta=0.5) as e:
        e.upload(d["pos"], d["m"])
        ref = e.direct_fp64_targets(idx, 1.0, 0.01)
        e.compute_force("hip_barnes_hut", 1.0, 0.01)
        half = e.accelerations()[idx]
        e.theta = 1.0
        e.set_mlp(mlp)
        e.compute_force("hip_bh_learned", 1.0, 0.01)
        corr = e.accelerations()[idx]
        e.set_mlp(None)
    mlp.close()
    r = np.linalg.norm(ref, axis=1)
    e_half = np.median(np.linalg.norm(half - ref, axis=1) / r)
    e_corr = np.median(np.linalg.norm(corr - ref, axis=1) / r)
    print(f"median error: plain theta=0.5 {e_half:.4e}, theta=1.0 + model {e_corr:.4e}")
    assert e_corr < e_half


def test_network_gain_over_the_quadrupole_alone(nbx, learned):
    """What the network adds: on a held-out 2M two-galaxy collision (neither
    the size nor the system is in the training set), theta = 1.0 with the
    quadrupole force and the network has a lower median error (FP64 direct
    sum on 8192 seeded particles) than the same walk with the quadrupole force
    alone, i.e. the zero model."""
    n = 2_000_000
    d = nbx.two_galaxy(n, seed=44)
    idx = np.sort(np.random.default_rng(7).choice(n, 8192, replace=False))
    mlp = learned.Mlp.load(learned.WEIGHTS)
    with nbx.Engine(n, theta=1.0) as e, learned.Mlp() as zero:
        zero.set_params(np.zeros(zero.n_params), np.zeros(16), np.ones(16), 1.0)
        e.upload(d["pos"], d["m"])
        ref = e.direct_fp64_targets(idx, 1.0, 0.01)
        out = {}
        for name, model in (("quadrupole", zero), ("network", mlp)):
            e.set_mlp(model)
            e.compute_force("hip_bh_learned", 1.0, 0.01)
            a = e.accelerations()[idx]
            out[name] = float(np.median(np.linalg.norm(a - ref, axis=1)
                                        / np.linalg.norm(ref, axis=1)))
        e.set_mlp(None)
    mlp.close()
    print(f"median error: quadrupole alone {out['quadrupole']:.4e}, with the network "
          f"{out['network']:.4e} (ratio {out['network'] / out['quadrupole']:.3f})")
    assert out["network"] < out["quadrupole"]
