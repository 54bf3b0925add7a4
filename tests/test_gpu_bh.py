"""Barnes-Hut parity on the GPU: Morton keys, sorted order, root cube and tree
topology bit-exact against the oracle; hip_barnes_hut_fp64 accelerations and
statistics bit-identical to the reference (committed vectors and oracle);
hip_barnes_hut (FP32 walk on FP32-rounded inputs): with per-particle reference
lists close to the FP64 result and to the reference's interaction counts; the
default group walk refines those lists."""
import ctypes as C

import numpy as np
import pytest

from nbx_testlib import BarnesHutStats, Particles, gravity, rms_rel_err

pytestmark = pytest.mark.gpu


def fixture_particles(golden, name):
    n = golden[f"{name}/px"].size
    return Particles(n, {k: golden[f"{name}/{k}"] for k in ("px", "py", "pz", "vx", "vy", "vz", "m")})


def from_gen(d, fp32=False):
    pos, m = d["pos"], d["m"]
    if fp32:
        pos = pos.astype(np.float32).astype(np.float64)
        m = m.astype(np.float32).astype(np.float64)
    return Particles(len(m), {"px": pos[:, 0], "py": pos[:, 1], "pz": pos[:, 2], "m": m})


def oracle_bh(oracle, p, theta, eps):
    q = p.copy()
    t = oracle.nbo_bh_create()
    oracle.nbo_bh_set_theta(t, theta)
    st = oracle.nbo_bh_compute(t, C.byref(q.view), C.byref(gravity(1.0, eps)))
    s = BarnesHutStats()
    oracle.nbo_bh_stats(t, C.byref(s))
    keys = np.zeros(p.n, dtype=np.uint64)
    perm = np.zeros(p.n, dtype=np.int64)
    root = np.zeros(4)
    rows = None
    if st == 0:
        oracle.nbo_bh_keys(t, keys.ctypes.data_as(C.POINTER(C.c_uint64)),
                           perm.ctypes.data_as(C.POINTER(C.c_int64)))
        oracle.nbo_bh_root(t, root.ctypes.data_as(C.POINTER(C.c_double)))
        rows = np.zeros((oracle.nbo_bh_node_count(t), 11))
        oracle.nbo_bh_export_nodes(t, rows.ctypes.data_as(C.POINTER(C.c_double)))
    oracle.nbo_bh_destroy(t)
    return st, q.acc(), s, keys, perm, root, rows


def gpu_bh(nbx, p, theta, eps, algorithm, reference_lists=False):
    e = nbx.Engine(p.n, theta=theta, reference_lists=reference_lists)
    e.upload_view(p.view)
    e.compute_force(algorithm, eps=eps)
    return e, e.accelerations(), e.stats()


def sort_rows(rows):
    keys = tuple(rows[:, k] for k in (10, 3, 2, 1, 0))
    return rows[np.lexsort(keys)]


def stats_tuple(s):
    if isinstance(s, dict):
        return (s["node_count"], s["leaf_count"], s["internal_node_count"],
                s["accepted_approximations"], s["exact_interactions"], s["maximum_depth"])
    return (s.node_count, s.leaf_count, s.internal_node_count, s.accepted_approximations,
            s.exact_interactions, s.maximum_depth)


@pytest.mark.parametrize("theta", [0.1, 0.3, 0.7, 1.0, 1.2])
def test_bh_fp64_bit_identical_to_reference_vectors(nbx, golden, theta):
    p = fixture_particles(golden, "cloud400_s7")
    e, acc, s = gpu_bh(nbx, p, theta, 0.02, "hip_barnes_hut_fp64")
    e.close()
    assert np.array_equal(acc, golden[f"cloud400_s7/bh_acc_theta{theta}"])
    assert list(stats_tuple(s)) == list(golden[f"cloud400_s7/bh_stats_theta{theta}"])


@pytest.mark.parametrize("name,eps", [("spiral400_s99", 0.05), ("cloud2048_s42", 0.01)])
def test_bh_fp64_other_reference_vectors(nbx, golden, name, eps):
    p = fixture_particles(golden, name)
    e, acc, s = gpu_bh(nbx, p, 0.7, eps, "hip_barnes_hut_fp64")
    e.close()
    assert np.array_equal(acc, golden[f"{name}/bh_acc_theta0.7"])
    assert list(stats_tuple(s)) == list(golden[f"{name}/bh_stats_theta0.7"])


def crafted_inputs():
    rng = np.random.default_rng(3)
    out = {}
    # lattice: many particles exactly on cell-centre planes (Morton/octant ties)
    g = np.arange(-4, 5, dtype=np.float64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    out["lattice"] = (pos, np.ones(len(pos)))
    # the lattice plus 100 particles within 1e-6 of one point: one run of
    # equal top 36 key bits longer than the Morton sort's finish kernel takes,
    # next to the lattice's points on cell-centre planes
    cluster = np.array([0.3, 0.3, 0.3]) + rng.uniform(-1e-6, 1e-6, size=(100, 3))
    out["lattice_cluster"] = (np.concatenate([pos, cluster]), np.ones(len(pos) + 100))
    # pairs closer than one 21-bit Morton cell (tie path: deeper than 21 levels)
    base = rng.uniform(-1, 1, size=(300, 3))
    close = base + rng.uniform(-1, 1, size=(300, 3)) * 1e-9
    pos = np.concatenate([base, close, rng.uniform(-1, 1, size=(400, 3))])
    out["close_pairs"] = (pos, rng.uniform(0.5, 1.0, len(pos)))
    # a cluster + far outlier (deep chains of single-child cells)
    pos = np.concatenate([rng.normal(scale=1e-4, size=(500, 3)), [[100.0, -50.0, 20.0]]])
    out["outlier"] = (pos, np.ones(len(pos)))
    # massless particles: a cell whose particles all have zero mass keeps the
    # position of its lowest-Morton-rank particle; with no mass at all the
    # root keeps particle 0's (barnes_hut.c:230-250, :504-512)
    pos = rng.uniform(-1, 1, size=(800, 3))
    m = rng.uniform(0.5, 1.0, len(pos))
    m[pos[:, 0] < -0.3] = 0.0
    out["massless_region"] = (pos, m)
    out["massless_all"] = (rng.uniform(-1, 1, size=(300, 3)), np.zeros(300))
    return out


@pytest.mark.parametrize("case", ["lattice", "close_pairs", "outlier", "massless_region",
                                  "massless_all", "lattice_cluster"])
@pytest.mark.parametrize("theta", [0.3, 0.7])
def test_bh_fp64_crafted_inputs_bit_identical(nbx, oracle, case, theta):
    pos, m = crafted_inputs()[case]
    p = Particles(len(m), {"px": pos[:, 0], "py": pos[:, 1], "pz": pos[:, 2], "m": m})
    st, oacc, os_, okeys, operm, oroot, orows = oracle_bh(oracle, p, theta, 0.01)
    assert st == 0
    e, acc, s = gpu_bh(nbx, p, theta, 0.01, "hip_barnes_hut_fp64")
    root, keys, perm = e.bh_morton()
    rows = e.bh_export_nodes()
    e.close()
    assert np.array_equal(root, oroot)
    assert np.array_equal(keys, okeys)
    assert np.array_equal(perm, operm)
    assert stats_tuple(s) == stats_tuple(os_)
    assert np.array_equal(sort_rows(rows), sort_rows(orows))
    assert np.array_equal(acc, oacc)


def test_bh_two_galaxy_fp32_inputs_keys_order_topology(nbx, oracle):
    """Config 2 shape (scaled to 200k): FP32-rounded positions/masses; the
    63-bit keys (and hence key >> 33), sorted order and tree are bit-exact."""
    d = nbx.two_galaxy(200_000, seed=42)
    p = from_gen(d, fp32=True)
    st, oacc, os_, okeys, operm, oroot, orows = oracle_bh(oracle, p, 0.7, 0.01)
    assert st == 0
    raw = from_gen(d, fp32=False)
    e, acc, s = gpu_bh(nbx, raw, 0.7, 0.01, "hip_barnes_hut", reference_lists=True)
    root, keys, perm = e.bh_morton()
    rows = e.bh_export_nodes()
    e.close()
    assert np.array_equal(root, oroot)
    assert np.array_equal(keys, okeys)
    assert np.array_equal(keys >> np.uint64(33), okeys >> np.uint64(33))
    assert np.array_equal(perm, operm)
    s_t, o_t = stats_tuple(s), stats_tuple(os_)
    assert s_t[:3] == o_t[:3] and s_t[5] == o_t[5]
    assert np.array_equal(sort_rows(rows)[:, :7], sort_rows(orows)[:, :7])
    # FP32 walk vs the FP64 reference walk on identical inputs
    err = rms_rel_err(acc, oacc)
    print(f"fp32 walk rms rel err {err:.3e}; interactions gpu {s_t[3] + s_t[4]} cpu {o_t[3] + o_t[4]}")
    assert err < 1e-4
    assert abs((s_t[3] + s_t[4]) - (o_t[3] + o_t[4])) <= 1e-3 * (o_t[3] + o_t[4])


@pytest.mark.parametrize("case", ["lattice", "outlier", "massless_region", "lattice_cluster"])
def test_bh_fp32_inputs_crafted_keys_order_topology(nbx, oracle, case):
    """hip_barnes_hut (FP32-rounded inputs, float4 build inputs) on the
    crafted inputs: keys, sorted order, root cube and node rows bit-exact
    against the oracle's build on the same rounded inputs, and the default
    group walk close to the reference walk. The cluster plus far outlier puts
    500 particles in one run of equal top 36 key bits, more than the Morton
    sort's finish kernel takes: that build sorts again on all 63 bits."""
    pos, m = crafted_inputs()[case]
    raw = Particles(len(m), {"px": pos[:, 0], "py": pos[:, 1], "pz": pos[:, 2], "m": m})
    pos32 = pos.astype(np.float32).astype(np.float64)
    m32 = m.astype(np.float32).astype(np.float64)
    p32 = Particles(len(m), {"px": pos32[:, 0], "py": pos32[:, 1], "pz": pos32[:, 2], "m": m32})
    st, oacc, os_, okeys, operm, oroot, orows = oracle_bh(oracle, p32, 0.7, 0.01)
    assert st == 0
    e, acc, s = gpu_bh(nbx, raw, 0.7, 0.01, "hip_barnes_hut")
    root, keys, perm = e.bh_morton()
    rows = e.bh_export_nodes()
    e.close()
    assert np.array_equal(root, oroot)
    assert np.array_equal(keys, okeys)
    assert np.array_equal(perm, operm)
    assert np.array_equal(sort_rows(rows)[:, :7], sort_rows(orows)[:, :7])
    assert rms_rel_err(acc, oacc) < 0.05


@pytest.mark.parametrize("n", [200_000, 4_500_000])
def test_group_walk_refines_reference_lists(nbx, oracle, n):
    """The FP32 group walk accepts a cell only when the whole group's box
    passes the reference test (NBX_FLAG_BH_BOX_TEST) or, by default, also when
    every member's own reference test passes, so it makes at least the
    reference's interactions and is at least as accurate (FP64 direct sum on
    a sample; the default lists are closest to the reference lists, so their
    median gets 2% slack)."""
    d = nbx.two_galaxy(n, seed=42)
    idx = np.sort(np.random.default_rng(3).choice(n, 2048, replace=False))
    out = {}
    for key, kw in (("reference", {"reference_lists": True}), ("default", {}),
                    ("box", {"member_test": False})):
        with nbx.Engine(n, theta=0.7, **kw) as e:
            e.upload(d["pos"], d["m"])
            e.compute_force("hip_barnes_hut", eps=0.01)
            s = e.stats()
            acc = e.accelerations()
            ref = e.direct_fp64_targets(idx, 1.0, 0.01)
        err = np.linalg.norm(acc[idx] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        out[key] = (s["accepted_approximations"] + s["exact_interactions"], err)
        print(f"{key}: interactions {out[key][0]} "
              f"median err {np.median(err):.3e} p99 {np.quantile(err, 0.99):.3e}")
    assert out["box"][0] > out["default"][0] >= out["reference"][0]
    assert np.median(out["box"][1]) <= np.median(out["reference"][1])
    assert np.quantile(out["box"][1], 0.99) <= 1.05 * np.quantile(out["reference"][1], 0.99)
    assert np.median(out["default"][1]) <= 1.02 * np.median(out["reference"][1])
    assert np.quantile(out["default"][1], 0.99) <= 1.05 * np.quantile(out["reference"][1], 0.99)
    # the group walk visits the same tree: momentum of the force field stays ~0
    with nbx.Engine(n, theta=0.7) as e:
        e.upload(d["pos"], d["m"])
        e.compute_force("hip_barnes_hut", eps=0.01)
        a = e.accelerations()
    f = (d["m"][:, None] * a).sum(0)
    assert np.linalg.norm(f) < 1e-3 * (d["m"][:, None] * np.abs(a)).sum()


def test_bh_fp64_two_galaxy_bit_identical(nbx, oracle):
    d = nbx.two_galaxy(30_000, seed=5)
    p = from_gen(d)
    st, oacc, os_, *_ = oracle_bh(oracle, p, 0.7, 0.01)
    e, acc, s = gpu_bh(nbx, p, 0.7, 0.01, "hip_barnes_hut_fp64")
    e.close()
    assert np.array_equal(acc, oacc)
    assert stats_tuple(s) == stats_tuple(os_)


def test_bh_single_particle_and_pair(nbx, oracle):
    for pos, m in [([[0.3, -0.2, 0.1]], [2.0]), ([[-0.5, 0, 0], [0.5, 0, 0]], [1.0, 1.0])]:
        pos = np.array(pos, dtype=np.float64)
        p = Particles(len(m), {"px": pos[:, 0], "py": pos[:, 1], "pz": pos[:, 2], "m": m})
        st, oacc, os_, *_ = oracle_bh(oracle, p, 0.7, 0.02)
        e, acc, s = gpu_bh(nbx, p, 0.7, 0.02, "hip_barnes_hut_fp64")
        e.close()
        assert np.array_equal(acc, oacc)
        assert stats_tuple(s) == stats_tuple(os_)


def test_bh_duplicate_positions_overflow(nbx, oracle):
    pos = np.array([[0.1, 0.2, 0.3]] * 2 + [[1.0, 1.0, 1.0]])
    p = Particles(3, {"px": pos[:, 0], "py": pos[:, 1], "pz": pos[:, 2], "m": np.ones(3)})
    st, *_ = oracle_bh(oracle, p, 0.7, 0.01)
    assert st == 6  # OVERFLOW (barnes_hut.c:149-153)
    e = nbx.Engine(3)
    e.upload_view(p.view)
    with pytest.raises(nbx.NBXError) as ex:
        e.compute_force("hip_barnes_hut_fp64", eps=0.01)
    assert ex.value.status == 6
    e.close()


def test_bh_empty_system_is_invalid(nbx):
    e = nbx.Engine(4)
    with pytest.raises(nbx.NBXError) as ex:
        e.compute_force("hip_barnes_hut", eps=0.01)
    assert ex.value.status == 1
    e.close()


def test_group_walk_fallback_is_the_reference_list_walk(nbx):
    """With the split off (NBX_FLAG_BH_NO_SPLIT) a group that is not compact
    walks each particle with its reference list: those particles get exactly
    the per-particle walk's accelerations, every other particle exactly the
    default group walk's (its group is compact, one shared walk either way)."""
    n = 200_000
    d = nbx.two_galaxy(n, seed=42)
    acc = {}
    for key, kw in (("default", {}), ("no_split", {"split": False}),
                    ("reference", {"reference_lists": True})):
        with nbx.Engine(n, theta=0.7, **kw) as e:
            e.upload(d["pos"], d["m"])
            e.compute_force("hip_barnes_hut", eps=0.01)
            acc[key] = e.accelerations()
    same_group = np.all(acc["no_split"] == acc["default"], axis=1)
    same_ref = np.all(acc["no_split"] == acc["reference"], axis=1)
    assert np.all(same_group | same_ref)
    frac = 1.0 - same_group.mean()
    print(f"particles on the per-particle fallback: {frac:.3%}")
    assert 0.02 < frac < 0.3          # the non-compact groups (11.6% at 1M)
    assert not np.any(same_ref & ~same_group & np.all(acc["reference"] == acc["default"], axis=1))


def test_bh_phase_timings_are_consistent(nbx):
    """The per-phase HIP-event timings of a Barnes-Hut evaluation: every phase
    is positive, and bounding cube/keys, sort, tree construction, centre of
    mass and walk fit inside the whole evaluation (their brackets do not
    overlap)."""
    n = 100_000
    d = nbx.two_galaxy(n, seed=5)
    with nbx.Engine(n, theta=0.7) as e:
        e.upload(d["pos"], d["m"])
        for alg in ("hip_barnes_hut", "hip_barnes_hut_fp64", "hip_bh_learned"):
            e.compute_force(alg, eps=0.01)
            t = e.timings()
            phases = ("bbox_keys_ms", "sort_ms", "tree_ms", "com_ms", "walk_ms")
            assert all(t[k] > 0.0 for k in phases), (alg, t)
            assert sum(t[k] for k in phases) <= t["force_ms"] * 1.001 + 1e-3, (alg, t)
            assert t["kernel_ms"] == t["walk_ms"]


@pytest.mark.parametrize("theta", [0.7, 1.0])
def test_member_test_still_refines_reference_lists(nbx, theta):
    """The group walk's member test (its default): a cell the group's box
    test opens is accepted when every member's own test accepts it. The
    shared lists shrink (fewer interactions than the box test alone,
    NBX_FLAG_BH_BOX_TEST) but still refine every member's reference list (no
    fewer than the reference lists), and the accuracy against the FP64 direct
    sum stays at least the reference lists'."""
    n = 200_000
    d = nbx.two_galaxy(n, seed=42)
    idx = np.sort(np.random.default_rng(3).choice(n, 2048, replace=False))
    out = {}
    for key, kw in (("reference", {"reference_lists": True}), ("box", {"member_test": False}),
                    ("member", {})):
        with nbx.Engine(n, theta=theta, **kw) as e:
            e.upload(d["pos"], d["m"])
            e.compute_force("hip_barnes_hut", eps=0.01)
            s = e.stats()
            acc = e.accelerations()
            ref = e.direct_fp64_targets(idx, 1.0, 0.01)
        err = np.linalg.norm(acc[idx] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        out[key] = (s["accepted_approximations"] + s["exact_interactions"], np.median(err))
        print(f"theta={theta} {key}: interactions {out[key][0]} median err {out[key][1]:.3e}")
    assert out["reference"][0] <= out["member"][0] < out["box"][0]
    # nested lists: the box lists refine the member lists, which refine the
    # reference lists (medians of 2048 sampled particles, 2% slack)
    assert out["box"][1] <= 1.02 * out["member"][1]
    assert out["member"][1] <= 1.02 * out["reference"][1]
