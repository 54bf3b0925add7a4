"""CPU tests that pin the oracle (oracle/*.c) before it is trusted as the checker:
committed reference vectors (tests/golden, written by the reference engine built
from its own sources), known-answer tests from the reference's test suite,
closed forms, brute force and textbook routines."""
import ctypes as C

import numpy as np
import pytest

from nbx_testlib import (BarnesHutStats, Diagnostics, Error, Particles, gravity, oracle_preset,
                         ref_preset, rms_rel_err, PRESET_GLOBULAR_CLUSTER, PRESET_RANDOM_CLOUD,
                         PRESET_SPIRAL_GALAXY, PRESET_TWO_BODY)

FIX = {
    "cloud400_s7": (PRESET_RANDOM_CLOUD, 400, 7),
    "cloud300_s7": (PRESET_RANDOM_CLOUD, 300, 7),
    "cloud2048_s42": (PRESET_RANDOM_CLOUD, 2048, 42),
    "spiral400_s99": (PRESET_SPIRAL_GALAXY, 400, 99),
    "globular1000_s3": (PRESET_GLOBULAR_CLUSTER, 1000, 3),
    "twobody": (PRESET_TWO_BODY, 2, 1),
}


def fixture_particles(golden, name):
    n = golden[f"{name}/px"].size
    return Particles(n, {k: golden[f"{name}/{k}"] for k in ("px", "py", "pz", "vx", "vy", "vz", "m")})


@pytest.mark.parametrize("name", sorted(FIX))
def test_oracle_presets_bit_identical_to_reference_vectors(oracle, golden, name):
    preset, n, seed = FIX[name]
    p = oracle_preset(oracle, preset, n, seed)
    for k in ("px", "py", "pz", "vx", "vy", "vz", "m"):
        assert np.array_equal(p[k], golden[f"{name}/{k}"]), k


@pytest.mark.parametrize("name,eps", [("cloud2048_s42", 0.01), ("cloud300_s7", 0.02),
                                      ("globular1000_s3", 0.01)])
def test_oracle_direct_bit_identical_to_reference_vectors(oracle, golden, name, eps):
    p = fixture_particles(golden, name)
    g = gravity(1.0, eps)
    assert oracle.nbo_direct_fp64(C.byref(p.view), C.byref(g)) == 0
    assert np.array_equal(p.acc(), golden[f"{name}/acc_reference_eps{eps}"])
    q = fixture_particles(golden, name)
    assert oracle.nbo_direct_openmp(C.byref(q.view), C.byref(g)) == 0
    assert np.array_equal(q.acc(), p.acc())


def test_reference_openmp_avx2_within_reference_tolerance(golden):
    # simd_equivalence_test.c:35-53: max component relative error < 1e-10 (1e-12 floor)
    a = golden["cloud300_s7/acc_openmp_avx2_eps0.02"]
    b = golden["cloud300_s7/acc_reference_eps0.02"]
    scale = np.where(np.abs(b) > 1e-12, np.abs(b), 1.0)
    assert np.max(np.abs(a - b) / scale) < 1e-10


@pytest.mark.parametrize("theta", [0.1, 0.3, 0.7, 1.0, 1.2])
def test_oracle_barnes_hut_bit_identical_to_reference_vectors(oracle, golden, theta):
    p = fixture_particles(golden, "cloud400_s7")
    t = oracle.nbo_bh_create()
    oracle.nbo_bh_set_theta(t, theta)
    assert oracle.nbo_bh_compute(t, C.byref(p.view), C.byref(gravity(1.0, 0.02))) == 0
    s = BarnesHutStats()
    oracle.nbo_bh_stats(t, C.byref(s))
    oracle.nbo_bh_destroy(t)
    assert np.array_equal(p.acc(), golden[f"cloud400_s7/bh_acc_theta{theta}"])
    st = golden[f"cloud400_s7/bh_stats_theta{theta}"]
    got = [s.node_count, s.leaf_count, s.internal_node_count, s.accepted_approximations,
           s.exact_interactions, s.maximum_depth]
    assert got == list(st)


@pytest.mark.parametrize("name,eps", [("spiral400_s99", 0.05), ("cloud2048_s42", 0.01)])
def test_oracle_barnes_hut_other_inputs(oracle, golden, name, eps):
    p = fixture_particles(golden, name)
    t = oracle.nbo_bh_create()
    assert oracle.nbo_bh_compute(t, C.byref(p.view), C.byref(gravity(1.0, eps))) == 0
    s = BarnesHutStats()
    oracle.nbo_bh_stats(t, C.byref(s))
    oracle.nbo_bh_destroy(t)
    assert np.array_equal(p.acc(), golden[f"{name}/bh_acc_theta0.7"])
    st = golden[f"{name}/bh_stats_theta0.7"]
    assert [s.node_count, s.leaf_count, s.internal_node_count, s.accepted_approximations,
            s.exact_interactions, s.maximum_depth] == list(st)
    # barnes_hut_test.c:194-197 structure invariants
    n = p.n
    assert s.leaf_count == n and s.internal_node_count > 0 and s.maximum_depth >= 2
    if name == "spiral400_s99":
        assert s.node_count <= 2 * n


def test_oracle_kdk_bit_identical_to_reference_vectors(oracle, golden):
    p = fixture_particles(golden, "globular1000_s3")
    g = gravity(1.0, 0.01)
    oracle.nbo_direct_fp64(C.byref(p.view), C.byref(g))
    fn = C.cast(oracle.nbo_direct_fp64, C.c_void_p)
    for _ in range(3):
        assert oracle.nbo_leapfrog_advance(C.byref(p.view), C.byref(g), 1e-3, fn) == 0
    for k in ("px", "py", "pz", "vx", "vy", "vz", "m"):
        assert np.array_equal(p[k], golden[f"globular1000_s3/kdk3/{k}"]), k
    assert np.array_equal(p.acc(), golden["globular1000_s3/kdk3/acc"])
    q = Diagnostics()
    oracle.nbo_diagnostics_global(C.byref(p.view), C.byref(q))
    got = np.array([getattr(q, f) for f, _ in q._fields_])
    assert np.array_equal(got, golden["globular1000_s3/kdk3/diagnostics"])
    u = C.c_double()
    oracle.nbo_potential_energy(C.byref(p.view), C.byref(g), C.byref(u))
    assert u.value == golden["globular1000_s3/kdk3/potential"][0]


def test_two_body_known_answer(oracle):
    # gravity_test.c:30-60: eps = 0, a0 = (1,0,0), a1 = (-1,0,0)
    p = Particles(2, {"px": [-0.5, 0.5], "py": [0, 0], "pz": [0, 0], "m": [1.0, 1.0]})
    oracle.nbo_direct_fp64(C.byref(p.view), C.byref(gravity(1.0, 0.0)))
    assert np.allclose(p.acc(), [[1, 0, 0], [-1, 0, 0]], atol=1e-15, rtol=0)


def test_self_force_excluded(oracle):
    p = Particles(1, {"px": [1.0], "py": [1.0], "pz": [1.0], "m": [1000.0]})
    oracle.nbo_direct_fp64(C.byref(p.view), C.byref(gravity(1.0, 0.0)))
    assert np.all(p.acc() == 0.0)


def test_direct_matches_long_double_brute_force(oracle):
    rng = np.random.default_rng(11)
    n = 48
    pos = rng.normal(size=(n, 3))
    m = rng.uniform(0.1, 1.0, size=n)
    p = Particles(n, {"px": pos[:, 0], "py": pos[:, 1], "pz": pos[:, 2], "m": m})
    eps = 0.05
    oracle.nbo_direct_fp64(C.byref(p.view), C.byref(gravity(1.0, eps)))
    P = pos.astype(np.longdouble)
    d = P[None, :, :] - P[:, None, :]
    r2 = (d ** 2).sum(-1) + np.longdouble(eps) ** 2
    w = m.astype(np.longdouble)[None, :] / (r2 * np.sqrt(r2))
    np.fill_diagonal(w, 0)
    exact = (w[:, :, None] * d).sum(1)
    rel = np.linalg.norm(p.acc() - exact.astype(np.float64), axis=1) / np.linalg.norm(
        exact.astype(np.float64), axis=1)
    assert rel.max() < 1e-13


def test_diagnostics_closed_forms(oracle):
    # diagnostics_test.c:13-51, 53-80, 82-103
    p = Particles(2, {"px": [1, 4], "py": [2, 5], "pz": [3, 6], "vx": [1, 0], "vy": [0, 1],
                      "vz": [0, 0], "m": [2.0, 3.0]})
    q = Diagnostics()
    oracle.nbo_diagnostics_global(C.byref(p.view), C.byref(q))
    assert q.kinetic_energy == pytest.approx(2.5, abs=1e-15)
    assert (q.total_momentum_x, q.total_momentum_y, q.total_momentum_z) == (2.0, 3.0, 0.0)
    assert (q.total_angular_momentum_x, q.total_angular_momentum_y,
            q.total_angular_momentum_z) == (-18.0, 6.0, 8.0)
    assert q.center_of_mass_x == pytest.approx(14 / 5, abs=1e-15)
    assert q.center_of_mass_y == pytest.approx(19 / 5, abs=1e-15)
    assert q.center_of_mass_z == pytest.approx(24 / 5, abs=1e-15)
    pair = Particles(2, {"px": [-0.5, 0.5], "py": [0, 0], "pz": [0, 0], "m": [1.0, 1.0]})
    u = C.c_double()
    oracle.nbo_potential_energy(C.byref(pair.view), C.byref(gravity(1.0, 0.0)), C.byref(u))
    assert u.value == pytest.approx(-1.0, abs=1e-15)


def test_splitmix_xoshiro_published_outputs(oracle):
    # splitmix64 from state 0 (Vigna's reference implementation): first four outputs
    expected_state = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F,
                      0xF88BB8A8724C81EC]
    buf = (C.c_uint64 * 4)()
    oracle.nbo_rng_init(C.byref(buf), 0)
    assert list(buf) == expected_state
    s1 = expected_state[1]
    mask = (1 << 64) - 1
    rot = lambda v, k: ((v << k) | (v >> (64 - k))) & mask  # noqa: E731
    assert oracle.nbo_rng_next_u64(C.byref(buf)) == (rot((s1 * 5) & mask, 7) * 9) & mask


def _keys(oracle, p):
    t = oracle.nbo_bh_create()
    assert oracle.nbo_bh_build(t, C.byref(p.view)) == 0
    keys = np.zeros(p.n, dtype=np.uint64)
    perm = np.zeros(p.n, dtype=np.int64)
    oracle.nbo_bh_keys(t, keys.ctypes.data_as(C.POINTER(C.c_uint64)),
                       perm.ctypes.data_as(C.POINTER(C.c_int64)))
    root = np.zeros(4)
    oracle.nbo_bh_root(t, root.ctypes.data_as(C.POINTER(C.c_double)))
    s = BarnesHutStats()
    oracle.nbo_bh_stats(t, C.byref(s))
    rows = np.zeros((oracle.nbo_bh_node_count(t), 11))
    oracle.nbo_bh_export_nodes(t, rows.ctypes.data_as(C.POINTER(C.c_double)))
    oracle.nbo_bh_destroy(t)
    return keys, perm, root, s, rows


def test_morton_key_corners_and_centre(oracle):
    p = Particles(3, {"px": [0.0, 1.0, 0.5], "py": [0.0, 1.0, 0.5], "pz": [0.0, 1.0, 0.5],
                      "m": [1.0, 1.0, 1.0]})
    keys, perm, root, _, _ = _keys(oracle, p)
    assert keys[0] == 0
    assert keys[1] == (1 << 63) - 1          # all 63 bits set
    assert keys[2] == (7 << 60)              # centre: code 2^20 per axis -> bits 60..62
    assert list(perm) == [0, 2, 1]
    assert root[0] == 0.5 and root[3] == 0.5 * (1.0 + 1.0e-9)


def test_morton_key_is_interleaved_axis_quantisation(oracle):
    rng = np.random.default_rng(5)
    n = 3000
    pos = rng.uniform(-3, 2, size=(n, 3))
    p = Particles(n, {"px": pos[:, 0], "py": pos[:, 1], "pz": pos[:, 2], "m": np.ones(n)})
    keys, perm, root, _, _ = _keys(oracle, p)
    c, h = root[:3], root[3]
    u = (pos - c) / (2.0 * h) + 0.5
    u = np.where(u < 0, 0.0, np.where(u >= 1.0, 1.0 - 1e-15, u))
    code = (u * 2097152.0).astype(np.uint64)
    expect = np.zeros(n, dtype=np.uint64)
    for b in range(21):
        for a in range(3):
            expect |= ((code[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    assert np.array_equal(keys, expect)
    # the 30-bit key of a 10-bit-per-axis quantisation is key >> 33
    code10 = (u * 1024.0).astype(np.uint64)
    k30 = np.zeros(n, dtype=np.uint64)
    for b in range(10):
        for a in range(3):
            k30 |= ((code10[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    assert np.array_equal(keys >> np.uint64(33), k30)
    # stable sort on the full key (ties by original index)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))


def test_tree_invariants_and_theta_zero_limit(oracle, golden):
    p = fixture_particles(golden, "cloud400_s7")
    keys, perm, root, s, rows = _keys(oracle, p)
    assert s.leaf_count == p.n and rows.shape[0] == s.node_count
    top = rows[rows[:, 0] == 0][0]
    assert top[6] == pytest.approx(p["m"].sum(), rel=1e-14)
    com = (p.pos() * p["m"][:, None]).sum(0) / p["m"].sum()
    assert np.allclose(top[7:10], com, rtol=0, atol=1e-14)
    # theta -> 0: every interaction is an exact leaf term
    q = p.copy()
    t = oracle.nbo_bh_create()
    oracle.nbo_bh_set_theta(t, 1e-9)
    oracle.nbo_bh_compute(t, C.byref(q.view), C.byref(gravity(1.0, 0.02)))
    st = BarnesHutStats()
    oracle.nbo_bh_stats(t, C.byref(st))
    oracle.nbo_bh_destroy(t)
    assert st.accepted_approximations == 0 and st.exact_interactions == p.n * (p.n - 1)
    d = p.copy()
    oracle.nbo_direct_fp64(C.byref(d.view), C.byref(gravity(1.0, 0.02)))
    assert rms_rel_err(q.acc(), d.acc()) < 1e-14


def test_oracle_matches_reference_build_on_fresh_inputs(oracle, ref):
    """Cross-check against oracle/_ref on inputs the fixtures do not hold."""
    for preset, n, seed, eps, theta in [(PRESET_RANDOM_CLOUD, 3000, 123, 0.01, 0.5),
                                        (PRESET_SPIRAL_GALAXY, 2500, 5, 0.02, 0.9),
                                        (PRESET_GLOBULAR_CLUSTER, 1800, 77, 0.0, 0.7)]:
        a = ref_preset(ref, preset, n, seed)
        b = oracle_preset(oracle, preset, n, seed)
        for k in ("px", "py", "pz", "vx", "vy", "vz", "m"):
            assert np.array_equal(a[k], b[k])
        g = gravity(1.0, eps)
        err = Error()
        ref.n_body_sim_pro_gravity_compute_acceleration_reference(C.byref(a.view), C.byref(g),
                                                                  None, C.byref(err))
        oracle.nbo_direct_openmp(C.byref(b.view), C.byref(g))
        assert np.array_equal(a.acc(), b.acc())
        t = ref.n_body_sim_pro_barnes_hut_tree_create(C.byref(err))
        ref.n_body_sim_pro_barnes_hut_tree_set_theta(t, theta)
        ref.n_body_sim_pro_barnes_hut_compute_acceleration(C.byref(a.view), C.byref(g), t,
                                                           C.byref(err))
        s1 = BarnesHutStats()
        ref.n_body_sim_pro_barnes_hut_tree_stats(t, C.byref(s1))
        ref.n_body_sim_pro_barnes_hut_tree_destroy(t)
        u = oracle.nbo_bh_create()
        oracle.nbo_bh_set_theta(u, theta)
        oracle.nbo_bh_compute(u, C.byref(b.view), C.byref(g))
        s2 = BarnesHutStats()
        oracle.nbo_bh_stats(u, C.byref(s2))
        oracle.nbo_bh_destroy(u)
        assert np.array_equal(a.acc(), b.acc())
        assert (s1.node_count, s1.accepted_approximations, s1.exact_interactions,
                s1.maximum_depth) == (s2.node_count, s2.accepted_approximations,
                                      s2.exact_interactions, s2.maximum_depth)
