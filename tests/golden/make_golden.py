"""Writes the committed golden fixtures from the reference engine built from its
own sources into oracle/_ref (``make -C oracle ref``). Run in a container that
has the reference checkout; the fixtures are data only (inputs and outputs).

    python tests/golden/make_golden.py
"""
import ctypes as C
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

from nbx_testlib import (  # noqa: E402
    BarnesHutStats, Diagnostics, Error, gravity, load_ref, ref_preset,
    PRESET_RANDOM_CLOUD, PRESET_SPIRAL_GALAXY, PRESET_TWO_BODY, PRESET_GLOBULAR_CLUSTER,
)


def state(p):
    return {k: p[k].copy() for k in ("px", "py", "pz", "vx", "vy", "vz", "m")}


def bh(lib, p, g, theta):
    err = Error()
    t = lib.n_body_sim_pro_barnes_hut_tree_create(C.byref(err))
    lib.n_body_sim_pro_barnes_hut_tree_set_theta(t, theta)
    st = lib.n_body_sim_pro_barnes_hut_compute_acceleration(C.byref(p.view), C.byref(g), t,
                                                            C.byref(err))
    assert st == 0
    s = BarnesHutStats()
    lib.n_body_sim_pro_barnes_hut_tree_stats(t, C.byref(s))
    lib.n_body_sim_pro_barnes_hut_tree_destroy(t)
    return p.acc().copy(), np.array([s.node_count, s.leaf_count, s.internal_node_count,
                                     s.accepted_approximations, s.exact_interactions,
                                     s.maximum_depth], dtype=np.int64)


def main():
    lib = load_ref()
    out = {}
    err = Error()

    # (i) reference presets: random_cloud seed 7 (N=400, N=300), seed 42 (N=2048),
    # spiral galaxy seed 99 (N=400), globular cluster seed 3 (N=1000), two-body.
    for name, preset, n, seed in [("cloud400_s7", PRESET_RANDOM_CLOUD, 400, 7),
                                  ("cloud300_s7", PRESET_RANDOM_CLOUD, 300, 7),
                                  ("cloud2048_s42", PRESET_RANDOM_CLOUD, 2048, 42),
                                  ("spiral400_s99", PRESET_SPIRAL_GALAXY, 400, 99),
                                  ("globular1000_s3", PRESET_GLOBULAR_CLUSTER, 1000, 3),
                                  ("twobody", PRESET_TWO_BODY, 2, 1)]:
        p = ref_preset(lib, preset, n, seed)
        for k, v in state(p).items():
            out[f"{name}/{k}"] = v

    # (ii) reference all-pairs accelerations (scalar oracle and openmp_avx2).
    for name, eps in [("cloud2048_s42", 0.01), ("cloud300_s7", 0.02), ("globular1000_s3", 0.01)]:
        n = out[f"{name}/px"].size
        p = ref_preset(lib, {"cloud2048_s42": PRESET_RANDOM_CLOUD, "cloud300_s7":
                             PRESET_RANDOM_CLOUD, "globular1000_s3": PRESET_GLOBULAR_CLUSTER}[name],
                       n, {"cloud2048_s42": 42, "cloud300_s7": 7, "globular1000_s3": 3}[name])
        g = gravity(1.0, eps)
        lib.n_body_sim_pro_gravity_compute_acceleration_reference(C.byref(p.view), C.byref(g),
                                                                  None, C.byref(err))
        out[f"{name}/acc_reference_eps{eps}"] = p.acc().copy()
        lib.n_body_sim_pro_gravity_compute_acceleration_openmp_avx2(C.byref(p.view), C.byref(g),
                                                                    None, C.byref(err))
        out[f"{name}/acc_openmp_avx2_eps{eps}"] = p.acc().copy()

    # (iii) Barnes-Hut accelerations and statistics: theta sweep on the 400 cloud
    # (eps 0.02, as barnes_hut_test.c:116-117), spiral galaxy (eps 0.05, theta 0.7).
    for theta in (0.1, 0.3, 0.7, 1.0, 1.2):
        p = ref_preset(lib, PRESET_RANDOM_CLOUD, 400, 7)
        acc, stats = bh(lib, p, gravity(1.0, 0.02), theta)
        out[f"cloud400_s7/bh_acc_theta{theta}"] = acc
        out[f"cloud400_s7/bh_stats_theta{theta}"] = stats
    p = ref_preset(lib, PRESET_SPIRAL_GALAXY, 400, 99)
    acc, stats = bh(lib, p, gravity(1.0, 0.05), 0.7)
    out["spiral400_s99/bh_acc_theta0.7"] = acc
    out["spiral400_s99/bh_stats_theta0.7"] = stats
    p = ref_preset(lib, PRESET_RANDOM_CLOUD, 2048, 42)
    acc, stats = bh(lib, p, gravity(1.0, 0.01), 0.7)
    out["cloud2048_s42/bh_acc_theta0.7"] = acc
    out["cloud2048_s42/bh_stats_theta0.7"] = stats

    # (iv) leapfrog KDK with the reference scalar kernel: 3 steps, dt 1e-3, eps 0.01.
    p = ref_preset(lib, PRESET_GLOBULAR_CLUSTER, 1000, 3)
    g = gravity(1.0, 0.01)
    fn = C.cast(lib.n_body_sim_pro_gravity_compute_acceleration_reference, C.c_void_p)
    lib.n_body_sim_pro_gravity_compute_acceleration_reference(C.byref(p.view), C.byref(g), None,
                                                              C.byref(err))
    for _ in range(3):
        assert lib.n_body_sim_pro_integrator_advance(C.byref(p.view), C.byref(g), 1, 1e-3, fn,
                                                     None, C.byref(err)) == 0
    for k, v in state(p).items():
        out[f"globular1000_s3/kdk3/{k}"] = v
    out["globular1000_s3/kdk3/acc"] = p.acc().copy()
    q = Diagnostics()
    lib.n_body_sim_pro_diagnostics_compute_global(C.byref(p.view), C.byref(q), C.byref(err))
    u = C.c_double()
    lib.n_body_sim_pro_diagnostics_compute_potential_energy(C.byref(p.view), C.byref(g),
                                                            C.byref(u), C.byref(err))
    out["globular1000_s3/kdk3/diagnostics"] = np.array([getattr(q, f) for f, _ in q._fields_])
    out["globular1000_s3/kdk3/potential"] = np.array([u.value])

    np.savez_compressed(HERE / "reference_vectors.npz", **out)
    print("wrote", HERE / "reference_vectors.npz", len(out), "arrays")


if __name__ == "__main__":
    main()
