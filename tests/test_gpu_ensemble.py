"""Batched ensemble on the GPU: per-cluster accelerations against the FP64
oracle, graph replays against eager steps, FP64 per-cluster energies against
the oracle, energy drift of a 50-step leapfrog against the oracle's FP64
trajectory, and the autotuner."""
import ctypes as C

import numpy as np
import pytest

from nbx_testlib import Particles, gravity, rel_err_per_particle

pytestmark = pytest.mark.gpu


def cluster_particles(d, off, n):
    sl = slice(off, off + n)
    return Particles(n, {"px": d["pos"][sl, 0], "py": d["pos"][sl, 1], "pz": d["pos"][sl, 2],
                         "vx": d["vel"][sl, 0], "vy": d["vel"][sl, 1], "vz": d["vel"][sl, 2],
                         "m": d["m"][sl]})


@pytest.fixture(scope="module")
def small_ensemble(nbx):
    sizes = nbx.ensemble_sizes(48, lo=64, hi=700, seed=5)
    d = nbx.ensemble_clusters(sizes, seed=9)
    return sizes, d


def test_ensemble_force_against_fp64_oracle(nbx, oracle, small_ensemble):
    sizes, d = small_ensemble
    with nbx.Ensemble(sizes) as ens:
        ens.upload(d["pos"], d["vel"], d["m"])
        ens.force(eps=0.01)
        acc = ens.download()["acc"]
    errs = []
    off = 0
    for n in sizes:
        p = cluster_particles(d, off, int(n))
        oracle.nbo_direct_openmp(C.byref(p.view), C.byref(gravity(1.0, 0.01)))
        errs.append(rel_err_per_particle(acc[off:off + n], p.acc()))
        off += int(n)
    err = np.concatenate(errs)
    print(f"ensemble fp32 median {np.median(err):.3e} max {err.max():.3e}")
    assert np.median(err) <= 1e-6
    assert err.max() < 1e-4


def test_ensemble_tile_shapes_give_identical_forces(nbx, small_ensemble):
    sizes, d = small_ensemble
    with nbx.Ensemble(sizes) as ens:
        ens.upload(d["pos"], d["vel"], d["m"])
        outs = []
        for r, bpc in ((1, 1), (2, 4), (4, 8)):
            ens.set_tile([r] * 8, bpc, 256)
            ens.force(eps=0.01)
            outs.append(ens.download()["acc"])
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_graph_replays_match_eager_steps(nbx, small_ensemble):
    """Two 10-step graph replays, with eager work and syncs between them, must
    reproduce 20 eager steps bit for bit after each replay."""
    sizes, d = small_ensemble
    with nbx.Ensemble(sizes) as a, nbx.Ensemble(sizes) as b:
        for e in (a, b):
            e.upload(d["pos"], d["vel"], d["m"])
            e.force(eps=0.01)
        for _ in range(2):
            a.advance(1e-3, 10, graph_steps=10, eps=0.01)
            b.advance(1e-3, 10, graph_steps=0, eps=0.01)
            sa, sb = a.download(), b.download()
            assert np.array_equal(sa["pos"], sb["pos"]) and np.array_equal(sa["vel"], sb["vel"])
            a.energies(eps=0.01)   # eager work between replays


def test_cluster_energies_against_oracle(nbx, oracle, small_ensemble):
    sizes, d = small_ensemble
    with nbx.Ensemble(sizes) as ens:
        ens.upload(d["pos"], d["vel"], d["m"])
        en = ens.energies(eps=0.01)
    off = 0
    for k, n in enumerate(sizes[:12]):
        p = cluster_particles(d, off, int(n))
        from nbx_testlib import Diagnostics
        q = Diagnostics()
        oracle.nbo_diagnostics_global(C.byref(p.view), C.byref(q))
        u = C.c_double()
        oracle.nbo_potential_energy(C.byref(p.view), C.byref(gravity(1.0, 0.01)), C.byref(u))
        assert abs(en[k, 0] - q.kinetic_energy) <= 1e-12 * abs(q.kinetic_energy)
        assert abs(en[k, 1] - u.value) <= 1e-12 * abs(u.value)
        off += int(n)


def test_energy_drift_against_fp64_oracle_trajectory(nbx, oracle, small_ensemble):
    """50 graph-captured KDK steps (dt=1e-3, eps=0.01) of FP32 forces vs the
    oracle's FP64 leapfrog on the same clusters: per-cluster energies agree."""
    sizes, d = small_ensemble
    with nbx.Ensemble(sizes) as ens:
        ens.upload(d["pos"], d["vel"], d["m"])
        ens.force(eps=0.01)
        e0 = ens.energies(eps=0.01).sum(1)
        ens.advance(1e-3, 50, graph_steps=50, eps=0.01)
        e1 = ens.energies(eps=0.01).sum(1)
    g = gravity(1.0, 0.01)
    fn = C.cast(oracle.nbo_direct_fp64, C.c_void_p)
    off = 0
    for k, n in enumerate(sizes[:6]):
        p = cluster_particles(d, off, int(n))
        oracle.nbo_direct_fp64(C.byref(p.view), C.byref(g))
        for _ in range(50):
            oracle.nbo_leapfrog_advance(C.byref(p.view), C.byref(g), 1e-3, fn)
        from nbx_testlib import Diagnostics
        q = Diagnostics()
        oracle.nbo_diagnostics_global(C.byref(p.view), C.byref(q))
        u = C.c_double()
        oracle.nbo_potential_energy(C.byref(p.view), C.byref(g), C.byref(u))
        eo = q.kinetic_energy + u.value
        drift_gpu = (e1[k] - e0[k]) / abs(e0[k])
        drift_cpu = (eo - e0[k]) / abs(e0[k])
        print(f"cluster {k} n={n} drift gpu {drift_gpu:.3e} fp64 oracle {drift_cpu:.3e}")
        assert abs(drift_gpu - drift_cpu) < 1e-6
        off += int(n)


def test_autotune_returns_valid_cached_config(nbx, small_ensemble, tmp_path):
    sizes, d = small_ensemble
    cache = tmp_path / "tune.txt"
    with nbx.Ensemble(sizes) as ens:
        ens.upload(d["pos"], d["vel"], d["m"])
        ens.force(eps=0.01)
        ref = ens.download()["acc"]
        rep = ens.autotune(eps=0.01, cache_path=str(cache))
        assert rep["cached"] is False and rep["blocks_per_cu"] in (1, 2, 4, 8)
        ens.force(eps=0.01)
        acc = ens.download()["acc"]
        rep2 = ens.autotune(eps=0.01, cache_path=str(cache))
        assert rep2["cached"] is True
        assert rep2["targets_per_lane"] == rep["targets_per_lane"]
    err = rel_err_per_particle(acc, ref)
    assert np.median(err) < 1e-6
