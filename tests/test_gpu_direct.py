"""All-pairs parity on the GPU: hip_direct_fp64 must be bit-identical to the
reference's scalar kernel (via the oracle and the committed reference vectors);
hip_direct_fp32 must meet median relative error <= 1e-6 against it."""
import ctypes as C

import numpy as np
import pytest

from nbx_testlib import Error, Particles, gravity, rel_err_per_particle

pytestmark = pytest.mark.gpu


def fixture_particles(golden, name):
    n = golden[f"{name}/px"].size
    return Particles(n, {k: golden[f"{name}/{k}"] for k in ("px", "py", "pz", "vx", "vy", "vz", "m")})


def gpu_force(nbx, p, algorithm, eps, G=1.0):
    with nbx.Engine(max(p.n, 1)) as e:
        e.upload_view(p.view)
        e.compute_force(algorithm, G=G, eps=eps)
        return e.accelerations()


def from_gen(d):
    return Particles(len(d["m"]), {"px": d["pos"][:, 0], "py": d["pos"][:, 1],
                                   "pz": d["pos"][:, 2], "vx": d["vel"][:, 0],
                                   "vy": d["vel"][:, 1], "vz": d["vel"][:, 2], "m": d["m"]})


@pytest.mark.parametrize("name,eps", [("cloud2048_s42", 0.01), ("cloud300_s7", 0.02),
                                      ("globular1000_s3", 0.01)])
def test_direct_fp64_bit_identical_to_reference_vectors(nbx, golden, name, eps):
    p = fixture_particles(golden, name)
    acc = gpu_force(nbx, p, "hip_direct_fp64", eps)
    assert np.array_equal(acc, golden[f"{name}/acc_reference_eps{eps}"])


def test_direct_fp64_plummer_16384_bit_identical(nbx, oracle):
    """Config 1 size: N=16384 Plummer (DF velocities), eps=0.01."""
    p = from_gen(nbx.plummer(16384, seed=42))
    acc = gpu_force(nbx, p, "hip_direct_fp64", 0.01)
    q = p.copy()
    oracle.nbo_direct_openmp(C.byref(q.view), C.byref(gravity(1.0, 0.01)))
    assert np.array_equal(acc, q.acc())
    assert rel_err_per_particle(acc, q.acc()).max() <= 1e-12


def test_direct_fp64_plummer_16384_three_kdk_steps_bit_identical(nbx, oracle):
    """Config 1: N=16384 Plummer, eps=0.01, leapfrog dt=1e-3, 3 steps (4 force
    evaluations incl. the initial one) against the oracle's KDK with the
    reference's scalar force: positions, velocities and accelerations agree
    bit for bit (so within the 1e-12 relative bound)."""
    p = from_gen(nbx.plummer(16384, seed=42))
    g = gravity(1.0, 0.01)
    with nbx.Engine(16384) as e:
        e.upload_view(p.view)
        e.compute_force("hip_direct_fp64", eps=0.01)
        e.advance("hip_direct_fp64", dt=1e-3, steps=3, eps=0.01)
        s = e.download()
    q = p.copy()
    fn = C.cast(oracle.nbo_direct_openmp, C.c_void_p)
    oracle.nbo_direct_openmp(C.byref(q.view), C.byref(g))
    for _ in range(3):
        assert oracle.nbo_leapfrog_advance(C.byref(q.view), C.byref(g), 1e-3, fn) == 0
    assert np.array_equal(s["pos"], q.pos())
    assert np.array_equal(s["vel"], np.stack([q["vx"], q["vy"], q["vz"]], 1))
    assert np.array_equal(s["acc"], q.acc())
    assert rel_err_per_particle(s["acc"], q.acc()).max() <= 1e-12


def test_direct_fp32_energy_drift_100_steps(nbx):
    """Config 3 shape at N=16384: 100 KDK steps (dt=1e-3, eps=0.01) with FP32
    pair forces and the FP64 integrator; the relative energy drift matches the
    drift of the same run with FP64 forces (hip_direct_fp64, bit-identical to
    the reference kernel) to 1e-6, and stays below 1e-4."""
    d = nbx.plummer(16384, seed=3)
    drift = {}
    for algo in ("hip_direct_fp32", "hip_direct_fp64"):
        with nbx.Engine(16384) as e:
            e.upload(d["pos"], d["m"], d["vel"])
            e.compute_force(algo, eps=0.01)
            q0 = e.diagnostics(eps=0.01)
            e.advance(algo, dt=1e-3, steps=100, eps=0.01)
            q1 = e.diagnostics(eps=0.01)
        e0 = q0["kinetic_energy"] + q0["potential_energy"]
        e1 = q1["kinetic_energy"] + q1["potential_energy"]
        drift[algo] = (e1 - e0) / abs(e0)
    print("relative energy drift over 100 steps", drift)
    assert abs(drift["hip_direct_fp32"] - drift["hip_direct_fp64"]) < 1e-6
    assert abs(drift["hip_direct_fp32"]) < 1e-4


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 257])
def test_direct_fp64_ragged_sizes_and_zero_softening(nbx, oracle, n):
    rng = np.random.default_rng(n)
    pos = rng.normal(size=(n, 3))
    p = Particles(n, {"px": pos[:, 0], "py": pos[:, 1], "pz": pos[:, 2],
                      "m": rng.uniform(0.5, 1.5, n)})
    for eps in (0.0, 0.01):
        acc = gpu_force(nbx, p, "hip_direct_fp64", eps)
        q = p.copy()
        oracle.nbo_direct_fp64(C.byref(q.view), C.byref(gravity(1.0, eps)))
        assert np.array_equal(acc, q.acc())
        assert np.all(np.isfinite(acc))


def test_two_body_known_answer_and_self_force(nbx):
    p = Particles(2, {"px": [-0.5, 0.5], "py": [0, 0], "pz": [0, 0], "m": [1.0, 1.0]})
    acc = gpu_force(nbx, p, "hip_direct_fp64", 0.0)
    assert np.allclose(acc, [[1, 0, 0], [-1, 0, 0]], atol=1e-15, rtol=0)
    one = Particles(1, {"px": [1.0], "py": [1.0], "pz": [1.0], "m": [1000.0]})
    for algo in ("hip_direct_fp64", "hip_direct_fp32"):
        assert np.all(gpu_force(nbx, one, algo, 0.0) == 0.0)


def test_dropin_force_function_overwrites_accelerations(nbx, golden):
    """The NBodySimProForceFunction drop-in with context NULL (library engine)."""
    p = fixture_particles(golden, "cloud2048_s42")
    p.a["ax"][:] = 123.0
    err = Error()
    fn = nbx.force_function("hip_direct_fp64")
    assert fn(C.byref(p.view), C.byref(gravity(1.0, 0.01)), None, C.byref(err)) == 0
    assert np.array_equal(p.acc(), golden["cloud2048_s42/acc_reference_eps0.01"])
    bad = fn(None, C.byref(gravity(1.0, 0.01)), None, C.byref(err))
    assert bad == 1 and b"must not be null" in err.message


def test_direct_fp32_error_against_fp64_oracle(nbx, oracle):
    """Median relative error <= 1e-6 (the FP32 target) on a 16384 Plummer sphere."""
    p = from_gen(nbx.plummer(16384, seed=7))
    acc = gpu_force(nbx, p, "hip_direct_fp32", 0.01)
    q = p.copy()
    oracle.nbo_direct_openmp(C.byref(q.view), C.byref(gravity(1.0, 0.01)))
    err = rel_err_per_particle(acc, q.acc())
    print(f"fp32 median {np.median(err):.3e} max {err.max():.3e}")
    assert np.median(err) <= 1e-6
    assert err.max() < 1e-4


def test_direct_fp32_subsample_at_65536(nbx, oracle):
    n = 65536
    p = from_gen(nbx.plummer(n, seed=11))
    acc = gpu_force(nbx, p, "hip_direct_fp32", 0.01)
    idx = np.random.default_rng(7).choice(n, 1024, replace=False).astype(np.int64)
    ref = np.zeros((idx.size, 3))
    oracle.nbo_direct_openmp_subset(C.byref(p.view), C.byref(gravity(1.0, 0.01)),
                                    idx.ctypes.data_as(C.POINTER(C.c_int64)), idx.size,
                                    ref.ctypes.data_as(C.POINTER(C.c_double)))
    err = rel_err_per_particle(acc[idx], ref)
    print(f"fp32@65536 median {np.median(err):.3e} max {err.max():.3e}")
    assert np.median(err) <= 1e-6


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 255, 256, 257, 300, 1023, 1025, 2049])
def test_direct_fp32_ragged_sizes(nbx, oracle, n):
    """Sizes around the wave (64), the compensation tile (256 sources) and the
    split of each target's sources into four ranges of whole tiles (1025:
    ranges of 2, 2, 1 and 0 tiles), with and without softening and with
    G != 1 (G multiplies each target's sum once)."""
    rng = np.random.default_rng(1000 + n)
    pos = rng.normal(size=(n, 3))
    p = Particles(n, {"px": pos[:, 0], "py": pos[:, 1], "pz": pos[:, 2],
                      "m": rng.uniform(0.5, 1.5, n)})
    for G, eps in ((1.0, 0.0), (1.0, 0.01), (2.5, 0.01)):
        acc = gpu_force(nbx, p, "hip_direct_fp32", eps, G=G)
        q = p.copy()
        oracle.nbo_direct_fp64(C.byref(q.view), C.byref(gravity(G, eps)))
        ref = q.acc()
        assert np.all(np.isfinite(acc))
        if n == 1:
            assert np.all(acc == 0.0)
            continue
        err = rel_err_per_particle(acc, ref)
        assert np.median(err) <= 1e-6, (G, eps, np.median(err))
        assert err.max() < 1e-4, (G, eps, err.max())


def test_direct_fp32_zero_softening_is_finite(nbx):
    d = nbx.random_cloud(1000, seed=3)
    p = from_gen(d)
    acc = gpu_force(nbx, p, "hip_direct_fp32", 0.0)
    assert np.all(np.isfinite(acc))
