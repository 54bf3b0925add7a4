"""C-ABI checks that need no GPU: libnbx.so loads, exports every function the
headers in include/ declare, and its host-only entry points (names, error
paths before any device work, initial-condition generators) behave."""
import ctypes as C
import pathlib
import re

import numpy as np
import pytest

from conftest import ROOT
from nbx_testlib import Particles

INCLUDE = ROOT / "include"
OK, INVALID_ARGUMENT = 0, 1   # NBodySimProStatus (src/c/core/status.h order)
DECL = re.compile(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(nbx_\w+)\s*\(", re.M)


def declared_functions():
    names = set()
    for h in sorted(INCLUDE.glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in DECL.finditer(text):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    import n_body_sim_pro_amd as nbx
    return nbx.load()


def test_headers_declare_the_drop_ins():
    names = declared_functions()
    for fn in ("nbx_hip_direct_fp64", "nbx_hip_direct_fp32", "nbx_hip_barnes_hut",
               "nbx_hip_barnes_hut_fp64", "nbx_hip_bh_learned", "nbx_engine_create",
               "nbx_ensemble_create", "nbx_mlp_create"):
        assert fn in names
    assert len(names) >= 50


def test_library_exports_every_declared_function(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared in include/ but not exported: {missing}"


def test_algorithm_names_round_trip(lib):
    lib.nbx_algorithm_name.restype = C.c_char_p
    names = ["hip_direct_fp64", "hip_direct_fp32", "hip_barnes_hut", "hip_barnes_hut_fp64",
             "hip_bh_learned"]
    for k, name in enumerate(names):
        assert lib.nbx_algorithm_from_name(name.encode()) == k
        assert lib.nbx_algorithm_name(k) == name.encode()
    assert lib.nbx_algorithm_from_name(b"reference") == -1
    assert lib.nbx_algorithm_from_name(None) == -1


def test_mlp_parameter_count_matches_layout(lib):
    import sys
    sys.path.insert(0, str(ROOT / "oracle"))
    import mlp_oracle as mo
    lib.nbx_mlp_param_count.restype = C.c_size_t
    assert lib.nbx_mlp_param_count() == mo.N_PARAMS == 6288


def test_drop_in_argument_errors_before_device_work(lib):
    import n_body_sim_pro_amd as nbx
    err = nbx.Error()
    g = nbx.gravity(1.0, 0.01)
    fn = nbx.force_function("hip_direct_fp64")
    assert fn(None, C.byref(g), None, C.byref(err)) == INVALID_ARGUMENT
    assert b"must not be null" in err.message
    empty = nbx.make_view(0)
    assert fn(C.byref(empty), C.byref(g), None, C.byref(err)) == OK   # nothing to do
    for name in ("hip_barnes_hut", "hip_barnes_hut_fp64", "hip_bh_learned"):
        err = nbx.Error()
        st = nbx.force_function(name)(C.byref(empty), C.byref(g), None, C.byref(err))
        assert st == INVALID_ARGUMENT and b"empty system" in err.message


def test_engine_and_ensemble_reject_bad_configs(lib):
    import n_body_sim_pro_amd as nbx
    err = nbx.Error()
    cfg = nbx.EngineConfig(capacity=0, device=0, theta=0.7, flags=0)
    assert not lib.nbx_engine_create(C.byref(cfg), C.byref(err))
    assert err.status == INVALID_ARGUMENT
    with pytest.raises(nbx.NBXError):
        nbx.Ensemble(np.array([], dtype=np.uint32))


def test_random_cloud_generator_matches_reference_fixture(golden):
    import n_body_sim_pro_amd as nbx
    for key, n, seed in (("cloud400_s7", 400, 7), ("cloud2048_s42", 2048, 42)):
        d = nbx.random_cloud(n, seed)
        for c, f in enumerate(("px", "py", "pz")):
            assert np.array_equal(d["pos"][:, c], golden[f"{key}/{f}"])
        for c, f in enumerate(("vx", "vy", "vz")):
            assert np.array_equal(d["vel"][:, c], golden[f"{key}/{f}"])
        assert np.array_equal(d["m"], golden[f"{key}/m"])


def test_random_cloud_generator_matches_oracle(oracle):
    import n_body_sim_pro_amd as nbx
    from nbx_testlib import oracle_preset
    p = oracle_preset(oracle, 1, 5000, 123)
    d = nbx.random_cloud(5000, 123)
    assert np.array_equal(d["pos"], p.pos())
    assert np.array_equal(d["m"], p["m"])


def test_plummer_invariants():
    import n_body_sim_pro_amd as nbx
    n = 200_000
    d = nbx.plummer(n, seed=3)
    m, x, v = d["m"], d["pos"], d["vel"]
    assert np.allclose(m, 1.0 / n) and m.sum() == pytest.approx(1.0, rel=1e-12)
    assert np.abs((m[:, None] * x).sum(0)).max() < 1e-12
    assert np.abs((m[:, None] * v).sum(0)).max() < 1e-12
    r = np.linalg.norm(x, axis=1)
    # half-mass radius of a Plummer sphere: a / sqrt(2^(2/3) - 1) = 1.3048 a
    assert np.median(r) == pytest.approx(1.3048, rel=0.02)
    # virial equilibrium: T = -E = 3 pi G M^2 / (64 a) = 0.1473; isotropic velocities
    T = 0.5 * (m * (v * v).sum(1)).sum()
    assert T == pytest.approx(3 * np.pi / 64, rel=0.02)
    assert np.allclose((v * v).mean(0) / (v * v).sum(1).mean(), 1 / 3, atol=0.01)
    # deterministic per seed, different across seeds
    assert np.array_equal(nbx.plummer(1000, seed=3)["pos"], nbx.plummer(1000, seed=3)["pos"])
    assert not np.array_equal(nbx.plummer(1000, seed=3)["pos"], nbx.plummer(1000, seed=4)["pos"])


def test_two_galaxy_invariants():
    import n_body_sim_pro_amd as nbx
    n = 100_000
    d = nbx.two_galaxy(n, seed=42)
    m, x, v = d["m"], d["pos"], d["vel"]
    assert m.sum() == pytest.approx(1.0, rel=1e-12)
    assert np.abs((m[:, None] * x).sum(0)).max() < 1e-9
    assert np.abs((m[:, None] * v).sum(0)).max() < 1e-9
    half = n // 2
    c1 = x[:half].mean(0)
    c2 = x[half:].mean(0)
    assert np.linalg.norm(c2 - c1) == pytest.approx(20.0, rel=0.02)
    # parabolic relative orbit: relative speed^2 = 2 G M / d at separation d
    v1, v2 = v[:half].mean(0), v[half:].mean(0)
    vrel = np.linalg.norm(v2 - v1)
    assert vrel == pytest.approx(np.sqrt(2 * 1.0 / 20.0), rel=0.05)


def _disk_model_checks(x, v, m, nb, scale_length=1.0, scale_height=0.1, bulge_scale=0.2,
                       bulge_fraction=0.1):
    """One galaxy in its own frame, disk in the xy plane, bulge particles
    first: exponential surface density (R ~ Gamma(2, Rd): mean 2 Rd, a
    fraction 1 - 2/e inside Rd), sech^2 vertical profile (median |z| =
    z0 atanh(1/2)), Plummer bulge (half-mass radius 1.3048 b) holding
    `bulge_fraction` of the mass, disk rotating at the circular speed of the
    disk + bulge mass model."""
    M = m.sum()
    assert m[:nb].sum() / M == pytest.approx(bulge_fraction, rel=1e-3)
    disk_x, disk_v = x[nb:], v[nb:]
    R = np.hypot(disk_x[:, 0], disk_x[:, 1])
    z = disk_x[:, 2]
    assert R.mean() == pytest.approx(2.0 * scale_length, rel=0.02)
    assert (R < scale_length).mean() == pytest.approx(1.0 - 2.0 / np.e, abs=0.01)
    assert np.median(np.abs(z)) == pytest.approx(scale_height * np.arctanh(0.5), rel=0.03)
    rb = np.linalg.norm(x[:nb], axis=1)
    assert np.median(rb) == pytest.approx(1.3048 * bulge_scale, rel=0.03)
    r = np.hypot(R, z)
    xr = r / scale_length
    menc = (M * (1 - bulge_fraction) * (1.0 - (1.0 + xr) * np.exp(-xr))
            + M * bulge_fraction * r ** 3 / (r * r + bulge_scale ** 2) ** 1.5)
    vphi = (disk_x[:, 0] * disk_v[:, 1] - disk_x[:, 1] * disk_v[:, 0]) / R
    assert np.median(vphi / np.sqrt(menc / r)) == pytest.approx(1.0, abs=0.03)


def test_disk_galaxy_structure():
    import n_body_sim_pro_amd as nbx
    n = 200_000
    d = nbx.disk_galaxy(n, seed=11)
    m, x, v = d["m"], d["pos"], d["vel"]
    assert m.sum() == pytest.approx(1.0, rel=1e-12)
    assert np.abs((m[:, None] * x).sum(0)).max() < 1e-9
    _disk_model_checks(x, v, m, int(round(0.1 * n)))


def test_two_galaxy_members_follow_the_disk_model():
    """Each galaxy of the pair, taken in its own centre-of-mass frame: the
    first keeps its disk in the xy plane, the second is tilted by 45 degrees
    about x (its spin along (0, -sin 45, cos 45)); untilted, both follow the
    disk + bulge model."""
    import n_body_sim_pro_amd as nbx
    n = 400_000
    d = nbx.two_galaxy(n, seed=42)
    m, x, v = d["m"], d["pos"], d["vel"]
    half = n // 2
    c, s = np.cos(np.pi / 4), np.sin(np.pi / 4)
    untilt = np.array([[1, 0, 0], [0, c, s], [0, -s, c]])   # inverse of the tilt about x
    for k, sl in enumerate((slice(0, half), slice(half, n))):
        gm, gx, gv = m[sl], x[sl], v[sl]
        gx = gx - (gm[:, None] * gx).sum(0) / gm.sum()
        gv = gv - (gm[:, None] * gv).sum(0) / gm.sum()
        spin = (gm[:, None] * np.cross(gx, gv)).sum(0)
        spin /= np.linalg.norm(spin)
        want = np.array([0.0, 0.0, 1.0]) if k == 0 else np.array([0.0, -s, c])
        assert np.abs(spin - want).max() < 0.02
        if k == 1:
            gx, gv = gx @ untilt.T, gv @ untilt.T
        _disk_model_checks(gx, gv, gm, int(round(0.1 * half)))


def test_ensemble_sizes_are_log_uniform_in_range():
    import n_body_sim_pro_amd as nbx
    s = nbx.ensemble_sizes(4096, lo=64, hi=4096, seed=42)
    assert s.min() >= 64 and s.max() <= 4096
    # log-uniform: each octave [64,128), [128,256) ... holds about 1/6 of the clusters
    hist, _ = np.histogram(np.log2(s), bins=6, range=(6, 12))
    assert np.all(np.abs(hist / len(s) - 1 / 6) < 0.03)
    assert np.array_equal(s, nbx.ensemble_sizes(4096, lo=64, hi=4096, seed=42))


def test_oracle_loads_without_product(oracle):
    # the checker and the product are separate libraries
    assert hasattr(oracle, "nbo_direct_fp64")
    p = Particles(2, {"px": np.array([0.0, 1.0]), "m": np.array([1.0, 1.0])})
    assert p.view.particle_count == 2
