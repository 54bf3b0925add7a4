"""Leapfrog KDK and FP64 diagnostics on the GPU against the reference (vectors
and oracle), plus the reference's two-body orbit regression."""
import ctypes as C
import math

import numpy as np
import pytest

from nbx_testlib import Diagnostics, Particles, gravity

pytestmark = pytest.mark.gpu


def fixture_particles(golden, name):
    n = golden[f"{name}/px"].size
    return Particles(n, {k: golden[f"{name}/{k}"] for k in ("px", "py", "pz", "vx", "vy", "vz", "m")})


def test_kdk_three_steps_bit_identical_to_reference_vectors(nbx, golden):
    """integrator.c:47-85 with the scalar kernel, 3 steps, dt=1e-3, eps=0.01."""
    p = fixture_particles(golden, "globular1000_s3")
    with nbx.Engine(p.n) as e:
        e.upload_view(p.view)
        e.compute_force("hip_direct_fp64", eps=0.01)
        e.advance("hip_direct_fp64", dt=1e-3, steps=3, eps=0.01)
        out = e.download()
        d = e.diagnostics(eps=0.01, potential=True)
    for k, col in (("px", 0), ("py", 1), ("pz", 2)):
        assert np.array_equal(out["pos"][:, col], golden[f"globular1000_s3/kdk3/{k}"])
    for k, col in (("vx", 0), ("vy", 1), ("vz", 2)):
        assert np.array_equal(out["vel"][:, col], golden[f"globular1000_s3/kdk3/{k}"])
    assert np.array_equal(out["acc"], golden["globular1000_s3/kdk3/acc"])
    ref = golden["globular1000_s3/kdk3/diagnostics"]
    names = [f for f, _ in Diagnostics._fields_]
    for k, name in enumerate(names):
        # FP64 tree reduction vs the reference's sequential sum: 1e-12 relative
        scale = max(abs(ref[k]), 1e-12 if "momentum" in name or "center" in name else 1e-300)
        assert abs(d[name] - ref[k]) <= 1e-12 * scale + 1e-15, name
    u = golden["globular1000_s3/kdk3/potential"][0]
    assert abs(d["potential_energy"] - u) <= 1e-12 * abs(u)


def test_diagnostics_closed_forms(nbx):
    p = Particles(2, {"px": [1, 4], "py": [2, 5], "pz": [3, 6], "vx": [1, 0], "vy": [0, 1],
                      "vz": [0, 0], "m": [2.0, 3.0]})
    with nbx.Engine(2) as e:
        e.upload_view(p.view)
        d = e.diagnostics(eps=0.0, potential=False)
    assert d["kinetic_energy"] == 2.5
    assert (d["total_momentum_x"], d["total_momentum_y"], d["total_momentum_z"]) == (2, 3, 0)
    assert (d["total_angular_momentum_x"], d["total_angular_momentum_y"],
            d["total_angular_momentum_z"]) == (-18, 6, 8)
    assert d["center_of_mass_x"] == pytest.approx(14 / 5, abs=1e-15)
    pair = Particles(2, {"px": [-0.5, 0.5], "py": [0, 0], "pz": [0, 0], "m": [1.0, 1.0]})
    with nbx.Engine(2) as e:
        e.upload_view(pair.view)
        assert e.diagnostics(eps=0.0)["potential_energy"] == pytest.approx(-1.0, abs=1e-15)


@pytest.mark.parametrize("integrator", [1, 2])
def test_two_body_orbit_regression(nbx, integrator):
    """two_body_test.c:57-184: 8 orbits x 500 steps, eps=0."""
    s = math.sqrt(2.0) / 2.0
    p = Particles(2, {"px": [-0.5, 0.5], "py": [0, 0], "pz": [0, 0], "vx": [0, 0],
                      "vy": [s, -s], "vz": [0, 0], "m": [1.0, 1.0]})
    period = 2.0 * math.pi / math.sqrt(2.0)
    dt = period / 500
    with nbx.Engine(2) as e:
        e.upload_view(p.view)
        e.compute_force("hip_direct_fp64", eps=0.0)
        d0 = e.diagnostics(eps=0.0)
        e.advance("hip_direct_fp64", dt=dt, steps=8 * 500, eps=0.0, integrator=integrator)
        d1 = e.diagnostics(eps=0.0)
        out = e.download()
    e0 = d0["kinetic_energy"] + d0["potential_energy"]
    e1 = d1["kinetic_energy"] + d1["potential_energy"]
    assert abs((e1 - e0) / e0) < 1e-3
    sep = np.linalg.norm(out["pos"][1] - out["pos"][0])
    assert abs(sep - 1.0) < 1e-2
    assert abs(d1["total_momentum_x"]) < 1e-12 and abs(d1["total_momentum_y"]) < 1e-12
    assert abs(d1["center_of_mass_x"]) < 1e-12 and abs(d1["center_of_mass_y"]) < 1e-12


def test_invalid_timestep_rejected(nbx):
    p = Particles(2, {"px": [-0.5, 0.5], "py": [0, 0], "pz": [0, 0], "m": [1.0, 1.0]})
    with nbx.Engine(2) as e:
        e.upload_view(p.view)
        with pytest.raises(nbx.NBXError) as ex:
            e.advance("hip_direct_fp64", dt=0.0, steps=1)
        assert ex.value.status == 1


def test_barnes_hut_kdk_conserves_momentum(nbx):
    d = nbx.two_galaxy(100_000, seed=1)
    p = Particles(len(d["m"]), {"px": d["pos"][:, 0], "py": d["pos"][:, 1], "pz": d["pos"][:, 2],
                                "vx": d["vel"][:, 0], "vy": d["vel"][:, 1], "vz": d["vel"][:, 2],
                                "m": d["m"]})
    with nbx.Engine(p.n) as e:
        e.upload_view(p.view)
        e.compute_force("hip_barnes_hut", eps=0.01)
        q0 = e.diagnostics(potential=False)
        e.advance("hip_barnes_hut", dt=1e-3, steps=3, eps=0.01)
        q1 = e.diagnostics(potential=False)
    for c in "xyz":
        assert abs(q1[f"total_momentum_{c}"] - q0[f"total_momentum_{c}"]) < 1e-5
