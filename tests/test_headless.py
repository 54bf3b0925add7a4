"""The command-line runner (n_body_sim_pro_amd/host/nbx_headless.c): the
reference's `hardware`, `benchmark`, `resume` and `save` subcommands
(src/main.cpp:50-380) on the device-resident engine, with the headless
report of src/cpp/benchmark/HeadlessRunner.cpp:106-164.

CPU: argument handling and exit statuses. GPU: the JSON report, `save` files
against the named presets (pinned bit for bit against the reference's
generator on the CPU) and the FP64 oracle's accelerations, and `resume`
against the same steps taken through the library."""
import ctypes as C
import json
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from nbx_testlib import Particles, gravity

EXE = ROOT / "n_body_sim_pro_amd" / "nbx_headless"
REFERENCE_KEYS = ("cpu", "simd", "threads", "particles", "algorithm", "theta", "steps",
                  "estimated_bytes", "avg_step_ms", "first_step_ms", "energy_drift",
                  "momentum_error")


def run(*args, timeout=120):
    assert EXE.exists(), f"{EXE} is missing: run python -m n_body_sim_pro_amd.build"
    cmd = [str(EXE), *map(str, args)]
    print("run:", " ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)


def report(r):
    lines = r.stdout.strip().splitlines()
    k = lines.index("--- machine-readable ---")
    return json.loads(lines[k + 1])


def test_usage_and_argument_errors():
    r = run("--help")
    assert r.returncode == 0 and "benchmark" in r.stdout and "resume FILE" in r.stdout
    assert run().returncode == 1                                     # no subcommand
    for args, message in [(("benchmark", "--bogus", "1"), "unknown argument --bogus"),
                          (("benchmark", "--particles"), "missing value for --particles"),
                          (("benchmark", "--particles", "0"), "invalid value 0"),
                          (("benchmark", "--particles", "12x"), "invalid value 12x"),
                          (("benchmark", "--algorithm", "openmp_avx2"), "unknown algorithm"),
                          (("benchmark", "--algorithm", "hip_bh_learned"), "trained model"),
                          (("benchmark", "--preset", "nebula"), "unknown preset nebula"),
                          (("benchmark", "--dt", "-1"), "invalid value -1"),
                          (("resume",), "missing checkpoint path"),
                          (("resume", "x.hpcs", "--theta", "0.5"), "unknown argument --theta"),
                          (("save",), "missing checkpoint path"),
                          (("save", "x.hpcs", "--steps", "3"), "unknown argument --steps")]:
        r = run(*args)
        assert r.returncode == 1, args
        assert message in r.stderr, (args, r.stderr)


def test_resume_rejects_bad_files(tmp_path):
    bad = tmp_path / "bad.hpcs"
    bad.write_bytes(b"not a checkpoint at all, just some bytes")
    r = run("resume", bad)
    assert r.returncode == 1 and "failed to load checkpoint" in r.stderr
    r = run("resume", tmp_path / "missing.hpcs")
    assert r.returncode == 1 and "failed to load checkpoint" in r.stderr


def test_hardware_and_no_device():
    import n_body_sim_pro_amd as nbx
    r = run("hardware")
    assert r.returncode == 0 and "HIP devices" in r.stdout
    if nbx.device_count() > 0:
        return
    r = run("benchmark", "--particles", 100, "--steps", 1)
    assert r.returncode == 1 and "no HIP device" in r.stderr


@pytest.mark.gpu
def test_benchmark_all_pairs_report():
    r = run("benchmark", "--particles", 4096, "--steps", 3, "--preset", "globular_cluster",
            "--algorithm", "all_pairs")
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stderr
    j = report(r)
    for k in REFERENCE_KEYS:
        assert k in j, k
    assert j["particles"] == 4096 and j["steps"] == 3 and j["algorithm"] == "all-pairs"
    assert j["engine_algorithm"] == "hip_direct_fp64" and j["preset"] == "globular_cluster"
    assert j["avg_step_ms"] > 0 and j["first_step_ms"] > 0 and j["estimated_bytes"] > 0
    # 4 KDK steps of dt 1e-3 on an unsoftened Plummer cluster
    assert j["energy_drift"] is not None and j["energy_drift"] < 1e-3
    assert j["momentum_error"] < 1e-10


@pytest.mark.gpu
def test_benchmark_barnes_hut_default_above_20000():
    """No --algorithm: Barnes-Hut above 20000 particles (main.cpp:157-159);
    200000 galaxy_collision particles use the per-particle generator."""
    r = run("benchmark", "--particles", 200000, "--steps", 2)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stderr
    j = report(r)
    assert j["algorithm"] == "barnes_hut" and j["engine_algorithm"] == "hip_barnes_hut"
    assert j["preset"] == "galaxy_collision" and j["energy_drift"] is None
    assert j["tree_build_ms"] > 0 and j["force_evaluation_ms"] > 0
    assert j["momentum_error"] < 1e-4


@pytest.mark.gpu
def test_save_matches_reference_controller(tmp_path, oracle):
    """`save` = the controller's defaults (SimulationController.hpp:57-75) on
    the sequential preset, with all-pairs FP64 initial accelerations."""
    from n_body_sim_pro_amd import checkpoint, preset
    path = tmp_path / "spiral.hpcs"
    r = run("save", path, "--particles", 3000, "--preset", "spiral_galaxy", "--seed", 9)
    assert r.returncode == 0, r.stderr
    h, a = checkpoint.read(path)
    assert (h.particle_count, h.simulation_time, h.timestep, h.integrator, h.theta,
            h.barnes_hut_enabled, h.random_seed, h.preset) == (3000, 0.0, 0.005, 1, 0.7, 0, 9, 5)
    want = preset("spiral_galaxy", 3000, 9)
    assert np.array_equal(np.stack([a["px"], a["py"], a["pz"]], 1), want["pos"])
    assert np.array_equal(np.stack([a["vx"], a["vy"], a["vz"]], 1), want["vel"])
    assert np.array_equal(a["m"], want["m"])
    q = Particles(3000, {"px": a["px"], "py": a["py"], "pz": a["pz"], "m": a["m"]})
    oracle.nbo_direct_fp64(C.byref(q.view), C.byref(gravity(1.0, 0.0)))
    assert np.array_equal(np.stack([a["ax"], a["ay"], a["az"]], 1), q.acc())


@pytest.mark.gpu
def test_resume_continues_like_the_library(tmp_path, nbx):
    """`resume FILE --steps 4 --save OUT` = load + 1 + 4 steps with the
    checkpoint's dt, integrator and force (Barnes-Hut off -> all-pairs)."""
    from n_body_sim_pro_amd import checkpoint
    start, out = tmp_path / "start.hpcs", tmp_path / "out.hpcs"
    r = run("save", start, "--particles", 2048, "--preset", "open_cluster", "--seed", 3)
    assert r.returncode == 0, r.stderr
    r = run("resume", start, "--steps", 4, "--save", out)
    print(r.stdout[-1200:])
    assert r.returncode == 0, r.stderr
    assert "Resumed 4 steps from" in r.stdout
    j = report(r)
    assert j["resumed"] is True and j["engine_algorithm"] == "hip_direct_fp64"
    h, got = checkpoint.read(out)
    assert h.simulation_time == pytest.approx(5 * 0.005, rel=1e-15)
    assert (h.timestep, h.integrator, h.barnes_hut_enabled, h.random_seed, h.preset) == (
        0.005, 1, 0, 3, 3)
    with nbx.Engine(2048) as e:
        checkpoint.load_engine(e, start)
        e.advance("hip_direct_fp64", dt=0.005, steps=1, eps=0.0)
        e.advance("hip_direct_fp64", dt=0.005, steps=4, eps=0.0)
        want = e.download()
    for k, cols in (("pos", ("px", "py", "pz")), ("vel", ("vx", "vy", "vz")),
                    ("acc", ("ax", "ay", "az"))):
        assert np.array_equal(np.stack([got[c] for c in cols], 1), want[k]), k


@pytest.mark.gpu
def test_resume_reference_written_checkpoint_barnes_hut(tmp_path, nbx):
    """A checkpoint with the Barnes-Hut flag set resumes on hip_barnes_hut;
    the file is written by this library's writer, which is byte-identical
    to the reference's (tests/test_checkpoint.py)."""
    from n_body_sim_pro_amd import checkpoint, preset
    n = 30000
    p = preset("globular_cluster", n, 5)
    arrays = {"px": p["pos"][:, 0], "py": p["pos"][:, 1], "pz": p["pos"][:, 2],
              "vx": p["vel"][:, 0], "vy": p["vel"][:, 1], "vz": p["vel"][:, 2],
              "ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n), "m": p["m"]}
    with nbx.Engine(n) as e:
        e.upload(p["pos"], p["m"], vel=p["vel"])
        e.compute_force("hip_barnes_hut", eps=0.0)
        acc = e.accelerations()
    arrays.update(ax=acc[:, 0], ay=acc[:, 1], az=acc[:, 2])
    path = tmp_path / "bh.hpcs"
    checkpoint.write(path, arrays, checkpoint.header(n, simulation_time=2.0, timestep=2e-3,
                                                     integrator=1, theta=0.6, barnes_hut=True,
                                                     seed=5, preset=4))
    r = run("resume", path, "--steps", 3)
    print(r.stdout[-1200:])
    assert r.returncode == 0, r.stderr
    j = report(r)
    assert j["algorithm"] == "barnes_hut" and j["theta"] == 0.6 and j["particles"] == n
    assert j["tree_build_ms"] > 0 and j["momentum_error"] < 1e-4
