"""The C17 force benchmark (n_body_sim_pro_amd/host/nbx_benchmark.c), the
counterpart of the reference's benchmarks/benchmark_force.c with the hip_*
algorithms: argument handling on the CPU, every algorithm on the GPU."""
import subprocess

import pytest

from conftest import ROOT

EXE = ROOT / "n_body_sim_pro_amd" / "nbx_benchmark"
ALGORITHMS = ["hip_direct_fp64", "hip_direct_fp32", "hip_barnes_hut", "hip_barnes_hut_fp64",
              "hip_bh_learned"]


def run(*args, timeout=120):
    assert EXE.exists(), f"{EXE} is missing: run python -m n_body_sim_pro_amd.build"
    return subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


def test_usage_and_argument_errors():
    r = run("--help")
    assert r.returncode == 0 and "--algorithm" in r.stderr and "hip_barnes_hut" in r.stderr
    assert run("--particles").returncode == 1                     # incomplete
    assert run("--bogus", "1").returncode == 1
    r = run("--algorithm", "openmp_avx2")                         # a CPU name: not ours
    assert r.returncode == 1 and "unknown algorithm" in r.stderr
    assert run("--particles", "0").returncode == 1


def test_without_device_reports_and_fails():
    import n_body_sim_pro_amd as nbx
    if nbx.device_count() > 0:
        pytest.skip("a HIP device is visible")
    r = run("--particles", "100", "--steps", "1")
    assert r.returncode == 1 and "no HIP device" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("algorithm", ALGORITHMS)
def test_every_algorithm_runs_both_modes(algorithm):
    r = run("--particles", 4096, "--steps", 3, "--warmup", 1, "--algorithm", algorithm,
            "--preset", "plummer")
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    header = lines.index("mode,ms_first,ms_mean,ms_median,kernel_ms,ginteractions_s,"
                         "kernel_tflops,frac_peak")
    rows = [ln.split(",") for ln in lines[header + 1:]]
    assert [row[0] for row in rows] == ["resident", "dropin"]
    for row in rows:
        vals = list(map(float, row[1:]))
        assert all(v >= 0 for v in vals) and vals[2] > 0 and vals[4] > 0
