"""The drop-in force functions driven by the reference's own code, compiled
from its sources into oracle/_ref: its leapfrog integrator
(n_body_sim_pro_integrator_advance, src/c/physics/integrator.c:141-172) takes
nbx_hip_* as its NBodySimProForceFunction with an NBXEngine* (or NULL) as the
force context, and its force benchmark (benchmarks/benchmark_force.c with the
dispatch of INTEGRATION.md section 2 added) runs the hip_* names next to its
CPU names."""
import ctypes as C
import pathlib
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from nbx_testlib import (PRESET_RANDOM_CLOUD, Error, Particles, gravity, ref_preset,
                         rel_err_per_particle)

pytestmark = pytest.mark.gpu

LEAPFROG = 1
REF_BENCH = ROOT / "oracle" / "_ref" / "benchmark_force_nbx"


def golden_state(golden, name):
    n = golden[f"{name}/px"].size
    return Particles(n, {k: golden[f"{name}/{k}"] for k in ("px", "py", "pz", "vx", "vy", "vz",
                                                              "m")})


def drop_in(nbx, name):
    fn = getattr(nbx.load(), name)
    return fn, C.cast(fn, C.c_void_p)


def run_reference_integrator(ref, p, g, fn_ptr, fn_call, ctx, steps, dt=1e-3):
    """Initial accelerations through the force function, then `steps` calls of
    the reference's integrator with that function and context."""
    err = Error()
    assert fn_call(C.byref(p.view), C.byref(g), ctx, C.byref(err)) == 0, err.message
    for _ in range(steps):
        st = ref.n_body_sim_pro_integrator_advance(C.byref(p.view), C.byref(g), LEAPFROG, dt,
                                                   fn_ptr, ctx, C.byref(err))
        assert st == 0, err.message
    return p


@pytest.mark.parametrize("with_engine", [True, False])
def test_reference_integrator_with_hip_direct_fp64_is_bit_identical(nbx, ref, golden,
                                                                     with_engine):
    """3 leapfrog steps (dt=1e-3, eps=0.01) of the 1000-particle globular
    cluster through the reference's integrator with nbx_hip_direct_fp64 give
    the committed reference vectors (the same integrator with the reference's
    scalar kernel) bit for bit."""
    p = golden_state(golden, "globular1000_s3")
    fn, ptr = drop_in(nbx, "nbx_hip_direct_fp64")
    eng = nbx.Engine(p.n) if with_engine else None
    try:
        run_reference_integrator(ref, p, gravity(1.0, 0.01), ptr, fn,
                                 C.c_void_p(eng.h) if eng else None, 3)
    finally:
        if eng is not None:
            eng.close()
    for k in ("px", "py", "pz", "vx", "vy", "vz"):
        assert np.array_equal(p[k], golden[f"globular1000_s3/kdk3/{k}"]), k
    assert np.array_equal(p.acc(), golden["globular1000_s3/kdk3/acc"])


def test_reference_integrator_with_hip_barnes_hut_fp64_is_bit_identical(nbx, ref):
    """3 leapfrog steps of a 20000-particle random cloud (the reference's
    generator, seed 42), theta=0.7, eps=0.01: the reference's integrator with
    nbx_hip_barnes_hut_fp64 (engine context) and with the reference's own
    Barnes-Hut (tree context) agree bit for bit."""
    base = ref_preset(ref, PRESET_RANDOM_CLOUD, 20000, 42)
    g = gravity(1.0, 0.01)
    mine = base.copy()
    fn, ptr = drop_in(nbx, "nbx_hip_barnes_hut_fp64")
    with nbx.Engine(mine.n, theta=0.7) as e:
        run_reference_integrator(ref, mine, g, ptr, fn, C.c_void_p(e.h), 3)
    theirs = base.copy()
    err = Error()
    tree = ref.n_body_sim_pro_barnes_hut_tree_create(C.byref(err))
    ref.n_body_sim_pro_barnes_hut_tree_set_theta(tree, 0.7)
    try:
        rfn = ref.n_body_sim_pro_barnes_hut_compute_acceleration
        run_reference_integrator(ref, theirs, g, C.cast(rfn, C.c_void_p), rfn,
                                 C.c_void_p(tree), 3)
    finally:
        ref.n_body_sim_pro_barnes_hut_tree_destroy(tree)
    for k in ("px", "py", "pz", "vx", "vy", "vz", "ax", "ay", "az"):
        assert np.array_equal(mine[k], theirs[k]), k


def test_reference_integrator_with_hip_direct_fp32(nbx, ref):
    """The FP32 drop-in through the reference's integrator: after 3 steps its
    accelerations stay within the FP32 all-pairs bar (median <= 1e-6) of the
    same integrator with the reference's scalar kernel."""
    base = ref_preset(ref, PRESET_RANDOM_CLOUD, 8192, 7)
    g = gravity(1.0, 0.01)
    mine = base.copy()
    fn, ptr = drop_in(nbx, "nbx_hip_direct_fp32")
    with nbx.Engine(mine.n) as e:
        run_reference_integrator(ref, mine, g, ptr, fn, C.c_void_p(e.h), 3)
    theirs = base.copy()
    rfn = ref.n_body_sim_pro_gravity_compute_acceleration_reference
    run_reference_integrator(ref, theirs, g, C.cast(rfn, C.c_void_p), rfn, None, 3)
    err = rel_err_per_particle(mine.acc(), theirs.acc())
    print(f"fp32 drop-in through the reference integrator: median {np.median(err):.3e} "
          f"max {err.max():.3e}")
    assert np.median(err) <= 1e-6
    assert np.abs(mine.pos() - theirs.pos()).max() <= 1e-9


def _ref_bench(*args):
    if not REF_BENCH.exists():
        pytest.skip("oracle/_ref/benchmark_force_nbx was not built (needs the reference sources)")
    r = subprocess.run([str(REF_BENCH), *args], capture_output=True, text=True, timeout=120)
    return r


@pytest.mark.parametrize("algorithm", ["hip_direct_fp64", "hip_direct_fp32", "hip_barnes_hut",
                                       "hip_barnes_hut_fp64", "hip_bh_learned", "openmp_avx2",
                                       "barnes_hut"])
def test_reference_benchmark_harness_runs_hip_and_cpu_names(algorithm):
    """The reference's benchmark_force.c, with the libnbx.so dispatch added,
    accepts the hip_* names beside its own CPU names and reports its CSV."""
    r = _ref_bench("--particles", "20000", "--steps", "3", "--threads", "4",
                   "--algorithm", algorithm)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0] == f"particles=20000 steps=3 algorithm={algorithm}"
    assert lines[-2] == "threads,ms,speedup,efficiency"
    threads, ms, *_ = lines[-1].split(",")
    assert float(ms) > 0.0
