"""Headline benchmark: Barnes-Hut force evaluation at N=1,000,000 plus all-pairs
Ginteractions/s, each at its stated acceleration error, on one MI355X.

Workload (a "step"): one leapfrog kick-drift-kick step of a 1,000,000-particle
two-galaxy collision (each galaxy an exponential disk, scale length 1, height
0.1, 90% of its mass, plus a Plummer bulge, 10%; 20 units apart on a parabolic
approach), FP32 positions and masses, theta = 0.7, eps = 0.01, dt = 1e-3,
G = 1, seed 42. Each step rebuilds the octree on the device (bounding cube,
Morton keys, radix sort, nodes, centre of mass) and walks it. The state stays
resident in HBM; nothing crosses PCIe inside the timed region except the
per-evaluation statistics.

value = mean BH force-evaluation time (ms, HIP events around build + walk),
with the walk's median/p99 relative error against the FP64 direct sum on a
seeded 4096-particle sample stated beside it.
The second headline number, all-pairs Ginteractions/s, comes from the FP32
tiled kernel on an N=262,144 Plummer sphere (DF velocities), with its median
and max relative acceleration error against the FP64 direct sum.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]
"""
import argparse
import ctypes as C
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from n_body_sim_pro_amd.abi import BarnesHutStats, Error  # noqa: E402

METRIC = "BH force-eval ms @1M + all-pairs Ginteractions/s, at stated accel error"
FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 vector/matrix peak
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--theta", type=float, default=0.7)
    ap.add_argument("--eps", type=float, default=0.01)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--ap-n", type=int, default=262_144)
    ap.add_argument("--ap-evals", type=int, default=5)
    ap.add_argument("--ap-drift-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-direct-fp64", action="store_true")
    ap.add_argument("--dp-n", type=int, default=16_384)
    ap.add_argument("--no-all-pairs", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ens-clusters", type=int, default=4096)
    ap.add_argument("--ens-steps", type=int, default=50)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-n", type=int, default=10_000_000)
    ap.add_argument("--learned-evals", type=int, default=3)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain elementwise copy of this many MiB (3 launches): "
                         "the known-size kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE")
    return ap.parse_args()


class CpuBaseline:
    """The reference's CPU force paths timed on this host's cores, in the same
    run as the GPU: the reference compiled from its own sources into
    oracle/_ref (kind "reference"), or the oracle's C restatement (kind
    "port", scalar/OpenMP/Barnes-Hut only) when that library was not built.
    Used as the timed baseline and never as the measured path.

    Timing follows benchmark_force.c:84-147 (CLOCK_MONOTONIC wall time; one
    Barnes-Hut tree created per measurement and reused for its `evals`
    evaluations; mean, no warm-up), and, where `median` is asked, one
    warm-up evaluation then the median of three."""

    ALGOS = {
        "reference": "n_body_sim_pro_gravity_compute_acceleration_reference",
        "openmp": "n_body_sim_pro_gravity_compute_acceleration_openmp",
        "openmp_avx2": "n_body_sim_pro_gravity_compute_acceleration_openmp_avx2",
        "openmp_avx512": "n_body_sim_pro_gravity_compute_acceleration_openmp_avx512",
        "barnes_hut": "n_body_sim_pro_barnes_hut_compute_acceleration",
        "barnes_hut_openmp_avx2": "n_body_sim_pro_barnes_hut_compute_acceleration_openmp_avx2",
        "barnes_hut_openmp_avx512":
            "n_body_sim_pro_barnes_hut_compute_acceleration_openmp_avx512",
    }

    def __init__(self):
        self.threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ref_so = ROOT / "oracle" / "_ref" / "libnbsp_ref.so"
        flags = ""
        try:
            flags = next(line for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines()
                         if line.startswith("flags"))
        except (OSError, StopIteration):
            pass
        self.avx512 = " avx512f" in flags
        self.avx2 = " avx2" in flags and " fma" in flags
        if ref_so.exists():
            self.kind = "reference"
            self.lib = C.CDLL(str(ref_so))
            for name in self.ALGOS.values():
                getattr(self.lib, name).argtypes = [C.c_void_p] * 4
                getattr(self.lib, name).restype = C.c_int
            self.lib.n_body_sim_pro_threading_set_thread_count.argtypes = [C.c_int]
            self.lib.n_body_sim_pro_barnes_hut_tree_create.restype = C.c_void_p
            self.lib.n_body_sim_pro_barnes_hut_tree_create.argtypes = [C.c_void_p]
            self.lib.n_body_sim_pro_barnes_hut_tree_set_theta.argtypes = [C.c_void_p, C.c_double]
            self.lib.n_body_sim_pro_barnes_hut_tree_destroy.argtypes = [C.c_void_p]
            self.lib.n_body_sim_pro_barnes_hut_tree_stats.argtypes = [C.c_void_p, C.c_void_p]
            self.lib.n_body_sim_pro_cpu_brand_string.restype = C.c_char_p
        else:
            self.kind = "port"
            sys.path.insert(0, str(ROOT / "tests"))
            from nbx_testlib import load_oracle
            self.lib = load_oracle()

    def available(self, algo):
        if self.kind == "port":
            return algo in ("reference", "openmp", "barnes_hut")
        if algo.endswith("avx512"):
            return self.avx512
        if algo.endswith("avx2"):
            return self.avx2
        return True

    def environment(self):
        brand = None
        if self.kind == "reference":
            brand = self.lib.n_body_sim_pro_cpu_brand_string().decode(errors="replace").strip()
        if not brand:
            try:
                brand = next(line.split(":", 1)[1].strip() for line in
                             pathlib.Path("/proc/cpuinfo").read_text().splitlines()
                             if line.startswith("model name"))
            except (OSError, StopIteration):
                brand = "unknown"
        omp = None
        try:
            gomp = C.CDLL("libgomp.so.1")
            omp = int(gomp.omp_get_max_threads())
        except OSError:
            pass
        return {"cpu_brand": brand, "nproc": os.cpu_count(),
                "affinity_cpus": len(os.sched_getaffinity(0)),
                "omp_get_max_threads": omp, "threads_used": self.threads,
                "avx2": self.avx2, "avx512f": self.avx512, "kind": self.kind}

    def _set_threads(self, threads):
        if self.kind == "reference":
            self.lib.n_body_sim_pro_threading_set_thread_count(int(threads))

    def _eval(self, algo, view, g, tree):
        if self.kind == "reference":
            err = Error()
            st = getattr(self.lib, self.ALGOS[algo])(C.byref(view), C.byref(g), tree, C.byref(err))
            if st != 0:
                raise RuntimeError(f"reference {algo} failed: {err.message}")
        elif algo == "reference":
            self.lib.nbo_direct_fp64(C.byref(view), C.byref(g))
        elif algo == "openmp":
            self.lib.nbo_direct_openmp(C.byref(view), C.byref(g))
        else:
            self.lib.nbo_bh_compute(tree, C.byref(view), C.byref(g))

    def _tree(self, theta):
        if theta is None:
            return None
        if self.kind == "reference":
            err = Error()
            t = self.lib.n_body_sim_pro_barnes_hut_tree_create(C.byref(err))
            self.lib.n_body_sim_pro_barnes_hut_tree_set_theta(t, theta)
            return t
        t = self.lib.nbo_bh_create()
        self.lib.nbo_bh_set_theta(t, theta)
        return t

    def _stats(self, tree):
        s = BarnesHutStats()
        if self.kind == "reference":
            self.lib.n_body_sim_pro_barnes_hut_tree_stats(tree, C.byref(s))
        else:
            self.lib.nbo_bh_stats(tree, C.byref(s))
        return s

    def _destroy(self, tree):
        if tree is None:
            return
        if self.kind == "reference":
            self.lib.n_body_sim_pro_barnes_hut_tree_destroy(tree)
        else:
            self.lib.nbo_bh_destroy(tree)

    def time(self, algo, view, g, evals=1, theta=None, median=False, threads=None):
        """ms per evaluation: the reference harness's no-warm-up mean over
        `evals`, and with `median` also warm-up + median of three."""
        threads = self.threads if threads is None else threads
        self._set_threads(threads)
        out = {"algorithm": algo, "threads": threads if "openmp" in algo or
               algo.startswith("barnes_hut") else 1}
        tree = self._tree(theta)
        try:
            t0 = time.perf_counter()
            for _ in range(evals):
                self._eval(algo, view, g, tree)
            out["no_warmup_mean_ms"] = 1e3 * (time.perf_counter() - t0) / evals
            out["no_warmup_evals"] = evals
            if median:
                self._eval(algo, view, g, tree)
                ts = []
                for _ in range(3):
                    t0 = time.perf_counter()
                    self._eval(algo, view, g, tree)
                    ts.append(time.perf_counter() - t0)
                out["median_ms"] = 1e3 * float(np.median(ts))
            if theta is not None:
                s = self._stats(tree)
                out["theta"] = theta
                out["interactions"] = int(s.accepted_approximations + s.exact_interactions)
                out["build_ms_last"] = 1e3 * s.build_time_seconds
                out["walk_ms_last"] = 1e3 * s.evaluation_time_seconds
        finally:
            self._destroy(tree)
        out["ms"] = out.get("median_ms", out["no_warmup_mean_ms"])
        return out


def host_view(pos, m):
    """A View over fresh float64 columns (positions, masses, zeroed accelerations)."""
    from n_body_sim_pro_amd.engine import make_view
    n = len(m)
    cols = {"px": np.ascontiguousarray(pos[:, 0], dtype=np.float64),
            "py": np.ascontiguousarray(pos[:, 1], dtype=np.float64),
            "pz": np.ascontiguousarray(pos[:, 2], dtype=np.float64),
            "m": np.ascontiguousarray(m, dtype=np.float64),
            "ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n)}
    return make_view(n, **cols), cols


def host_acc(cols):
    return np.stack([cols["ax"], cols["ay"], cols["az"]], 1)


def rel_err(a, ref):
    den = np.linalg.norm(ref, axis=1)
    return np.linalg.norm(a - ref, axis=1) / np.where(den > 0, den, 1.0)


def err_summary(a, ref):
    e = rel_err(a, ref)
    return {"median_rel_err": float(np.median(e)), "max_rel_err": float(e.max())}


def direct_fp64(nbx, cpu, n, eps, seed, dt, steps=3, evals=5):
    """Config 1: FP64 all-pairs on an N=16384 Plummer sphere, bit-identical to
    the reference's scalar kernel; a 3-step leapfrog (4 evaluations with the
    initial one). CPU baseline: the reference's scalar, OpenMP and
    OpenMP+AVX2 (AVX-512) kernels on the same input."""
    d = nbx.plummer(n, seed=seed)
    with nbx.Engine(n) as e:
        e.upload(d["pos"], d["m"], d["vel"])
        e.compute_force("hip_direct_fp64", eps=eps)   # warm-up, and the initial forces
        acc = e.accelerations()
        ks = []
        for _ in range(evals):
            e.compute_force("hip_direct_fp64", eps=eps)
            ks.append(e.timings()["kernel_ms"])
        e.compute_force("hip_direct_fp64", eps=eps)
        e.advance("hip_direct_fp64", dt=dt, steps=steps, eps=eps)
        kdk_ms = e.timings()["integrate_ms"]
    ms = float(np.mean(ks))
    pairs = float(n) * float(n - 1)
    out = {"n": n, "ms_per_eval": ms, "gpairs_s": pairs / (ms * 1e-3) / 1e9,
           "tflops_20flop": 20.0 * pairs / (ms * 1e-3) / 1e12,
           "frac_fp64_peak_20flop": 20.0 * pairs / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
           "kdk_steps": steps, "kdk_ms_total": kdk_ms, "dtype": "f64"}
    if cpu is not None:
        view, cols = host_view(d["pos"], d["m"])
        g = nbx.gravity(1.0, eps)
        runs = [cpu.time("reference", view, g, evals=1, threads=1)]
        ref_acc = host_acc(cols).copy()
        for algo in ("openmp", "openmp_avx2", "openmp_avx512"):
            if cpu.available(algo):
                runs.append(cpu.time(algo, view, g, evals=3, median=True))
        out["bit_identical_to_cpu_reference"] = bool(np.array_equal(acc, ref_acc))
        best = min(runs, key=lambda r: r["ms"])
        out["cpu_baseline"] = {"value": best["ms"], "unit": "ms per FP64 force eval",
                               "cores": best["threads"], "kind": cpu.kind,
                               "best": best["algorithm"],
                               "sample": f"full N={n} evaluation", "runs": runs,
                               "gpu_speedup_vs_best": best["ms"] / ms}
    return out


def all_pairs(nbx, cpu, n, evals, eps, seed, drift_steps, dt):
    d = nbx.plummer(n, seed=seed)
    with nbx.Engine(n) as e:
        e.upload(d["pos"], d["m"], d["vel"])
        e.compute_force("hip_direct_fp64", eps=eps)
        ref = e.accelerations()
        e.compute_force("hip_direct_fp32", eps=eps)   # warm-up
        ks = []
        for _ in range(evals):
            e.compute_force("hip_direct_fp32", eps=eps)
            ks.append(e.timings())
        acc = e.accelerations()
        drift = None
        if drift_steps > 0:
            # FP32 pair forces, FP64 KDK state and FP64 energies (O(N^2) potential)
            q0 = e.diagnostics(eps=eps)
            e.advance("hip_direct_fp32", dt=dt, steps=drift_steps, eps=eps)
            q1 = e.diagnostics(eps=eps)
            e0 = q0["kinetic_energy"] + q0["potential_energy"]
            e1 = q1["kinetic_energy"] + q1["potential_energy"]
            p0 = np.array([q0[f"total_momentum_{c}"] for c in "xyz"])
            p1 = np.array([q1[f"total_momentum_{c}"] for c in "xyz"])
            drift = {"steps": drift_steps, "dt": dt, "relative_energy_drift": (e1 - e0) / abs(e0),
                     "momentum_change": float(np.linalg.norm(p1 - p0))}
    idx = np.random.default_rng(7).choice(n, min(4096, n), replace=False)
    rel = rel_err(acc[idx], ref[idx])
    kern = float(np.mean([t["kernel_ms"] for t in ks]))
    force = float(np.mean([t["force_ms"] for t in ks]))
    pairs = float(n) * float(n - 1)
    out = {"n": n, "gint_s": pairs / (force * 1e-3) / 1e9, "ms_per_eval": force,
           "kernel_ms": kern, "tflops_20flop": 20.0 * pairs / (kern * 1e-3) / 1e12,
           "frac_fp32_peak": 20.0 * pairs / (kern * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
           "median_rel_err": float(np.median(rel)), "max_rel_err": float(rel.max()),
           "error_sample": f"{idx.size} seeded targets vs the FP64 direct sum "
                           "(hip_direct_fp64, bit-identical to the reference scalar kernel)",
           "dtype": "f32 pairs, compensated f32 sums", "energy_drift": drift}
    if cpu is not None:
        view, cols = host_view(d["pos"], d["m"])
        g = nbx.gravity(1.0, eps)
        runs = []
        for algo in ("openmp_avx2", "openmp_avx512") if cpu.kind == "reference" else ("openmp",):
            if cpu.available(algo):
                r = cpu.time(algo, view, g, evals=1)
                r.update(err_summary(host_acc(cols)[idx], ref[idx]))
                r["gint_s"] = pairs / (r["ms"] * 1e-3) / 1e9
                runs.append(r)
        if runs:
            best = min(runs, key=lambda r: r["ms"])
            out["cpu_baseline"] = {"value": best["gint_s"], "unit": "Ginteractions/s",
                                   "ms_per_eval": best["ms"], "cores": best["threads"],
                                   "kind": cpu.kind, "best": best["algorithm"],
                                   "sample": f"one full N={n} evaluation per kernel "
                                             "(the reference harness with --steps 1)",
                                   "runs": runs, "gpu_speedup_vs_best": best["ms"] / force}
    return out


def ensemble(nbx, cpu, clusters, steps, eps, dt, seed):
    """4096 Plummer clusters, sizes log-uniform in [64, 4096]: autotuned tile
    shapes, one persistent launch per force evaluation, a hipGraph of `steps`
    leapfrog steps; per-cluster FP64 energy drift over those steps."""
    sizes = nbx.ensemble_sizes(clusters, lo=64, hi=4096, seed=seed)
    d = nbx.ensemble_clusters(sizes, seed=seed)
    with nbx.Ensemble(sizes) as ens:
        ens.upload(d["pos"], d["vel"], d["m"])
        ens.force(eps=eps)
        tune = ens.autotune(eps=eps)
        ens.upload(d["pos"], d["vel"], d["m"])
        ens.force(eps=eps)
        acc0 = ens.download()["acc"]
        e0 = ens.energies(eps=eps).sum(1)
        ms = ens.advance(dt, steps, graph_steps=steps, eps=eps)     # builds + replays the graph
        ms2 = ens.advance(dt, steps, graph_steps=steps, eps=eps)    # replay only
        e1 = ens.energies(eps=eps).sum(1)
    n = sizes.astype(np.float64)
    pairs = float(np.sum(n * (n - 1)))
    drift = np.abs((e1 - e0) / e0)
    out = {"clusters": int(clusters), "particles": int(n.sum()), "steps": steps,
           "ms_per_step": ms2 / steps, "first_replay_ms_per_step": ms / steps,
           "pairs_per_step": pairs, "gint_s": pairs / (ms2 / steps * 1e-3) / 1e9,
           "frac_fp32_peak_20flop": 20.0 * pairs / (ms2 / steps * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
           "energy_drift_median": float(np.median(drift)), "energy_drift_max": float(drift.max()),
           "drift_interval_steps": 2 * steps, "autotune": tune}
    if cpu is not None:
        # one force evaluation of every cluster through the reference's best
        # all-pairs kernel (each cluster is its own particle system)
        algo = next(a for a in ("openmp_avx512", "openmp_avx2", "openmp") if cpu.available(a))
        g = nbx.gravity(1.0, eps)
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        views = [host_view(d["pos"][o:o + k], d["m"][o:o + k])
                 for o, k in zip(offs, sizes.astype(np.int64))]
        cpu._set_threads(cpu.threads)
        t0 = time.perf_counter()
        for view, _ in views:
            cpu._eval(algo, view, g, None)
        cpu_ms = 1e3 * (time.perf_counter() - t0)
        cpu_acc = np.concatenate([host_acc(c) for _, c in views])
        r = err_summary(acc0, cpu_acc)
        out["cpu_baseline"] = {"value": pairs / (cpu_ms * 1e-3) / 1e9, "unit": "Ginteractions/s",
                               "ms_per_force_eval": cpu_ms, "cores": cpu.threads,
                               "kind": cpu.kind, "algorithm": algo,
                               "sample": f"one force evaluation of all {clusters} clusters, "
                                         "one reference call per cluster, no warm-up",
                               "gpu_fp32_vs_cpu_fp64": r,
                               "gpu_speedup_vs_cpu_force": cpu_ms / (ms2 / steps)}
    return out


def learned_bh(nbx, cpu, n, seed, eps, evals, sample=8192):
    """Learned correction at N = n two-galaxy: theta = 1.0 walk + the committed
    MLP (n_body_sim_pro_amd/weights) against plain theta = 1.0, 0.7 and 0.5
    and against the theta = 1.0 walk corrected by its quadrupole force alone
    (no network), errors from the FP64 direct sum (reference scalar order) on
    a seeded `sample`-particle subsample."""
    from n_body_sim_pro_amd import learned as L
    t0 = time.perf_counter()
    d = nbx.two_galaxy(n, seed=seed)
    t_gen = time.perf_counter() - t0
    idx = np.sort(np.random.default_rng(7).choice(n, size=min(sample, n), replace=False))
    out = {"n": n, "sample": int(idx.size), "generation_s": t_gen}
    mlp = L.Mlp.load(L.WEIGHTS)
    with nbx.Engine(n, theta=1.0) as e:
        e.upload(d["pos"], d["m"])
        t0 = time.perf_counter()
        ref = e.direct_fp64_targets(idx, 1.0, eps)
        out["reference_fp64_subsample_s"] = time.perf_counter() - t0
        # the quadrupole-corrected walk without the network: a + q, q from the
        # walk's own features (5, 6: q along e1 and across it, in units of |a|)
        feats, frames = e.bh_features(1.0, eps, frames=True)
        base = e.accelerations()[idx]
        f = feats[idx].astype(np.float64)
        fr = frames[idx].astype(np.float64)
        quad = base + np.linalg.norm(base, axis=1)[:, None] * (f[:, 5:6] * fr[:, 0]
                                                                + f[:, 6:7] * fr[:, 1])
        del feats, frames
        for name, theta, algo, model in (("bh_theta1.0", 1.0, "hip_barnes_hut", None),
                                          ("learned_theta1.0", 1.0, "hip_bh_learned", mlp),
                                          ("bh_theta0.7", 0.7, "hip_barnes_hut", None),
                                          ("bh_theta0.5", 0.5, "hip_barnes_hut", None)):
            e.theta = theta
            e.set_mlp(model)
            e.compute_force(algo, 1.0, eps)   # warm-up
            ts = []
            for _ in range(evals):
                e.compute_force(algo, 1.0, eps)
                ts.append(e.timings())
            err = rel_err(e.accelerations()[idx], ref)
            ms = float(np.mean([t["force_ms"] for t in ts]))
            out[name] = {"ms": ms, "mparticles_s": n / (ms * 1e-3) / 1e6,
                         "walk_ms": float(np.mean([t["walk_ms"] for t in ts])),
                         "median_rel_err": float(np.median(err)),
                         "p90_rel_err": float(np.quantile(err, 0.9)),
                         "max_rel_err": float(err.max())}
            if model is not None:
                out[name]["mlp_ms"] = float(np.mean([t["mlp_ms"] for t in ts]))
            e.set_mlp(None)
    mlp.close()
    qe = rel_err(quad, ref)
    lrn, b5 = out["learned_theta1.0"], out["bh_theta0.5"]
    out["quad_theta1.0"] = {"median_rel_err": float(np.median(qe)),
                            "p90_rel_err": float(np.quantile(qe, 0.9)),
                            "max_rel_err": float(qe.max()),
                            "ms": lrn["ms"] - lrn["mlp_ms"],
                            "note": "theta=1.0 feature walk + its quadrupole force, no network; "
                                    "ms = the learned path less its correction kernel"}
    out["network_error_ratio_vs_quad"] = lrn["median_rel_err"] / out["quad_theta1.0"][
        "median_rel_err"]
    out["speedup_vs_theta0.5"] = b5["ms"] / lrn["ms"]
    out["error_ratio_vs_theta0.5"] = lrn["median_rel_err"] / b5["median_rel_err"]
    out["error_gain_vs_theta1.0"] = out["bh_theta1.0"]["median_rel_err"] / lrn["median_rel_err"]
    out["target_met"] = bool(lrn["median_rel_err"] <= b5["median_rel_err"]
                             and out["speedup_vs_theta0.5"] >= 2.0)
    if cpu is not None:
        view, cols = host_view(d["pos"], d["m"])
        g = nbx.gravity(1.0, eps)
        runs = []
        for theta in (0.5, 1.0):
            r = cpu.time("barnes_hut", view, g, evals=1, theta=theta)
            r.update(err_summary(host_acc(cols)[idx], ref))
            runs.append(r)
        out["cpu_baseline"] = {"value": runs[0]["ms"], "unit": "ms per BH force eval (theta=0.5)",
                               "cores": runs[0]["threads"], "kind": cpu.kind,
                               "sample": f"one full N={n} evaluation per theta (the reference "
                                         "harness with --steps 1, --algorithm barnes_hut)",
                               "runs": runs,
                               "learned_gpu_speedup_vs_cpu_theta0.5": runs[0]["ms"] / lrn["ms"]}
    return out


def progress(msg, t0=[time.perf_counter()]):
    """One line per finished block on stderr (stdout carries only the JSON line)."""
    print(f"bench: {msg} ({time.perf_counter() - t0[0]:.1f} s)", file=sys.stderr, flush=True)


def measured_traffic(name):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary (profiles/<name>_traffic.json), if one was recorded for this
    workload; None otherwise."""
    path = ROOT / "profiles" / f"{name}_traffic.json"
    if not path.exists():
        return None
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return None


def main():
    args = parse()
    if args.copy_probe_mib > 0:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.add(a, 1.0, out=b)   # reads 4n bytes, writes 4n bytes
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_mib": args.copy_probe_mib, "read_bytes": 4 * n,
                          "write_bytes": 4 * n, "launches": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(multi-GPU sharding is described, not implemented).", file=sys.stderr)
        return 2
    import n_body_sim_pro_amd as nbx
    nbx.load()
    n = args.n
    t_gen = time.perf_counter()
    d = nbx.two_galaxy(n, seed=args.seed)
    t_gen = time.perf_counter() - t_gen

    # accuracy of the timed walk: FP64 direct sum on a seeded 4096-particle sample
    sample = np.sort(np.random.default_rng(7).choice(n, min(4096, n), replace=False))
    with nbx.Engine(n, theta=args.theta, reference_lists=True) as er:
        er.upload(d["pos"], d["m"], d["vel"])
        ref_sample = er.direct_fp64_targets(sample, 1.0, args.eps)
        er.compute_force("hip_barnes_hut", eps=args.eps)
        reflist_acc = er.accelerations()
        reflist_stats = er.stats()
    rn = np.linalg.norm(ref_sample, axis=1)

    e = nbx.Engine(n, theta=args.theta)
    e.upload(d["pos"], d["m"], d["vel"])
    e.compute_force("hip_barnes_hut", eps=args.eps)
    first_acc = e.accelerations()
    err_first = np.linalg.norm(first_acc[sample] - ref_sample, axis=1) / rn
    err_reflist = np.linalg.norm(reflist_acc[sample] - ref_sample, axis=1) / rn
    for _ in range(args.warmup):
        e.advance("hip_barnes_hut", dt=args.dt, steps=1, eps=args.eps)
    force_ms, walk_ms, build_ms, flops, inter = [], [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e.advance("hip_barnes_hut", dt=args.dt, steps=1, eps=args.eps)
        t = e.timings()
        s = e.stats()
        force_ms.append(t["force_ms"])
        walk_ms.append(t["walk_ms"])
        build_ms.append(t["force_ms"] - t["walk_ms"])
        flops.append(t["kernel_flops"])
        inter.append(s["accepted_approximations"] + s["exact_interactions"])
    elapsed = time.perf_counter() - t0
    stats = e.stats()
    tm = e.timings()
    e.close()

    value = float(np.mean(force_ms))
    walk = float(np.mean(walk_ms))
    ref_inter = float(reflist_stats["accepted_approximations"] + reflist_stats["exact_interactions"])
    achieved = float(np.mean([f / (w * 1e-3) for f, w in zip(flops, walk_ms)])) / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 walk; f64 keys/topology/COM and integrator",
        "data": "synthetic (seeded two-galaxy generator)",
        "config": {"workload": f"two-galaxy collision N={n}, Barnes-Hut theta={args.theta}, "
                               f"eps={args.eps}, leapfrog KDK dt={args.dt}",
                   "n": n, "theta": args.theta, "eps": args.eps, "dt": args.dt,
                   "seed": args.seed, "algorithm": "hip_barnes_hut"},
        "bh": {"build_ms": float(np.mean(build_ms)), "walk_ms": walk,
               "walk": "group walk (64 consecutive particles share one list of the cells "
                       "accepted against their bounding box or by every member's own test, "
                       "found with a per-wave stack over level-indexed cells; each "
                       "particle's list refines its reference list)",
               "median_rel_err_vs_fp64_direct": float(np.median(err_first)),
               "p99_rel_err_vs_fp64_direct": float(np.quantile(err_first, 0.99)),
               "max_rel_err_vs_fp64_direct": float(err_first.max()),
               "error_sample": f"{sample.size} seeded particles, FP64 direct sum in the "
                               "reference's scalar order",
               "reference_lists": {
                   "interactions_per_eval": int(reflist_stats["accepted_approximations"]
                                                + reflist_stats["exact_interactions"]),
                   "median_rel_err_vs_fp64_direct": float(np.median(err_reflist)),
                   "p99_rel_err_vs_fp64_direct": float(np.quantile(err_reflist, 0.99)),
                   "max_rel_err_vs_fp64_direct": float(err_reflist.max())},
               "interactions_per_eval": float(np.mean(inter)),
               "interactions_vs_reference_lists": float(np.mean(inter)) / ref_inter,
               "ginteractions_s": float(np.mean(inter)) / (value * 1e-3) / 1e9,
               "node_count": stats["node_count"],
               "internal_node_count": stats["internal_node_count"],
               "maximum_depth": stats["maximum_depth"],
               "phase_ms_last": {k: tm[k] for k in ("bbox_keys_ms", "sort_ms", "tree_ms",
                                                    "com_ms", "walk_ms")},
               "generation_s": t_gen},
        "roofline": {"kernel": "group_walk_kernel (FP32 group walk)",
                     "bound": "valu", "pipe": "valu", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "flops_per_unit": "20 per interaction (approximations + exact)",
                     "launch_ms": walk,
                     # the same walk time priced at the reference lists' interactions
                     "frac_at_reference_work": 20.0 * ref_inter / (walk * 1e-3) / 1e12
                                               / FP32_PEAK_TFLOPS},
    }
    tr = measured_traffic("group_walk_two_galaxy_1M")
    if tr is not None:
        out["roofline"]["traffic"] = tr.get("bytes_per_launch")
        out["roofline"]["traffic_source"] = tr.get("source")
    progress("Barnes-Hut headline")
    cpu = None if args.no_cpu_baseline else CpuBaseline()
    if cpu is not None:
        out["cpu_environment"] = cpu.environment()
    if not args.no_direct_fp64:
        out["direct_fp64"] = direct_fp64(nbx, cpu, args.dp_n, args.eps, args.seed, args.dt)
        progress("FP64 all-pairs block")
    if not args.no_all_pairs:
        out["all_pairs"] = all_pairs(nbx, cpu, args.ap_n, args.ap_evals, args.eps, args.seed,
                                     args.ap_drift_steps, args.dt)
        progress("FP32 all-pairs block")
    if not args.no_ensemble:
        out["ensemble"] = ensemble(nbx, cpu, args.ens_clusters, args.ens_steps, args.eps,
                                   args.dt, args.seed)
        progress("ensemble block")
    if not args.no_learned:
        out["learned"] = learned_bh(nbx, cpu, args.learned_n, args.seed, args.eps,
                                    args.learned_evals)
        progress("learned block")
    if cpu is not None:
        # the reference's Barnes-Hut paths on the same FP32-rounded input
        view, cols = host_view(d["pos"].astype(np.float32).astype(np.float64),
                               d["m"].astype(np.float32).astype(np.float64))
        g = nbx.gravity(1.0, args.eps)
        runs = []
        for algo in ("barnes_hut_openmp_avx2", "barnes_hut_openmp_avx512", "barnes_hut"):
            if cpu.available(algo):
                r = cpu.time(algo, view, g, evals=2, theta=args.theta, median=True)
                r.update(err_summary(host_acc(cols)[sample], ref_sample))
                runs.append(r)
        cpu_acc = host_acc(cols)   # the scalar walk's (the last run)
        best = min(runs, key=lambda r: r["ms"])
        # same interaction lists (per-particle reference lists) on both sides
        rel = rel_err(reflist_acc, cpu_acc)
        out["cpu_baseline"] = {
            "value": best["ms"], "unit": "ms per BH force eval", "cores": best["threads"],
            "kind": cpu.kind, "best": best["algorithm"],
            "sample": f"full {n}-particle workload per kernel: the reference harness's "
                      "no-warm-up mean of 2 evaluations on one tree, then one warm-up and "
                      "the median of 3 (value)",
            "no_warmup_mean_ms": best["no_warmup_mean_ms"],
            "interactions": runs[-1]["interactions"], "runs": runs,
            "cpu_brand": out["cpu_environment"]["cpu_brand"],
            "omp_get_max_threads": out["cpu_environment"]["omp_get_max_threads"],
            "nproc": out["cpu_environment"]["nproc"],
            "gpu_reference_lists_vs_cpu_median_rel_diff": float(np.median(rel)),
            "gpu_reference_lists_vs_cpu_max_rel_diff": float(rel.max()),
            "median_rel_err_vs_fp64_direct": runs[-1]["median_rel_err"],
            "gpu_speedup_vs_best": best["ms"] / value}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
