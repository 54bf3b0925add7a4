"""Builds libnbx.so in-tree for gfx950: every HIP translation unit under csrc/
plus the C17 host sources under host/, linked into one shared library, and the
C17 benchmark driver nbx_benchmark and the headless runner nbx_headless.

    python -m n_body_sim_pro_amd.build
"""
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
OBJ = PKG / "_build"
LIB = PKG / "libnbx.so"
BENCH_EXE = PKG / "nbx_benchmark"
HEADLESS_EXE = PKG / "nbx_headless"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NBX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
C_FLAGS = ["-std=c17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
           f"-I{ROOT / 'include'}"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(map(str, cmd))}\n"
                           f"{r.stdout}\n{r.stderr}")
    return r


def build(verbose=False):
    OBJ.mkdir(exist_ok=True)
    hip_srcs = sorted(CSRC.glob("*.hip"))
    c_srcs = sorted(HOST.glob("nbx_*.c"))
    c_srcs = [p for p in c_srcs if p.name not in ("nbx_benchmark.c", "nbx_headless.c")]
    jobs = []
    for src in hip_srcs:
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(OBJ / (src.stem + ".o"))])
    for src in c_srcs:
        jobs.append(["gcc", *C_FLAGS, "-c", str(src), "-o", str(OBJ / (src.stem + ".c.o"))])
    workers = min(8, max(1, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(workers) as ex:
        for r in ex.map(_run, jobs):
            if verbose and r.stderr.strip():
                print(r.stderr, file=sys.stderr)
    objs = [str(OBJ / (s.stem + ".o")) for s in hip_srcs] + \
           [str(OBJ / (s.stem + ".c.o")) for s in c_srcs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    bench_src = HOST / "nbx_benchmark.c"
    if bench_src.exists():
        _run(["gcc", *C_FLAGS, "-fopenmp", str(bench_src), "-o", str(BENCH_EXE), f"-L{PKG}",
              "-lnbx", f"-Wl,-rpath,{PKG}", "-lm"])
    headless_src = HOST / "nbx_headless.c"
    if headless_src.exists():
        _run(["gcc", *C_FLAGS, str(headless_src), "-o", str(HEADLESS_EXE), f"-L{PKG}",
              "-lnbx", f"-Wl,-rpath,{PKG}", "-lm"])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
