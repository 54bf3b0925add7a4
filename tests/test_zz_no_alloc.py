"""nbx_direct_fp32_noalloc: the FP32 all-pairs entry point that allocates
nothing. Called through the C ABI with device pointers; its accelerations must
be bit-identical to the existing FP32 all-pairs path (nbx_engine_compute_force
with hip_direct_fp32, the engine behind nbx_hip_direct_fp32), on two calls in
a row, the second with a smaller input, and with and without softening.
Device memory for the inputs and outputs comes from hipMalloc of the HIP
runtime that libnbx.so itself links (a second HIP runtime in the process,
such as torch's bundled one, would not see the device)."""
import ctypes as C

import numpy as np
import pytest

from nbx_testlib import Error, gravity

pytestmark = pytest.mark.gpu

INVALID_ARGUMENT = 1
MAX_PARTICLES = 1 << 20


def _bind(lib):
    fn = lib.nbx_direct_fp32_noalloc
    fn.argtypes = [C.c_void_p] * 7 + [C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
    fn.restype = C.c_int
    return fn


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of libnbx.so's own HIP runtime."""

    H2D, D2H = 1, 2

    def __init__(self, nbx):
        nbx.load()
        self.lib = C.CDLL("libamdhip64.so.7")   # already loaded as libnbx.so's dependency
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipSetDevice.argtypes = [C.c_int]
        assert self.lib.hipSetDevice(0) == 0
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p

    def get(self, p, n):
        out = np.empty(n, dtype=np.float64)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.lib.hipFree(p)
        self.ptrs = []


def _noalloc(nbx, pos, m, G, eps):
    fn = _bind(nbx.load())
    hip = _Hip(nbx)
    n = len(m)
    try:
        cols = [hip.put(pos[:, k]) for k in range(3)]
        dm = hip.put(m)
        acc = [hip.put(np.full(n, np.nan)) for _ in range(3)]
        g = gravity(G, eps)
        err = Error()
        st = fn(*cols, dm, *acc, n, C.byref(g), None, C.byref(err))
        assert st == 0, err.message
        assert hip.lib.hipDeviceSynchronize() == 0
        return np.stack([hip.get(p, n) for p in acc], axis=1)
    finally:
        hip.free()


def _engine(nbx, pos, m, G, eps):
    with nbx.Engine(len(m)) as e:
        e.upload(pos, m)
        e.compute_force("hip_direct_fp32", G=G, eps=eps)
        return e.accelerations()


def test_noalloc_matches_direct_fp32_on_two_calls(nbx):
    big = nbx.plummer(70001, seed=5)
    small = nbx.plummer(4097, seed=6)
    print("noalloc: N=70001 Plummer seed 5, eps 0.01; then N=4097 seed 6, G=0.5", flush=True)
    a1 = _noalloc(nbx, big["pos"], big["m"], 1.0, 0.01)
    a2 = _noalloc(nbx, small["pos"], small["m"], 0.5, 0.01)
    assert np.all(np.isfinite(a1)) and np.all(np.isfinite(a2))
    assert np.array_equal(a1, _engine(nbx, big["pos"], big["m"], 1.0, 0.01))
    assert np.array_equal(a2, _engine(nbx, small["pos"], small["m"], 0.5, 0.01))


def test_noalloc_unsoftened_matches_direct_fp32(nbx):
    d = nbx.random_cloud(3001, seed=9)
    print("noalloc: N=3001 random cloud seed 9, eps 0", flush=True)
    a = _noalloc(nbx, d["pos"], d["m"], 1.0, 0.0)
    assert np.all(np.isfinite(a))
    assert np.array_equal(a, _engine(nbx, d["pos"], d["m"], 1.0, 0.0))


def test_noalloc_rejects_oversize_and_missing_arguments(nbx):
    """Rejected on the host before any device work (the pointers are never used)."""
    fn = _bind(nbx.load())
    err = Error()
    dummy = C.c_void_p(16)
    st = fn(*([dummy] * 7), MAX_PARTICLES + 1, C.byref(gravity()), None, C.byref(err))
    assert st == INVALID_ARGUMENT and b"1048576" in err.message
    st = fn(*([dummy] * 7), 10, None, None, C.byref(err))
    assert st == INVALID_ARGUMENT
    st = fn(*([None] * 7), 10, C.byref(gravity()), None, C.byref(err))
    assert st == INVALID_ARGUMENT
    assert fn(*([None] * 7), 0, C.byref(gravity()), None, None) == 0
