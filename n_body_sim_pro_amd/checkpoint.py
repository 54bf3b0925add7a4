"""".hpcs" checkpoints (include/nbx_checkpoint.h) from Python.

Host-side reading and writing (no GPU needed) and the engine's device-state
save / resume, over libnbx.so.
"""
import ctypes as C

import numpy as np

from .abi import Error, View
from .engine import Engine, _check, load, make_view

ARRAY_NAMES = ("px", "py", "pz", "vx", "vy", "vz", "ax", "ay", "az", "m")
MAGIC = 0x53504348
VERSION = 1


class CheckpointHeader(C.Structure):
    """NBodySimProCheckpointHeader (checkpoint.h:33-44), natural alignment."""
    _fields_ = [("magic", C.c_uint32), ("version", C.c_uint32),
                ("particle_count", C.c_size_t), ("simulation_time", C.c_double),
                ("timestep", C.c_double), ("integrator", C.c_int32), ("theta", C.c_double),
                ("barnes_hut_enabled", C.c_int32), ("random_seed", C.c_uint64),
                ("preset", C.c_int32)]


_bound = None


def _lib():
    global _bound
    if _bound is None:
        lib = load()
        hp, vp, ep = C.POINTER(CheckpointHeader), C.POINTER(View), C.POINTER(Error)
        lib.nbx_checkpoint_write.argtypes = [C.c_char_p, vp, hp, ep]
        lib.nbx_checkpoint_peek.argtypes = [C.c_char_p, hp, ep]
        lib.nbx_checkpoint_read.argtypes = [C.c_char_p, hp, vp, C.c_size_t, ep]
        lib.nbx_engine_checkpoint_save.argtypes = [C.c_void_p, C.c_char_p, hp, ep]
        lib.nbx_engine_checkpoint_load.argtypes = [C.c_void_p, C.c_char_p, hp, ep]
        for name in ("nbx_checkpoint_write", "nbx_checkpoint_peek", "nbx_checkpoint_read",
                     "nbx_engine_checkpoint_save", "nbx_engine_checkpoint_load"):
            getattr(lib, name).restype = C.c_int
        lib.nbx_engine_particle_count.argtypes = [C.c_void_p]
        lib.nbx_engine_particle_count.restype = C.c_size_t
        lib.nbx_engine_capacity.argtypes = [C.c_void_p]
        lib.nbx_engine_capacity.restype = C.c_size_t
        _bound = lib
    return _bound


def header(n, simulation_time=0.0, timestep=1e-3, integrator=1, theta=0.7,
           barnes_hut=True, seed=42, preset=0):
    h = CheckpointHeader()
    h.magic, h.version = MAGIC, VERSION
    h.particle_count = int(n)
    h.simulation_time, h.timestep = float(simulation_time), float(timestep)
    h.integrator, h.theta = int(integrator), float(theta)
    h.barnes_hut_enabled = 1 if barnes_hut else 0
    h.random_seed, h.preset = int(seed), int(preset)
    return h


def write(path, arrays, hdr):
    """arrays: dict of the ten float64 arrays (ARRAY_NAMES)."""
    a = {k: np.ascontiguousarray(arrays[k], dtype=np.float64) for k in ARRAY_NAMES}
    v = make_view(len(a["px"]), *(a[k] for k in ARRAY_NAMES))
    err = Error()
    _check(_lib().nbx_checkpoint_write(str(path).encode(), C.byref(v), C.byref(hdr),
                                       C.byref(err)), err)


def peek(path):
    h, err = CheckpointHeader(), Error()
    _check(_lib().nbx_checkpoint_peek(str(path).encode(), C.byref(h), C.byref(err)), err)
    return h


def read(path, capacity=None):
    """(header, dict of the ten arrays)."""
    if capacity is None:
        capacity = peek(path).particle_count
    a = {k: np.zeros(max(int(capacity), 1)) for k in ARRAY_NAMES}
    v = make_view(capacity, *(a[k] for k in ARRAY_NAMES))
    h, err = CheckpointHeader(), Error()
    _check(_lib().nbx_checkpoint_read(str(path).encode(), C.byref(h), C.byref(v), int(capacity),
                                      C.byref(err)), err)
    n = v.particle_count
    return h, {k: a[k][:n].copy() for k in ARRAY_NAMES}


def save_engine(engine: Engine, path, hdr):
    err = Error()
    _check(_lib().nbx_engine_checkpoint_save(engine.h, str(path).encode(), C.byref(hdr),
                                             C.byref(err)), err)


def load_engine(engine: Engine, path):
    h, err = CheckpointHeader(), Error()
    _check(_lib().nbx_engine_checkpoint_load(engine.h, str(path).encode(), C.byref(h),
                                             C.byref(err)), err)
    engine.n = int(h.particle_count)
    return h
