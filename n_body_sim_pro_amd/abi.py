"""ctypes mirrors of the plain C structs in include/nbx_types.h and include/nbx_engine.h.

The layouts match the reference engine's public headers
(include/n_body_sim_pro/core/particle_system.h:38-50, physics/gravity.h:25-29,
barnes_hut/barnes_hut.h:31-40, diagnostics/numerics.h:22-34, core/status.h:18-35).
"""
import ctypes as C

STATUS_NAMES = {
    0: "ok", 1: "invalid argument", 2: "out of memory", 3: "invalid state",
    4: "not implemented", 5: "unsupported platform", 6: "overflow", 7: "i/o error",
    8: "insufficient memory",
}
STATUS_OK = 0
STATUS_INVALID_ARGUMENT = 1
STATUS_OUT_OF_MEMORY = 2
STATUS_INVALID_STATE = 3
STATUS_UNSUPPORTED_PLATFORM = 5
STATUS_OVERFLOW = 6

INTEGRATOR_EULER = 0
INTEGRATOR_LEAPFROG = 1
INTEGRATOR_VELOCITY_VERLET = 2


class View(C.Structure):
    _fields_ = [("particle_count", C.c_size_t)] + [
        (n, C.POINTER(C.c_double)) for n in (
            "positions_x", "positions_y", "positions_z",
            "velocities_x", "velocities_y", "velocities_z",
            "accelerations_x", "accelerations_y", "accelerations_z", "masses")]


class Gravity(C.Structure):
    _fields_ = [("gravitational_constant", C.c_double), ("softening_length", C.c_double),
                ("softening_squared", C.c_double)]


class Error(C.Structure):
    _fields_ = [("status", C.c_int), ("source_file", C.c_char_p), ("source_line", C.c_int),
                ("message", C.c_char * 256)]


class BarnesHutStats(C.Structure):
    _fields_ = [("node_count", C.c_size_t), ("leaf_count", C.c_size_t),
                ("internal_node_count", C.c_size_t), ("accepted_approximations", C.c_size_t),
                ("exact_interactions", C.c_size_t), ("maximum_depth", C.c_size_t),
                ("build_time_seconds", C.c_double), ("evaluation_time_seconds", C.c_double)]


class Diagnostics(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "kinetic_energy", "total_mass", "total_momentum_x", "total_momentum_y",
        "total_momentum_z", "total_angular_momentum_x", "total_angular_momentum_y",
        "total_angular_momentum_z", "center_of_mass_x", "center_of_mass_y",
        "center_of_mass_z")]


class EngineConfig(C.Structure):
    """NBXConfig in include/nbx_engine.h."""
    _fields_ = [("capacity", C.c_size_t), ("device", C.c_int), ("theta", C.c_double),
                ("flags", C.c_uint32)]


class Timings(C.Structure):
    """NBXTimings in include/nbx_engine.h (milliseconds from HIP events)."""
    _fields_ = [(n, C.c_double) for n in (
        "force_ms", "bbox_keys_ms", "sort_ms", "tree_ms", "com_ms", "walk_ms", "kernel_ms",
        "integrate_ms")] + [("kernel_bytes", C.c_double), ("kernel_flops", C.c_double),
                            ("mlp_ms", C.c_double)]
