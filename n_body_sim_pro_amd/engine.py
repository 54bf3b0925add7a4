"""Python host layer over the C ABI in include/nbx_engine.h (libnbx.so).

Mirrors the reference's force-kernel plug point: ``force_function(name)``
returns the C function pointer with the NBodySimProForceFunction signature
(reference: include/n_body_sim_pro/physics/integrator.h:47-49), and ``Engine``
wraps the device-resident API. No fallback: when libnbx.so is missing or no
HIP device is visible, the calls raise.
"""
import ctypes as C
import pathlib

import numpy as np

from .abi import (BarnesHutStats, Diagnostics, EngineConfig, Error, Gravity, Timings, View,
                  STATUS_NAMES)

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "libnbx.so"

ALGORITHMS = {
    "hip_direct_fp64": 0,
    "hip_direct_fp32": 1,
    "hip_barnes_hut": 2,
    "hip_barnes_hut_fp64": 3,
    "hip_bh_learned": 4,
}
_FORCE_SYMBOLS = {
    "hip_direct_fp64": "nbx_hip_direct_fp64",
    "hip_direct_fp32": "nbx_hip_direct_fp32",
    "hip_barnes_hut": "nbx_hip_barnes_hut",
    "hip_barnes_hut_fp64": "nbx_hip_barnes_hut_fp64",
    "hip_bh_learned": "nbx_hip_bh_learned",
}

_lib = None


class NBXError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def load():
    """Loads libnbx.so (built in-tree by n_body_sim_pro_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m n_body_sim_pro_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    vp, gp, ep = C.POINTER(View), C.POINTER(Gravity), C.POINTER(Error)
    lib.nbx_engine_create.argtypes = [C.POINTER(EngineConfig), ep]
    lib.nbx_engine_create.restype = C.c_void_p
    lib.nbx_engine_destroy.argtypes = [C.c_void_p]
    lib.nbx_engine_set_theta.argtypes = [C.c_void_p, C.c_double]
    lib.nbx_engine_theta.argtypes = [C.c_void_p]
    lib.nbx_engine_theta.restype = C.c_double
    lib.nbx_engine_upload.argtypes = [C.c_void_p, vp, ep]
    lib.nbx_engine_download.argtypes = [C.c_void_p, vp, ep]
    lib.nbx_engine_compute_force.argtypes = [C.c_void_p, C.c_int, gp, ep]
    lib.nbx_engine_advance.argtypes = [C.c_void_p, C.c_int, gp, C.c_int, C.c_double, C.c_int, ep]
    lib.nbx_engine_diagnostics.argtypes = [C.c_void_p, gp, C.POINTER(Diagnostics),
                                           C.POINTER(C.c_double), ep]
    lib.nbx_engine_stats.argtypes = [C.c_void_p, C.POINTER(BarnesHutStats)]
    lib.nbx_engine_timings.argtypes = [C.c_void_p, C.POINTER(Timings)]
    lib.nbx_engine_bh_morton.argtypes = [C.c_void_p, C.POINTER(C.c_double),
                                         C.POINTER(C.c_uint64), C.POINTER(C.c_int64), ep]
    lib.nbx_engine_bh_export_nodes.argtypes = [C.c_void_p, C.POINTER(C.c_double), ep]
    lib.nbx_algorithm_from_name.argtypes = [C.c_char_p]
    lib.nbx_algorithm_name.argtypes = [C.c_int]
    lib.nbx_algorithm_name.restype = C.c_char_p
    lib.nbx_device_count.restype = C.c_int
    for sym in _FORCE_SYMBOLS.values():
        getattr(lib, sym).argtypes = [vp, gp, C.c_void_p, ep]
        getattr(lib, sym).restype = C.c_int
    _lib = lib
    return lib


def gravity(G=1.0, eps=0.01):
    g = Gravity()
    g.gravitational_constant = G
    g.softening_length = eps
    g.softening_squared = eps * eps
    return g


def force_function(name):
    """C function pointer for the drop-in force function `name` (hip_*)."""
    lib = load()
    return getattr(lib, _FORCE_SYMBOLS[name])


def _check(status, err):
    if status != 0:
        raise NBXError(status, err.message.decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_view(n, px=None, py=None, pz=None, vx=None, vy=None, vz=None, ax=None, ay=None,
              az=None, m=None):
    """A View over float64 numpy arrays (None fields stay NULL). Keep the arrays alive."""
    v = View()
    v.particle_count = int(n)
    for field, arr in (("positions_x", px), ("positions_y", py), ("positions_z", pz),
                       ("velocities_x", vx), ("velocities_y", vy), ("velocities_z", vz),
                       ("accelerations_x", ax), ("accelerations_y", ay), ("accelerations_z", az),
                       ("masses", m)):
        if arr is not None:
            assert arr.dtype == np.float64 and arr.flags.c_contiguous and arr.size >= n
            setattr(v, field, _ptr(arr))
    return v


class Engine:
    """Device-resident engine (one HIP stream and its buffers on one device)."""

    def __init__(self, capacity, device=0, theta=0.7, reference_lists=False, split=True,
                 member_test=True):
        """reference_lists: hip_barnes_hut walks each particle with its own
        reference interaction list (NBX_FLAG_BH_REFERENCE_LISTS) instead of the
        default group walk. split=False: groups that are not compact walk per
        particle instead of being split (NBX_FLAG_BH_NO_SPLIT). member_test
        (default): the group walk also accepts a cell that every member's own
        test accepts; False keeps the group's box test alone
        (NBX_FLAG_BH_BOX_TEST)."""
        self.lib = load()
        cfg = EngineConfig(capacity=int(capacity), device=int(device), theta=float(theta),
                           flags=(1 if reference_lists else 0) | (0 if split else 2)
                           | (0 if member_test else 8))
        err = Error()
        self.h = self.lib.nbx_engine_create(C.byref(cfg), C.byref(err))
        if not self.h:
            raise NBXError(err.status, err.message.decode(errors="replace"))
        self.capacity = int(capacity)
        self.device = int(device)
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.nbx_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def theta(self):
        return self.lib.nbx_engine_theta(self.h)

    @theta.setter
    def theta(self, value):
        self.lib.nbx_engine_set_theta(self.h, float(value))

    def upload_view(self, view):
        err = Error()
        _check(self.lib.nbx_engine_upload(self.h, C.byref(view), C.byref(err)), err)
        self.n = int(view.particle_count)

    def download_view(self, view):
        err = Error()
        _check(self.lib.nbx_engine_download(self.h, C.byref(view), C.byref(err)), err)

    def upload(self, pos, mass, vel=None, acc=None):
        """pos/vel/acc: (n, 3) arrays; mass: (n,)."""
        n = len(mass)
        cols = {}
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).T)
        cols.update(px=p[0].copy(), py=p[1].copy(), pz=p[2].copy())
        cols["m"] = np.ascontiguousarray(mass, dtype=np.float64).copy()
        if vel is not None:
            v = np.ascontiguousarray(np.asarray(vel, dtype=np.float64).T)
            cols.update(vx=v[0].copy(), vy=v[1].copy(), vz=v[2].copy())
        if acc is not None:
            a = np.ascontiguousarray(np.asarray(acc, dtype=np.float64).T)
            cols.update(ax=a[0].copy(), ay=a[1].copy(), az=a[2].copy())
        self.upload_view(make_view(n, **cols))

    def download(self):
        """Returns dict with pos, vel, acc (n, 3) and m (n,)."""
        n = self.n
        cols = {k: np.zeros(n) for k in ("px", "py", "pz", "vx", "vy", "vz", "ax", "ay", "az",
                                         "m")}
        self.download_view(make_view(n, **cols))
        return {"pos": np.stack([cols["px"], cols["py"], cols["pz"]], 1),
                "vel": np.stack([cols["vx"], cols["vy"], cols["vz"]], 1),
                "acc": np.stack([cols["ax"], cols["ay"], cols["az"]], 1), "m": cols["m"]}

    def accelerations(self):
        n = self.n
        cols = {k: np.zeros(n) for k in ("ax", "ay", "az")}
        self.download_view(make_view(n, **cols))
        return np.stack([cols["ax"], cols["ay"], cols["az"]], 1)

    def compute_force(self, algorithm, G=1.0, eps=0.01):
        algo = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
        err = Error()
        g = gravity(G, eps)
        _check(self.lib.nbx_engine_compute_force(self.h, algo, C.byref(g), C.byref(err)), err)

    def advance(self, algorithm, dt, steps, G=1.0, eps=0.01, integrator=1):
        algo = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
        err = Error()
        g = gravity(G, eps)
        _check(self.lib.nbx_engine_advance(self.h, algo, C.byref(g), int(integrator), float(dt),
                                           int(steps), C.byref(err)), err)

    def diagnostics(self, G=1.0, eps=0.01, potential=True):
        q = Diagnostics()
        u = C.c_double(0.0)
        err = Error()
        g = gravity(G, eps)
        _check(self.lib.nbx_engine_diagnostics(self.h, C.byref(g), C.byref(q),
                                               C.byref(u) if potential else None, C.byref(err)),
               err)
        out = {f: getattr(q, f) for f, _ in q._fields_}
        if potential:
            out["potential_energy"] = u.value
        return out

    def stats(self):
        s = BarnesHutStats()
        self.lib.nbx_engine_stats(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def timings(self):
        t = Timings()
        self.lib.nbx_engine_timings(self.h, C.byref(t))
        return {f: getattr(t, f) for f, _ in t._fields_}

    def bh_morton(self):
        n = self.n
        root = np.zeros(4)
        keys = np.zeros(n, dtype=np.uint64)
        perm = np.zeros(n, dtype=np.int64)
        err = Error()
        _check(self.lib.nbx_engine_bh_morton(self.h, root.ctypes.data_as(C.POINTER(C.c_double)),
                                             keys.ctypes.data_as(C.POINTER(C.c_uint64)),
                                             perm.ctypes.data_as(C.POINTER(C.c_int64)),
                                             C.byref(err)), err)
        return root, keys, perm

    def set_mlp(self, mlp):
        """Attach a learned correction (n_body_sim_pro_amd.learned.Mlp) or detach with None."""
        lib = _learned_lib()
        err = Error()
        _check(lib.nbx_engine_set_mlp(self.h, None if mlp is None else mlp.h, C.byref(err)), err)
        self._mlp = mlp   # keep the model alive while attached

    def bh_features(self, G=1.0, eps=0.01, frames=False):
        """Builds and walks the FP32 tree at the engine's theta; returns the raw
        (n, 16) learned-correction features in particle order, and with
        frames=True also the (n, 3, 3) local frames (rows e1, e2, e3). The
        walk's accelerations stay in the engine, as after hip_barnes_hut."""
        lib = _learned_lib()
        out = np.zeros((self.n, 16), dtype=np.float32)
        fr = np.zeros((self.n, 3, 3), dtype=np.float32) if frames else None
        err = Error()
        g = gravity(G, eps)
        fp = C.POINTER(C.c_float)
        _check(lib.nbx_engine_bh_features(self.h, C.byref(g), out.ctypes.data_as(fp),
                                          fr.ctypes.data_as(fp) if frames else None,
                                          C.byref(err)), err)
        return (out, fr) if frames else out

    def direct_fp64_targets(self, targets, G=1.0, eps=0.01):
        """FP64 direct accelerations (reference scalar order) of the given particles only."""
        lib = _learned_lib()
        idx = np.ascontiguousarray(targets, dtype=np.int64)
        out = np.zeros((len(idx), 3))
        err = Error()
        g = gravity(G, eps)
        _check(lib.nbx_engine_direct_fp64_targets(
            self.h, C.byref(g), idx.ctypes.data_as(C.POINTER(C.c_int64)), len(idx),
            out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(err)), err)
        return out

    def bh_export_nodes(self):
        count = self.stats()["node_count"]
        rows = np.zeros((count, 11))
        err = Error()
        _check(self.lib.nbx_engine_bh_export_nodes(
            self.h, rows.ctypes.data_as(C.POINTER(C.c_double)), C.byref(err)), err)
        return rows


def device_count():
    return int(load().nbx_device_count())


class GalaxyPairParams(C.Structure):
    """NBXGalaxyPairParams in include/nbx_presets.h."""
    _fields_ = [(n, C.c_double) for n in (
        "disk_scale_length", "disk_scale_height", "bulge_fraction", "bulge_scale",
        "separation", "pericenter", "inclination", "velocity_dispersion")]


def _generate(symbol, n, *args):
    lib = load()
    cols = {k: np.zeros(int(n)) for k in ("px", "py", "pz", "vx", "vy", "vz", "m")}
    v = make_view(n, **cols)
    fn = getattr(lib, symbol)
    fn.restype = C.c_int
    st = fn(C.byref(v), *args)
    if st != 0:
        raise NBXError(st, f"{symbol} failed")
    return {"pos": np.stack([cols["px"], cols["py"], cols["pz"]], 1),
            "vel": np.stack([cols["vx"], cols["vy"], cols["vz"]], 1), "m": cols["m"]}


def random_cloud(n, seed=42):
    """The reference's random_cloud preset (bit-identical)."""
    return _generate("nbx_preset_random_cloud", n, C.c_uint64(seed))


def plummer(n, seed=42, scale_radius=1.0, total_mass=1.0):
    """Plummer sphere with velocities from the distribution function."""
    return _generate("nbx_generate_plummer", n, C.c_double(scale_radius),
                     C.c_double(total_mass), C.c_uint64(seed))


def galaxy_pair_defaults():
    p = GalaxyPairParams()
    load().nbx_galaxy_pair_defaults(C.byref(p))
    return p


def two_galaxy(n, seed=42, params=None):
    """Two disk+bulge galaxies on a parabolic approach, 20 units apart."""
    return _generate("nbx_generate_two_galaxy", n,
                     C.byref(params) if params is not None else None, C.c_uint64(seed))


def disk_galaxy(n, seed=42, params=None):
    return _generate("nbx_generate_disk_galaxy", n,
                     C.byref(params) if params is not None else None, C.c_uint64(seed))


PRESETS = ("two_body", "random_cloud", "solar_system", "open_cluster", "globular_cluster",
           "spiral_galaxy", "elliptical_galaxy", "galaxy_collision", "triple_galaxy")


def preset(name, n, seed=42, parallel=False):
    """The reference's named preset `name` (presets.h:27-38) with its sequential
    generator (presets.c:387-467) or, with parallel=True, its per-particle
    generator (presets_parallel.c:286-482); both bit-identical to the reference."""
    lib = load()
    idx = PRESETS.index(name) if isinstance(name, str) else int(name)
    symbol = "nbx_preset_generate_parallel" if parallel else "nbx_preset_generate"
    fn = getattr(lib, symbol)
    fn.argtypes = [C.POINTER(View), C.c_int, C.c_uint64, C.POINTER(Error)]
    fn.restype = C.c_int
    cols = {k: np.zeros(int(n)) for k in ("px", "py", "pz", "vx", "vy", "vz", "m")}
    v = make_view(n, **cols)
    err = Error()
    st = fn(C.byref(v), idx, C.c_uint64(seed), C.byref(err))
    if st != 0:
        raise NBXError(st, f"{symbol}({name}, {n}) failed: {err.message.decode(errors='replace')}")
    return {"pos": np.stack([cols["px"], cols["py"], cols["pz"]], 1),
            "vel": np.stack([cols["vx"], cols["vy"], cols["vz"]], 1), "m": cols["m"]}


def set_host_threads(threads):
    """Worker threads of the host generators (0: all online CPUs)."""
    load().nbx_set_host_threads(C.c_int(int(threads)))


def ensemble_sizes(clusters, lo=64, hi=4096, seed=42):
    out = np.zeros(int(clusters), dtype=np.uint32)
    lib = load()
    lib.nbx_ensemble_sizes.restype = C.c_int
    st = lib.nbx_ensemble_sizes(C.c_size_t(clusters), C.c_uint32(lo), C.c_uint32(hi),
                                C.c_uint64(seed), out.ctypes.data_as(C.POINTER(C.c_uint32)))
    if st != 0:
        raise NBXError(st, "nbx_ensemble_sizes failed")
    return out


def _learned_lib():
    lib = load()
    if getattr(lib, "_nbx_learned_bound", False):
        return lib
    gp, ep = C.POINTER(Gravity), C.POINTER(Error)
    fp = C.POINTER(C.c_float)
    lib.nbx_mlp_param_count.restype = C.c_size_t
    lib.nbx_mlp_create.argtypes = [C.c_int, ep]
    lib.nbx_mlp_create.restype = C.c_void_p
    lib.nbx_mlp_destroy.argtypes = [C.c_void_p]
    lib.nbx_mlp_set_params.argtypes = [C.c_void_p, fp, fp, fp, C.c_float, ep]
    lib.nbx_mlp_get_params.argtypes = [C.c_void_p, fp, fp, fp, fp, ep]
    lib.nbx_mlp_forward.argtypes = [C.c_void_p, fp, C.c_size_t, fp, ep]
    lib.nbx_mlp_gradients.argtypes = [C.c_void_p, fp, fp, C.c_size_t, fp, fp, ep]
    lib.nbx_mlp_load_dataset.argtypes = [C.c_void_p, fp, fp, C.c_size_t, ep]
    lib.nbx_mlp_train.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_uint64, fp, ep]
    lib.nbx_engine_set_mlp.argtypes = [C.c_void_p, C.c_void_p, ep]
    lib.nbx_engine_bh_features.argtypes = [C.c_void_p, gp, fp, fp, ep]
    lib.nbx_engine_direct_fp64_targets.argtypes = [C.c_void_p, gp, C.POINTER(C.c_int64),
                                                   C.c_size_t, C.POINTER(C.c_double), ep]
    lib.nbx_mlp_probe_mfma.argtypes = [C.c_int, fp, fp, fp, ep]
    lib._nbx_learned_bound = True
    return lib


class TileConfig(C.Structure):
    """NBXTileConfig in include/nbx_ensemble.h."""
    _fields_ = [("targets_per_lane", C.c_int * 8), ("blocks_per_cu", C.c_int),
                ("source_tile", C.c_int)]


def _ensemble_lib():
    lib = load()
    if getattr(lib, "_nbx_ensemble_bound", False):
        return lib
    vp, gp, ep = C.POINTER(View), C.POINTER(Gravity), C.POINTER(Error)
    lib.nbx_ensemble_create.argtypes = [C.c_int, C.POINTER(C.c_uint32), C.c_size_t, ep]
    lib.nbx_ensemble_create.restype = C.c_void_p
    lib.nbx_ensemble_destroy.argtypes = [C.c_void_p]
    lib.nbx_ensemble_particle_count.argtypes = [C.c_void_p]
    lib.nbx_ensemble_particle_count.restype = C.c_size_t
    lib.nbx_ensemble_upload.argtypes = [C.c_void_p, vp, ep]
    lib.nbx_ensemble_download.argtypes = [C.c_void_p, vp, ep]
    lib.nbx_ensemble_force.argtypes = [C.c_void_p, gp, ep]
    lib.nbx_ensemble_advance.argtypes = [C.c_void_p, gp, C.c_double, C.c_int, C.c_int,
                                         C.POINTER(C.c_double), ep]
    lib.nbx_ensemble_energies.argtypes = [C.c_void_p, gp, C.POINTER(C.c_double), ep]
    lib.nbx_ensemble_get_tile.argtypes = [C.c_void_p, C.POINTER(TileConfig)]
    lib.nbx_ensemble_set_tile.argtypes = [C.c_void_p, C.POINTER(TileConfig), ep]
    lib.nbx_ensemble_autotune.argtypes = [C.c_void_p, gp, C.c_char_p, C.c_char_p, C.c_size_t, ep]
    lib._nbx_ensemble_bound = True
    return lib


class Ensemble:
    """Batched ensemble of independent clusters (include/nbx_ensemble.h)."""

    def __init__(self, sizes, device=0):
        self.lib = _ensemble_lib()
        self.sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        err = Error()
        self.h = self.lib.nbx_ensemble_create(
            int(device), self.sizes.ctypes.data_as(C.POINTER(C.c_uint32)), self.sizes.size,
            C.byref(err))
        if not self.h:
            raise NBXError(err.status, err.message.decode(errors="replace"))
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes.astype(np.int64))])
        self.n = int(self.offsets[-1])

    def close(self):
        if getattr(self, "h", None):
            self.lib.nbx_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, pos, vel, mass):
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).T).copy()
        v = np.ascontiguousarray(np.asarray(vel, dtype=np.float64).T).copy()
        m = np.ascontiguousarray(mass, dtype=np.float64).copy()
        zeros = np.zeros((3, self.n))
        view = make_view(self.n, px=p[0], py=p[1], pz=p[2], vx=v[0], vy=v[1], vz=v[2],
                         ax=zeros[0], ay=zeros[1], az=zeros[2], m=m)
        err = Error()
        _check(self.lib.nbx_ensemble_upload(self.h, C.byref(view), C.byref(err)), err)

    def download(self):
        cols = {k: np.zeros(self.n) for k in ("px", "py", "pz", "vx", "vy", "vz", "ax", "ay",
                                              "az", "m")}
        err = Error()
        _check(self.lib.nbx_ensemble_download(self.h, C.byref(make_view(self.n, **cols)),
                                              C.byref(err)), err)
        return {"pos": np.stack([cols["px"], cols["py"], cols["pz"]], 1),
                "vel": np.stack([cols["vx"], cols["vy"], cols["vz"]], 1),
                "acc": np.stack([cols["ax"], cols["ay"], cols["az"]], 1), "m": cols["m"]}

    def force(self, G=1.0, eps=0.01):
        err = Error()
        g = gravity(G, eps)
        _check(self.lib.nbx_ensemble_force(self.h, C.byref(g), C.byref(err)), err)

    def advance(self, dt, steps, graph_steps=50, G=1.0, eps=0.01):
        err = Error()
        g = gravity(G, eps)
        ms = C.c_double(0.0)
        _check(self.lib.nbx_ensemble_advance(self.h, C.byref(g), float(dt), int(steps),
                                             int(graph_steps), C.byref(ms), C.byref(err)), err)
        return ms.value

    def energies(self, G=1.0, eps=0.01):
        out = np.zeros(2 * self.sizes.size)
        err = Error()
        g = gravity(G, eps)
        _check(self.lib.nbx_ensemble_energies(self.h, C.byref(g),
                                              out.ctypes.data_as(C.POINTER(C.c_double)),
                                              C.byref(err)), err)
        return out.reshape(-1, 2)

    def tile(self):
        t = TileConfig()
        self.lib.nbx_ensemble_get_tile(self.h, C.byref(t))
        return {"targets_per_lane": list(t.targets_per_lane), "blocks_per_cu": t.blocks_per_cu,
                "source_tile": t.source_tile}

    def set_tile(self, targets_per_lane, blocks_per_cu, source_tile):
        t = TileConfig((C.c_int * 8)(*targets_per_lane), blocks_per_cu, source_tile)
        err = Error()
        _check(self.lib.nbx_ensemble_set_tile(self.h, C.byref(t), C.byref(err)), err)

    def autotune(self, G=1.0, eps=0.01, cache_path=None):
        err = Error()
        g = gravity(G, eps)
        buf = C.create_string_buffer(512)
        _check(self.lib.nbx_ensemble_autotune(
            self.h, C.byref(g), cache_path.encode() if cache_path else None, buf, 512,
            C.byref(err)), err)
        import json
        return json.loads(buf.value.decode())


def ensemble_clusters(sizes, seed=42):
    """Concatenated Plummer-DF clusters (a=1, M=1 each), one seed per cluster."""
    sizes = np.asarray(sizes, dtype=np.int64)
    pos, vel, m = [], [], []
    for k, n in enumerate(sizes):
        d = plummer(int(n), seed=seed * 1_000_003 + k)
        pos.append(d["pos"])
        vel.append(d["vel"])
        m.append(d["m"])
    return {"pos": np.concatenate(pos), "vel": np.concatenate(vel), "m": np.concatenate(m)}
