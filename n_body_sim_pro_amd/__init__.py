"""n_body_sim_pro_amd: an MI355X (gfx950) gravitational force engine.

The compute lives in libnbx.so (hand-written HIP kernels behind a C ABI,
include/nbx_engine.h, nbx_ensemble.h, nbx_learned.h); this package is the
Python host layer over it.
"""
from .abi import (BarnesHutStats, Diagnostics, EngineConfig, Error, Gravity, Timings, View)  # noqa: F401
from .engine import (ALGORITHMS, Engine, Ensemble, GalaxyPairParams, NBXError,  # noqa: F401
                     TileConfig, device_count, disk_galaxy, ensemble_clusters, ensemble_sizes,
                     force_function, galaxy_pair_defaults, gravity, load, make_view, plummer,
                     PRESETS, preset, random_cloud, set_host_threads, two_galaxy)
