/*
 * nbx_ensemble.h - batched ensemble of independent clusters (C ABI, libnbx.so).
 *
 * Thousands of independent self-gravitating clusters of uneven size share one
 * device: particles are stored back to back (cluster k owns
 * [offset_k, offset_k + size_k)); a force evaluation is ONE persistent launch
 * whose waves pull (cluster, target-tile) work items from a device queue,
 * costliest first (cost = targets x cluster size); leapfrog steps are
 * captured in a HIP graph and replayed. Pair arithmetic is FP32 with
 * tile-compensated FP32 sums (as hip_direct_fp32); positions, velocities and
 * the per-cluster energies stay FP64. There is no interaction between
 * clusters. The reference has no batched mode; per cluster the arithmetic
 * contract is the reference's all-pairs kernel (src/c/physics/gravity.c:15-70)
 * and leapfrog (src/c/physics/integrator.c:47-85).
 */
#ifndef NBX_ENSEMBLE_H
#define NBX_ENSEMBLE_H

#include "nbx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct NBXEnsemble NBXEnsemble;

/* Work-item shape: targets per lane for each cluster-size bucket
 * (bucket b holds sizes in [2^(b+6), 2^(b+7)), b = 0..7, larger sizes use
 * bucket 7), and the number of resident blocks per CU of the persistent launch. */
typedef struct NBXTileConfig {
    int targets_per_lane[8];  /* each 1, 2 or 4 */
    int blocks_per_cu;        /* 1..8 */
    int source_tile;          /* sources per compensated partial sum (multiple of 4) */
} NBXTileConfig;

NBXEnsemble* nbx_ensemble_create(int device, const uint32_t* sizes, size_t clusters,
                                 NBodySimProError* error);
void nbx_ensemble_destroy(NBXEnsemble* ensemble);
size_t nbx_ensemble_particle_count(const NBXEnsemble* ensemble);

/* Concatenated host view of every cluster (particle_count must equal the
 * ensemble total). */
NBodySimProStatus nbx_ensemble_upload(NBXEnsemble* ensemble,
                                      const NBodySimProParticleSystemView* view,
                                      NBodySimProError* error);
NBodySimProStatus nbx_ensemble_download(NBXEnsemble* ensemble,
                                        const NBodySimProParticleSystemView* view,
                                        NBodySimProError* error);
/* Accelerations of every particle from its own cluster (one launch). */
NBodySimProStatus nbx_ensemble_force(NBXEnsemble* ensemble, const NBodySimProGravity* gravity,
                                     NBodySimProError* error);
/*
 * `steps` leapfrog KDK steps of every cluster; the accelerations must be
 * valid first (call nbx_ensemble_force once). Steps run as replays of a
 * captured HIP graph of `graph_steps` steps (built on first use, rebuilt when
 * dt, gravity or the tile configuration change); a remainder runs eagerly.
 * `elapsed_ms` (nullable) receives the device time of the steps.
 */
NBodySimProStatus nbx_ensemble_advance(NBXEnsemble* ensemble, const NBodySimProGravity* gravity,
                                       double timestep, int steps, int graph_steps,
                                       double* elapsed_ms, NBodySimProError* error);
/* FP64 kinetic and softened potential energy per cluster: out[2k] = K_k,
 * out[2k+1] = U_k (src/c/diagnostics/numerics.c:5-116 restricted to cluster k). */
NBodySimProStatus nbx_ensemble_energies(NBXEnsemble* ensemble, const NBodySimProGravity* gravity,
                                        double* out, NBodySimProError* error);

/* Current / explicit work-item shape. */
void nbx_ensemble_get_tile(const NBXEnsemble* ensemble, NBXTileConfig* config);
NBodySimProStatus nbx_ensemble_set_tile(NBXEnsemble* ensemble, const NBXTileConfig* config,
                                        NBodySimProError* error);
/*
 * Autotuner: for every size bucket present, times the force launch with 1, 2
 * and 4 targets per lane on that bucket's clusters alone and keeps the
 * fastest; then times 1..8 blocks per CU and 64..512-source tiles on the whole
 * ensemble. Results are cached per (device architecture, bucket histogram) in
 * the file named by `cache_path` (NULL: no file). `report` (nullable) gets a
 * one-line JSON summary.
 */
NBodySimProStatus nbx_ensemble_autotune(NBXEnsemble* ensemble, const NBodySimProGravity* gravity,
                                        const char* cache_path, char* report, size_t report_size,
                                        NBodySimProError* error);

#ifdef __cplusplus
}
#endif

#endif /* NBX_ENSEMBLE_H */
