/*
 * nbx_engine.h - C ABI of the MI355X gravitational force engine (libnbx.so).
 *
 * Plain C: no HIP or C++ types cross this boundary. Built by hipcc for gfx950,
 * callable from C17. One engine = one HIP device + one HIP stream + its device
 * buffers; an engine is not reentrant (one host thread at a time), and distinct
 * engines are independent.
 *
 * Two ways in:
 *   1. Drop-in force functions with the exact NBodySimProForceFunction signature
 *      (reference: include/n_body_sim_pro/physics/integrator.h:47-49). They read
 *      positions and masses from the caller's host view and OVERWRITE its
 *      accelerations for [0, particle_count) (reference contract:
 *      include/n_body_sim_pro/physics/gravity.h:43-44). `context` is an
 *      NBXEngine* or NULL (then a library-owned engine on device 0 is used).
 *      Each call copies the view to the device and the accelerations back.
 *   2. The device-resident engine API below: upload once, then force
 *      evaluations, leapfrog steps and diagnostics run on device-resident FP64
 *      state with no per-evaluation PCIe traffic.
 *
 * Errors: every function returns an NBodySimProStatus and fills the optional
 * NBodySimProError (NULL allowed); nothing aborts. HIP errors map to
 * OUT_OF_MEMORY (allocation), UNSUPPORTED_PLATFORM (no device) or
 * INVALID_STATE (launch/runtime failure).
 */
#ifndef NBX_ENGINE_H
#define NBX_ENGINE_H

#include "nbx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct NBXEngine NBXEngine;

/* Force algorithms (the `--algorithm` names are given beside each). */
typedef enum NBXAlgorithm {
    /* hip_direct_fp64: all-pairs FP64, ascending-j order per target, no
     * contraction: bit-identical to n_body_sim_pro_gravity_compute_acceleration_reference
     * (src/c/physics/gravity.c:15-70). */
    NBX_ALGORITHM_DIRECT_FP64 = 0,
    /* hip_direct_fp32: all-pairs, FP32 pair arithmetic on FP32-rounded
     * positions/masses, tile-compensated FP32 accumulation. */
    NBX_ALGORITHM_DIRECT_FP32 = 1,
    /* hip_barnes_hut: octree built from FP32-rounded positions and masses
     * (keys, order and topology exact in FP64 on those values), FP32 walk:
     * the group walk by default, per-particle reference lists with
     * NBX_FLAG_BH_REFERENCE_LISTS. */
    NBX_ALGORITHM_BARNES_HUT = 2,
    /* hip_barnes_hut_fp64: same device tree on the FP64 positions with an
     * FP64 walk in the reference's visit order: bit-identical to
     * n_body_sim_pro_barnes_hut_compute_acceleration (src/c/barnes_hut/barnes_hut.c:638-690). */
    NBX_ALGORITHM_BARNES_HUT_FP64 = 3,
    /* hip_bh_learned: hip_barnes_hut plus the learned residual MLP. Falls back
     * to plain hip_barnes_hut when no trained weights are loaded. */
    NBX_ALGORITHM_BH_LEARNED = 4
} NBXAlgorithm;

/* NBXConfig.flags bits. */
enum {
    /* hip_barnes_hut walks every particle with its own reference interaction
     * list (the reference's acceptance per particle, reference counts in the
     * stats) instead of the default group walk, in which each wave of 64
     * consecutive particles shares one list of the cells accepted against the
     * group's bounding box or by every member (every particle's list then
     * refines its reference list: more interactions, same or smaller error). */
    NBX_FLAG_BH_REFERENCE_LISTS = 1u,
    /* group walk only: a group of 64 that is not compact (its smallest
     * enclosing cell holds more than 2048 particles) walks each of its
     * particles with its reference list instead of being split into compact
     * parts that each share a list. Slower; for comparisons and tests. */
    NBX_FLAG_BH_NO_SPLIT = 2u,
    /* group walk: a cell that the group's bounding-box test opens is still
     * accepted when every member's own reference test accepts it (fewer
     * interactions; each list still refines its members' reference lists).
     * This is the group walk's default; the bit only states it. */
    NBX_FLAG_BH_MEMBER_TEST = 4u,
    /* group walk only: the group's bounding-box test alone (no member test):
     * longer lists, same or smaller error. For comparisons and tests. */
    NBX_FLAG_BH_BOX_TEST = 8u
};

typedef struct NBXConfig {
    size_t capacity;  /* maximum particle count */
    int device;       /* HIP device ordinal */
    double theta;     /* Barnes-Hut opening angle (reference default 0.7) */
    uint32_t flags;   /* NBX_FLAG_* bits, 0 for the defaults */
} NBXConfig;

/* Per-evaluation timings from HIP events on the engine's stream (ms). */
typedef struct NBXTimings {
    double force_ms;      /* whole force evaluation */
    double bbox_keys_ms;  /* BH: bounding cube + keys */
    double sort_ms;       /* BH: radix sort into Morton order */
    double tree_ms;       /* BH: tree order, depths and node construction */
    double com_ms;        /* BH: centre-of-mass pass */
    double walk_ms;       /* BH: walk (the dominant BH kernel) */
    double kernel_ms;     /* dominant kernel of the last evaluation */
    double integrate_ms;  /* last advance(): total; 0 after a force evaluation alone */
    double kernel_bytes;  /* algorithmic bytes moved by the dominant kernel */
    double kernel_flops;  /* algorithmic flops of the dominant kernel (20 per interaction) */
    double mlp_ms;        /* hip_bh_learned: the learned correction kernel; 0 otherwise */
} NBXTimings;

/* Name <-> algorithm for the `--algorithm` vocabulary; -1 / NULL when unknown. */
int nbx_algorithm_from_name(const char* name);
const char* nbx_algorithm_name(int algorithm);
/* Name (hipDeviceProp_t::name, NUL-terminated in `len` bytes) and compute-unit
 * count of HIP device `device`; 0 on success, -1 without such a device. For
 * reports (the reference prints its CPU, HeadlessRunner.cpp:110-125). */
int nbx_device_info(int device, char* name, size_t len, int* compute_units);

NBXEngine* nbx_engine_create(const NBXConfig* config, NBodySimProError* error);
void nbx_engine_destroy(NBXEngine* engine);
void nbx_engine_set_theta(NBXEngine* engine, double theta);
double nbx_engine_theta(const NBXEngine* engine);
/* Particles currently on the device (the last upload's count) and the
 * engine's capacity (the particle system's count and capacity accessors,
 * particle_system.h). */
size_t nbx_engine_particle_count(const NBXEngine* engine);
size_t nbx_engine_capacity(const NBXEngine* engine);
/* Device memory the engine holds now: the particle arrays plus its grown
 * work buffers (the reference's n_body_sim_pro_memory_estimate_simulation,
 * src/c/memory/memory_estimate.c, reports the host-side counterpart). */
size_t nbx_engine_device_bytes(const NBXEngine* engine);

/* Host view -> device (positions, velocities, accelerations, masses). */
NBodySimProStatus nbx_engine_upload(NBXEngine* engine, const NBodySimProParticleSystemView* view,
                                    NBodySimProError* error);
/* Device -> host view (same fields). */
NBodySimProStatus nbx_engine_download(NBXEngine* engine, const NBodySimProParticleSystemView* view,
                                      NBodySimProError* error);
/* Accelerations for the device-resident state (overwritten). */
NBodySimProStatus nbx_engine_compute_force(NBXEngine* engine, NBXAlgorithm algorithm,
                                           const NBodySimProGravity* gravity,
                                           NBodySimProError* error);
/*
 * `steps` integrator steps on the device (reference: src/c/physics/integrator.c:141-172).
 * Like the reference, the accelerations on the device must be valid for the
 * current positions before the first leapfrog / Verlet step.
 */
NBodySimProStatus nbx_engine_advance(NBXEngine* engine, NBXAlgorithm algorithm,
                                     const NBodySimProGravity* gravity,
                                     NBodySimProIntegratorType integrator, double timestep,
                                     int steps, NBodySimProError* error);
/* FP64 O(N) quantities (src/c/diagnostics/numerics.c:5-81); when `potential`
 * is non-NULL also the FP64 O(N^2) softened potential (numerics.c:83-116). */
NBodySimProStatus nbx_engine_diagnostics(NBXEngine* engine, const NBodySimProGravity* gravity,
                                         NBodySimProDiagnosticsQuantities* quantities,
                                         double* potential, NBodySimProError* error);
/* Statistics of the last Barnes-Hut evaluation (barnes_hut.h:116-117 contract). */
int nbx_engine_stats(const NBXEngine* engine, NBodySimProBarnesHutStats* stats);
int nbx_engine_timings(const NBXEngine* engine, NBXTimings* timings);

/*
 * Parity probes of the last Barnes-Hut build (test and tooling use).
 * root[4] = centre xyz, half. keys: 63-bit Morton key per particle in
 * ORIGINAL order; permutation[i] = original index of the i-th particle in
 * ascending (key, original index) order, exactly the reference's stable
 * LSD order (barnes_hut.c:312-339). Any pointer may be NULL.
 */
NBodySimProStatus nbx_engine_bh_morton(NBXEngine* engine, double* root, uint64_t* keys,
                                       int64_t* permutation, NBodySimProError* error);
/* Node rows of the last build, 11 doubles each: depth, centre xyz, half,
 * particle count, mass, COM xyz, leaf (Morton rank, -1 internal). `rows`
 * holds node_count*11 doubles (node_count from nbx_engine_stats). */
NBodySimProStatus nbx_engine_bh_export_nodes(NBXEngine* engine, double* rows,
                                             NBodySimProError* error);

/* ---- drop-in force functions (NBodySimProForceFunction) ---- */
NBodySimProStatus nbx_hip_direct_fp64(const NBodySimProParticleSystemView* view,
                                      const NBodySimProGravity* gravity, void* context,
                                      NBodySimProError* error);
NBodySimProStatus nbx_hip_direct_fp32(const NBodySimProParticleSystemView* view,
                                      const NBodySimProGravity* gravity, void* context,
                                      NBodySimProError* error);
NBodySimProStatus nbx_hip_barnes_hut(const NBodySimProParticleSystemView* view,
                                     const NBodySimProGravity* gravity, void* context,
                                     NBodySimProError* error);
NBodySimProStatus nbx_hip_barnes_hut_fp64(const NBodySimProParticleSystemView* view,
                                          const NBodySimProGravity* gravity, void* context,
                                          NBodySimProError* error);
NBodySimProStatus nbx_hip_bh_learned(const NBodySimProParticleSystemView* view,
                                     const NBodySimProGravity* gravity, void* context,
                                     NBodySimProError* error);

/*
 * FP32 all-pairs accelerations that allocate nothing. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * Same arithmetic and kernels as nbx_hip_direct_fp32, hence identical
 * accelerations, but every pointer is a DEVICE pointer on the current HIP
 * device (count doubles each), and the work is only enqueued on `stream`
 * (a hipStream_t; NULL = the null stream): the caller synchronizes. The
 * FP32 staging lives in a fixed per-device array of the library, so
 * count <= 1048576 (larger counts: INVALID_ARGUMENT), and calls that overlap
 * on one GPU race on it.
 */
NBodySimProStatus nbx_direct_fp32_noalloc(const double* positions_x, const double* positions_y,
                                          const double* positions_z, const double* masses,
                                          double* accelerations_x, double* accelerations_y,
                                          double* accelerations_z, size_t count,
                                          const NBodySimProGravity* gravity, void* stream,
                                          NBodySimProError* error);

/* Number of visible HIP devices (0 when none). */
int nbx_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* NBX_ENGINE_H */
