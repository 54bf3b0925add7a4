/*
 * nbx_checkpoint.h - ".hpcs" simulation checkpoints (C ABI, libnbx.so).
 *
 * Reads and writes version-1 checkpoint files byte-compatible with the
 * reference engine's (include/n_body_sim_pro/checkpoint/checkpoint.h:13-66,
 * src/c/checkpoint/checkpoint.c:9-206), so a state saved by the reference's
 * `n_body_sim_pro save` resumes on the GPU engine and the other way round.
 *
 * File layout (little-endian, as the reference writes it on x86-64):
 *   0   u32 magic 0x53504348 (the bytes "HCPS")   4   u32 version (1)
 *   8   header record, 72 bytes: u32 magic, u32 version, u64 particle_count,
 *       f64 simulation_time, f64 timestep, i32 integrator, 4 pad, f64 theta,
 *       i32 barnes_hut_enabled, 4 pad, u64 random_seed, i32 preset, 4 pad
 *   80  ten arrays of particle_count f64: positions x y z, velocities x y z,
 *       accelerations x y z, masses
 * The reference writes the header record as a raw struct (padding included);
 * here every field is written explicitly and padding is zero, and reads parse
 * the fields at those offsets and check the file length.
 */
#ifndef NBX_CHECKPOINT_H
#define NBX_CHECKPOINT_H

#include "nbx_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef N_BODY_SIM_PRO_CHECKPOINT_CHECKPOINT_H
/* include/n_body_sim_pro/checkpoint/checkpoint.h:31-44 */
enum { N_BODY_SIM_PRO_CHECKPOINT_MAGIC = 0x53504348, N_BODY_SIM_PRO_CHECKPOINT_VERSION = 1 };

typedef struct NBodySimProCheckpointHeader {
    uint32_t magic;
    uint32_t version;
    size_t particle_count;
    double simulation_time;
    double timestep;
    int32_t integrator;
    double theta;
    int32_t barnes_hut_enabled;
    uint64_t random_seed;
    int32_t preset;
} NBodySimProCheckpointHeader;
#endif

enum { NBX_CHECKPOINT_DATA_OFFSET = 80 };

/* Replaces n_body_sim_pro_checkpoint_write (checkpoint.c:9-78): writes the
 * view's N particles and the header; N must equal header->particle_count. */
NBodySimProStatus nbx_checkpoint_write(const char* path, const NBodySimProParticleSystemView* view,
                                       const NBodySimProCheckpointHeader* header,
                                       NBodySimProError* error);

/* Replaces n_body_sim_pro_checkpoint_peek (checkpoint.c:104-133): the header
 * only, after the magic/version check. */
NBodySimProStatus nbx_checkpoint_peek(const char* path, NBodySimProCheckpointHeader* header,
                                      NBodySimProError* error);

/* Replaces n_body_sim_pro_checkpoint_read (checkpoint.c:135-206): reads into
 * the caller's ten arrays (`capacity` doubles each); on success
 * view->particle_count is the stored count. A stored count above `capacity`
 * is INVALID_ARGUMENT, as in the reference. */
NBodySimProStatus nbx_checkpoint_read(const char* path, NBodySimProCheckpointHeader* header,
                                      NBodySimProParticleSystemView* view, size_t capacity,
                                      NBodySimProError* error);

/* Device-resident state to a checkpoint: downloads the engine's particles and
 * writes them with `meta` (its particle_count is taken from the engine). No
 * reference counterpart (the reference saves from host memory,
 * SimulationController.cpp:213-241). */
NBodySimProStatus nbx_engine_checkpoint_save(NBXEngine* engine, const char* path,
                                             const NBodySimProCheckpointHeader* meta,
                                             NBodySimProError* error);

/* A checkpoint into the engine's device state (resume); returns the stored
 * header. The engine's capacity must hold the stored count, and the file must
 * hold at least one particle (INVALID_ARGUMENT otherwise, as the reference's
 * read into a smaller particle system, checkpoint.c:159-164). No reference
 * counterpart (the reference loads into host memory,
 * SimulationController.cpp:242-272). */
NBodySimProStatus nbx_engine_checkpoint_load(NBXEngine* engine, const char* path,
                                             NBodySimProCheckpointHeader* header,
                                             NBodySimProError* error);

#ifdef __cplusplus
}
#endif

#endif /* NBX_CHECKPOINT_H */
