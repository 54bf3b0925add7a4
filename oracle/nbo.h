/*
 * nbo.h - CPU oracle for the gravitational force path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP engine
 * is compared against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; nothing in the product links or calls it.
 *
 * Each function restates the reference engine's arithmetic operation for
 * operation (same association order, contraction off: built with
 * -std=c17 -ffp-contract=off), citing the reference file:line it follows.
 *
 * Parity pins (what checks each function beyond re-typing its formula):
 *   nbo_direct_fp64 ........ bit-identical to the reference build in
 *                            oracle/_ref on committed fixtures; two-body KAT
 *                            a=(+-1,0,0) at eps=0; N=1 self force = 0;
 *                            long-double brute force on tiny N.
 *   nbo_direct_openmp ...... bit-identical to nbo_direct_fp64.
 *   nbo_bh_* ............... keys/permutation/accelerations/stats bit-identical
 *                            to oracle/_ref (400-particle cloud, seed 7:
 *                            154,231 / 107,700 / 37,920 interactions at
 *                            theta 0.1 / 0.3 / 0.7); the reference tests'
 *                            two-body and theta trade-off assertions;
 *                            theta -> 0 reduces to a direct sum over leaves;
 *                            stable_sort((key, index)) for the permutation.
 *   nbo_leapfrog_advance ... bit-identical to oracle/_ref; two-body circular
 *                            orbit conservation bounds.
 *   nbo_diagnostics_* ...... closed forms (KE 2.5, P (2,3,0), L (-18,6,8),
 *                            COM (14/5,19/5,24/5); U=-1 for a unit pair).
 *   nbo_random_* / presets . splitmix64 / xoshiro256** published outputs;
 *                            bit-identical arrays to oracle/_ref.
 */
#ifndef NBO_H
#define NBO_H

#include "nbx_types.h"

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- all-pairs (src/c/physics/gravity.c, gravity_openmp.c) ---- */
int nbo_direct_fp64(const NBodySimProParticleSystemView* view, const NBodySimProGravity* gravity);
int nbo_direct_openmp(const NBodySimProParticleSystemView* view, const NBodySimProGravity* gravity);
/* Same sum restricted to targets [first, first+count): used for subsampled error checks. */
int nbo_direct_openmp_subset(const NBodySimProParticleSystemView* view,
                             const NBodySimProGravity* gravity, const int64_t* targets,
                             size_t target_count, double* out_xyz);

/* ---- integrator (src/c/physics/integrator.c:47-85) ---- */
typedef int (*nbo_force_fn)(const NBodySimProParticleSystemView*, const NBodySimProGravity*);
int nbo_leapfrog_advance(NBodySimProParticleSystemView* view, const NBodySimProGravity* gravity,
                         double timestep, nbo_force_fn force);

/* ---- diagnostics (src/c/diagnostics/numerics.c) ---- */
int nbo_diagnostics_global(const NBodySimProParticleSystemView* view,
                           NBodySimProDiagnosticsQuantities* q);
int nbo_potential_energy(const NBodySimProParticleSystemView* view,
                         const NBodySimProGravity* gravity, double* out);

/* ---- Barnes-Hut (src/c/barnes_hut/barnes_hut.c) ---- */
typedef struct nbo_bh nbo_bh;
nbo_bh* nbo_bh_create(void);
void nbo_bh_destroy(nbo_bh* tree);
void nbo_bh_set_theta(nbo_bh* tree, double theta);
/* Build + walk + scatter, as n_body_sim_pro_barnes_hut_compute_acceleration. */
int nbo_bh_compute(nbo_bh* tree, const NBodySimProParticleSystemView* view,
                   const NBodySimProGravity* gravity);
int nbo_bh_build(nbo_bh* tree, const NBodySimProParticleSystemView* view);
void nbo_bh_stats(const nbo_bh* tree, NBodySimProBarnesHutStats* stats);
/* Root cube of the last build: centre xyz and half size. */
void nbo_bh_root(const nbo_bh* tree, double out[4]);
/* Keys of the last build, in ORIGINAL particle order, and the sorted permutation. */
void nbo_bh_keys(const nbo_bh* tree, uint64_t* keys_out, int64_t* permutation_out);
size_t nbo_bh_node_count(const nbo_bh* tree);
/*
 * One row per node (in any order): depth, centre xyz, half, particle count,
 * total mass, COM xyz, leaf particle (sorted rank, -1 for internal) -> 11 doubles.
 */
void nbo_bh_export_nodes(const nbo_bh* tree, double* rows);

/* ---- random numbers and presets (src/c/generation/random.c, presets.c) ---- */
typedef struct nbo_rng {
    uint64_t s[4];
} nbo_rng;
void nbo_rng_init(nbo_rng* rng, uint64_t seed);
uint64_t nbo_rng_next_u64(nbo_rng* rng);
double nbo_rng_next_double(nbo_rng* rng);
double nbo_rng_next_gaussian(nbo_rng* rng);

enum {
    NBO_PRESET_TWO_BODY = 0,
    NBO_PRESET_RANDOM_CLOUD = 1,
    NBO_PRESET_GLOBULAR_CLUSTER = 4,
    NBO_PRESET_SPIRAL_GALAXY = 5,
    NBO_PRESET_GALAXY_COLLISION = 7
};
/* Fills pos/vel/mass of `view` (count = view->particle_count). Returns 0 on success. */
int nbo_preset_generate(NBodySimProParticleSystemView* view, int preset, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* NBO_H */
