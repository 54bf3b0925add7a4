/*
 * nbx_presets.h - host-side (C17) initial-condition generators in libnbx.so.
 *
 * They write positions, velocities and masses of an existing view (the caller
 * owns the arrays; accelerations are left untouched). All are deterministic
 * for a given seed; the same arrays then feed both the CPU and GPU paths.
 * Generators with a stream per particle run on the host worker threads
 * (nbx_set_host_threads) with the same output for any thread count.
 */
#ifndef NBX_PRESETS_H
#define NBX_PRESETS_H

#include "nbx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The reference's random_cloud preset, bit-identical for the same seed
 * (src/c/generation/presets.c:272-288; the force benchmark's default input,
 * benchmarks/benchmark_force.c:164-165). */
NBodySimProStatus nbx_preset_random_cloud(NBodySimProParticleSystemView* view, uint64_t seed);

/*
 * Plummer sphere (scale radius a, total mass M, G = 1, equal masses) with
 * radii from the cumulative mass profile and isotropic velocities sampled
 * from the Plummer distribution function (speed fraction q = v/v_esc by
 * rejection on g(q) = q^2 (1 - q^2)^(7/2)). Mass fractions above 0.999 are
 * resampled (r < ~38.7 a). Positions and velocities are shifted so the
 * centre of mass sits at rest at the origin. Each particle draws from its
 * own xoshiro256** stream seeded from (seed, index).
 */
NBodySimProStatus nbx_generate_plummer(NBodySimProParticleSystemView* view, double scale_radius,
                                       double total_mass, uint64_t seed);

typedef struct NBXGalaxyPairParams {
    double disk_scale_length;  /* exponential disk scale length (1.0) */
    double disk_scale_height;  /* sech^2 vertical scale height (0.1) */
    double bulge_fraction;     /* share of each galaxy's particles/mass in the bulge (0.1) */
    double bulge_scale;        /* Plummer bulge scale radius (0.2) */
    double separation;         /* initial centre separation (20.0) */
    double pericenter;         /* pericentre of the parabolic relative orbit (2.0) */
    double inclination;        /* tilt of the second disk about the x axis, radians (pi/4) */
    double velocity_dispersion;/* disk velocity dispersion, fraction of v_circ (0.1) */
} NBXGalaxyPairParams;

/* Defaults listed above. */
void nbx_galaxy_pair_defaults(NBXGalaxyPairParams* params);

/*
 * Two equal-mass galaxies (total mass 1, equal particle masses), each an
 * exponential disk plus a Plummer bulge, on a parabolic (zero-energy)
 * relative orbit starting `separation` apart, in the centre-of-mass frame.
 * `params` may be NULL for the defaults.
 */
NBodySimProStatus nbx_generate_two_galaxy(NBodySimProParticleSystemView* view,
                                          const NBXGalaxyPairParams* params, uint64_t seed);

/* One exponential disk + bulge galaxy of total mass 1 at the origin. */
NBodySimProStatus nbx_generate_disk_galaxy(NBodySimProParticleSystemView* view,
                                           const NBXGalaxyPairParams* params, uint64_t seed);

#ifndef N_BODY_SIM_PRO_GENERATION_PRESETS_H
/* include/n_body_sim_pro/generation/presets.h:27-38 */
typedef enum NBodySimProSimulationPreset {
    N_BODY_SIM_PRO_PRESET_TWO_BODY,
    N_BODY_SIM_PRO_PRESET_RANDOM_CLOUD,
    N_BODY_SIM_PRO_PRESET_SOLAR_SYSTEM,
    N_BODY_SIM_PRO_PRESET_OPEN_CLUSTER,
    N_BODY_SIM_PRO_PRESET_GLOBULAR_CLUSTER,
    N_BODY_SIM_PRO_PRESET_SPIRAL_GALAXY,
    N_BODY_SIM_PRO_PRESET_ELLIPTICAL_GALAXY,
    N_BODY_SIM_PRO_PRESET_GALAXY_COLLISION,
    N_BODY_SIM_PRO_PRESET_TRIPLE_GALAXY,
    N_BODY_SIM_PRO_PRESET_COUNT
} NBodySimProSimulationPreset;
#endif

enum {
    NBX_PRESET_TWO_BODY = N_BODY_SIM_PRO_PRESET_TWO_BODY,
    NBX_PRESET_RANDOM_CLOUD = N_BODY_SIM_PRO_PRESET_RANDOM_CLOUD,
    NBX_PRESET_SOLAR_SYSTEM = N_BODY_SIM_PRO_PRESET_SOLAR_SYSTEM,
    NBX_PRESET_OPEN_CLUSTER = N_BODY_SIM_PRO_PRESET_OPEN_CLUSTER,
    NBX_PRESET_GLOBULAR_CLUSTER = N_BODY_SIM_PRO_PRESET_GLOBULAR_CLUSTER,
    NBX_PRESET_SPIRAL_GALAXY = N_BODY_SIM_PRO_PRESET_SPIRAL_GALAXY,
    NBX_PRESET_ELLIPTICAL_GALAXY = N_BODY_SIM_PRO_PRESET_ELLIPTICAL_GALAXY,
    NBX_PRESET_GALAXY_COLLISION = N_BODY_SIM_PRO_PRESET_GALAXY_COLLISION,
    NBX_PRESET_TRIPLE_GALAXY = N_BODY_SIM_PRO_PRESET_TRIPLE_GALAXY,
    NBX_PRESET_COUNT = N_BODY_SIM_PRO_PRESET_COUNT
};

/* Replaces n_body_sim_pro_preset_string (presets.c:27-51); "unknown" out of range. */
const char* nbx_preset_name(int preset);
/* The preset whose name is `name`, or -1 (the CLI's lookup, main.cpp:136-148). */
int nbx_preset_from_name(const char* name);

/*
 * Replaces n_body_sim_pro_preset_generate (presets.c:387-467): the nine named
 * presets from one sequential stream, bit-identical to the reference for the
 * same seed. Writes view->particle_count particles (positions, velocities,
 * masses; accelerations untouched). Errors as the reference: a zero count,
 * two_body with a count other than 2, an unknown preset -> INVALID_ARGUMENT;
 * solar_system with no planet -> INVALID_STATE. Where the reference leaves
 * particles unset (the remainders of galaxy_collision's n/2 halves and
 * triple_galaxy's n/3 thirds) they are zero-mass particles at the origin.
 */
NBodySimProStatus nbx_preset_generate(NBodySimProParticleSystemView* view, int preset,
                                      uint64_t seed, NBodySimProError* error);

/*
 * Replaces n_body_sim_pro_preset_generate_parallel (presets_parallel.c:286-482):
 * a stream per particle seeded from (seed, index), over the host worker
 * threads; the same output for any thread count, bit-identical to the
 * reference's run on one thread.
 */
NBodySimProStatus nbx_preset_generate_parallel(NBodySimProParticleSystemView* view, int preset,
                                               uint64_t seed, NBodySimProError* error);

/* Host worker threads of the generators (the reference's
 * n_body_sim_pro_threading_set_thread_count, threading.c:15-25); 0 restores
 * the default, all online CPUs. Not to be changed while a generator runs. */
void nbx_set_host_threads(int threads);
int nbx_host_threads(void);

/* Cluster sizes log-uniform in [min_size, max_size] (inclusive), seeded. */
NBodySimProStatus nbx_ensemble_sizes(size_t clusters, uint32_t min_size, uint32_t max_size,
                                     uint64_t seed, uint32_t* sizes_out);

#ifdef __cplusplus
}
#endif

#endif /* NBX_PRESETS_H */
