/*
 * nbx_presets_named.c - the reference's nine named presets (include/nbx_presets.h).
 *
 * Two generators with the reference's exact sampling steps, so the same
 * preset and seed give the same arrays bit for bit (same libm):
 *   nbx_preset_generate           one sequential stream over the whole system
 *                                 (src/c/generation/presets.c:113-467);
 *   nbx_preset_generate_parallel  one stream per particle, seeded from
 *                                 (seed, index), split over host threads
 *                                 (src/c/generation/presets_parallel.c:36-482).
 * Each preset is a list of blocks (a galaxy, a cluster) drawn with one
 * per-particle sampler; a block's mean velocity is removed after it is drawn,
 * summed in index order. That order is also what makes the parallel variant
 * independent of the thread count; the reference sums random_cloud's means
 * with an OpenMP reduction (presets_parallel.c:225), which equals this at one
 * thread. Particles no block covers (the remainder of galaxy_collision's
 * halves and triple_galaxy's thirds) are left unset by the reference; here
 * they are zero-mass particles at rest at the origin, plus the collision's
 * approach velocity shift that the reference applies to every index.
 */
#include <stdio.h>
#include <string.h>

#include "nbx_host_internal.h"
#include "nbx_presets.h"

/* include/n_body_sim_pro/generation/presets.h:27-38, presets.c:27-51 */
static const char* const kPresetNames[NBX_PRESET_COUNT] = {
    "two_body",       "random_cloud",      "solar_system",
    "open_cluster",   "globular_cluster",  "spiral_galaxy",
    "elliptical_galaxy", "galaxy_collision", "triple_galaxy"};

const char* nbx_preset_name(int preset) {
    return preset >= 0 && preset < NBX_PRESET_COUNT ? kPresetNames[preset] : "unknown";
}

int nbx_preset_from_name(const char* name) {
    if (name == NULL) {
        return -1;
    }
    for (int p = 0; p < NBX_PRESET_COUNT; ++p) {
        if (strcmp(name, kPresetNames[p]) == 0) {
            return p;
        }
    }
    return -1;
}

static NBodySimProStatus fail(NBodySimProError* error, NBodySimProStatus status, int line,
                              const char* message) {
    if (error != NULL) {
        error->status = status;
        error->source_file = __FILE__;
        error->source_line = line;
        snprintf(error->message, sizeof(error->message), "%s", message);
    }
    return status;
}

#define FAIL(st, msg) return fail(error, (st), __LINE__, (msg))

typedef enum Sampler { CLOUD_RADIUS_FIRST, CLOUD_DIRECTION_FIRST, GAUSSIAN_BLOB, PLUMMER, DISK } Sampler;

typedef struct Block {
    Sampler sampler;
    size_t start, count;
    double mass;           /* per particle */
    /* PLUMMER: scale radius, the a^2 of the dispersion, total mass */
    double a, a2, total;
    /* DISK: scale radius, scale height, circular speed, dispersion, arm twist */
    double scale, height, speed, sigma, twist;
    int arms;
    int centred;
    double centre[3];
    uint64_t stream;       /* parallel variant: streams from (stream, index) */
} Block;

/*
 * One particle of a block from generator g. The order of the draws is the
 * reference's (presets.c:130-141, 144-154, 157-161, 170-257, 272-308;
 * presets_parallel.c:107-142, 174-210, 226-259, 285-301).
 */
static void sample(const Block* b, nbx_rng* g, double p[3], double v[3]) {
    double d[3];
    switch (b->sampler) {
        case CLOUD_RADIUS_FIRST:
        case CLOUD_DIRECTION_FIRST: {
            /* uniform in the unit ball, U(-0.05, 0.05) velocities */
            double r;
            if (b->sampler == CLOUD_RADIUS_FIRST) {
                r = 1.0 * cbrt(nbx_rng_unit(g));
                nbx_rng_direction(g, d);
            } else {
                nbx_rng_direction(g, d);
                r = cbrt(nbx_rng_unit(g));
            }
            for (int k = 0; k < 3; ++k) p[k] = r * d[k];
            for (int k = 0; k < 3; ++k) v[k] = -0.05 + (0.05 - -0.05) * nbx_rng_unit(g);
            break;
        }
        case GAUSSIAN_BLOB:
            /* open cluster: isotropic Gaussians, sigma_x 0.3, sigma_v 0.02 */
            for (int k = 0; k < 3; ++k) p[k] = nbx_rng_gauss(g) * 0.3;
            for (int k = 0; k < 3; ++k) v[k] = nbx_rng_gauss(g) * 0.02;
            break;
        case PLUMMER: {
            /* r = a / sqrt(u^(-2/3) - 1), u uniform in (0, 1]; Gaussian velocities
             * with sigma^2 = G M / (6 sqrt(a^2 + r^2)), G = 1 */
            const double u = 1.0 - nbx_rng_unit(g);
            const double r = b->a / sqrt(pow(u, -2.0 / 3.0) - 1.0);
            nbx_rng_direction(g, d);
            for (int k = 0; k < 3; ++k) p[k] = r * d[k];
            const double s = sqrt(b->total / (6.0 * sqrt(b->a2 + r * r)));
            for (int k = 0; k < 3; ++k) v[k] = nbx_rng_gauss(g) * s;
            break;
        }
        case DISK: {
            /* radius ~ Gamma(2, scale); angle uniform or around `arms` arms
             * wound by twist * r / scale; height Gaussian; circular speed
             * plus Gaussian dispersion (half of it vertically) */
            const double e1 = -log(1.0 - nbx_rng_unit(g));
            const double e2 = -log(1.0 - nbx_rng_unit(g));
            const double r = b->scale * (e1 + e2);
            double angle = 2.0 * M_PI * nbx_rng_unit(g);
            if (b->twist > 0.0 && b->arms > 0) {
                const double arm = nbx_rng_unit(g) * (double)b->arms;
                const double arm_angle =
                    2.0 * M_PI * arm / (double)b->arms + b->twist * r / b->scale;
                angle = arm_angle + nbx_rng_gauss(g) * 0.15;
            }
            const double h = nbx_rng_gauss(g) * b->height;
            p[0] = r * cos(angle);
            p[1] = r * sin(angle);
            p[2] = h;
            const double tx = -sin(angle), ty = cos(angle);
            v[0] = b->speed * tx + nbx_rng_gauss(g) * b->sigma;
            v[1] = b->speed * ty + nbx_rng_gauss(g) * b->sigma;
            v[2] = nbx_rng_gauss(g) * b->sigma * 0.5;
            break;
        }
    }
    if (b->centred) {
        for (int k = 0; k < 3; ++k) p[k] += b->centre[k];
    }
}

static void store(NBodySimProParticleSystemView* view, size_t i, const double p[3],
                  const double v[3], double m) {
    view->positions_x[i] = p[0];
    view->positions_y[i] = p[1];
    view->positions_z[i] = p[2];
    view->velocities_x[i] = v[0];
    view->velocities_y[i] = v[1];
    view->velocities_z[i] = v[2];
    view->masses[i] = m;
}

/* presets.c:80-111: subtract the block's mean velocity (equal masses). */
static void remove_mean_velocity(NBodySimProParticleSystemView* view, size_t start, size_t count) {
    if (count == 0) {
        return;
    }
    double mean[3] = {0.0, 0.0, 0.0};
    for (size_t i = start; i < start + count; ++i) {
        mean[0] += view->velocities_x[i];
        mean[1] += view->velocities_y[i];
        mean[2] += view->velocities_z[i];
    }
    for (int k = 0; k < 3; ++k) mean[k] /= (double)count;
    for (size_t i = start; i < start + count; ++i) {
        view->velocities_x[i] -= mean[0];
        view->velocities_y[i] -= mean[1];
        view->velocities_z[i] -= mean[2];
    }
}

static Block disk_block(size_t start, size_t count, double mass, double twist, int arms) {
    /* spiral_galaxy and each galaxy_collision disk: scale 0.7, height 0.1,
     * circular speed 1, dispersion 0.12 (presets.c:348-356, 453-456) */
    Block b = {DISK, start, count, mass, 0, 0, 0, 0.7, 0.1, 1.0, 0.12, twist, arms, 0, {0, 0, 0}, 0};
    return b;
}

static Block plummer_block(size_t start, size_t count, double a, double a2, double total) {
    Block b = {PLUMMER, start, count, total / (double)count, a, a2, total, 0, 0, 0, 0, 0, 0, 0,
               {0, 0, 0}, 0};
    return b;
}

static void centre(Block* b, double x, double y, double z) {
    b->centred = 1;
    b->centre[0] = x;
    b->centre[1] = y;
    b->centre[2] = z;
}

/* The blocks of a sampled preset (everything but two_body and solar_system). */
static int preset_blocks(int preset, size_t n, uint64_t seed, int parallel, Block out[3]) {
    const Block none = {GAUSSIAN_BLOB, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, {0, 0, 0}, 0};
    out[0] = out[1] = out[2] = none;
    switch (preset) {
        case NBX_PRESET_RANDOM_CLOUD:
            out[0] = none;
            out[0].sampler = parallel ? CLOUD_DIRECTION_FIRST : CLOUD_RADIUS_FIRST;
            out[0].count = n;
            out[0].mass = 1.0 / (double)n;
            out[0].stream = seed;
            return 1;
        case NBX_PRESET_OPEN_CLUSTER:
            out[0].count = n;
            out[0].mass = 1.0 / (double)n;
            out[0].stream = seed;
            return 1;
        case NBX_PRESET_GLOBULAR_CLUSTER:
        case NBX_PRESET_ELLIPTICAL_GALAXY: {
            const double a = preset == NBX_PRESET_GLOBULAR_CLUSTER ? 1.0 : 2.0;
            out[0] = plummer_block(0, n, a, a * a, 1.0);
            out[0].stream = seed;
            return 1;
        }
        case NBX_PRESET_SPIRAL_GALAXY:
            out[0] = disk_block(0, n, 1.0 / (double)n, 3.5, 2);
            out[0].stream = seed;
            return 1;
        case NBX_PRESET_GALAXY_COLLISION: {
            /* two disks of mass 1/2; the sequential generator centres them at
             * x = -2.5 and +2.5 (presets.c:343-372); the parallel one leaves
             * both at the origin (presets_parallel.c:443-455) */
            const size_t half = n / 2;
            const double m = 0.5 / (double)half;
            out[0] = disk_block(0, half, m, 0.0, 0);
            out[1] = disk_block(half, half, m, 0.0, 0);
            if (parallel) {
                out[0].stream = nbx_mix64(seed ^ (0 + 0x9E3779B97F4A7C15ull));
                out[1].stream = nbx_mix64(seed ^ (1 + 0x9E3779B97F4A7C15ull));
            } else {
                centre(&out[0], -2.5, 0.0, 0.0);
                centre(&out[1], 2.5, 0.0, 0.0);
            }
            return 2;
        }
        case NBX_PRESET_TRIPLE_GALAXY: {
            /* three Plummer spheres of mass 1/3, a = 0.7, at (-1.8,0,0),
             * (1.8,0,0), (0,1.8,0) (presets.c:374-385; presets_parallel.c:457-477
             * writes a^2 as the literal 0.49) */
            const size_t third = n / 3;
            const double c[3][3] = {{-1.8, 0.0, 0.0}, {1.8, 0.0, 0.0}, {0.0, 1.8, 0.0}};
            for (int k = 0; k < 3; ++k) {
                out[k] = plummer_block((size_t)k * third, third, 0.7, parallel ? 0.49 : 0.7 * 0.7,
                                       1.0 / 3.0);
                centre(&out[k], c[k][0], c[k][1], c[k][2]);
                out[k].stream = nbx_mix64(seed ^ ((uint64_t)k + 0x9E3779B97F4A7C15ull));
            }
            return 3;
        }
        default:
            return 0;
    }
}

/* presets.c:261-270 */
static void two_body(NBodySimProParticleSystemView* view) {
    const double s = sqrt(2.0) / 2.0;
    const double p0[3] = {-0.5, 0.0, 0.0}, v0[3] = {0.0, s, 0.0};
    const double p1[3] = {0.5, 0.0, 0.0}, v1[3] = {0.0, -s, 0.0};
    store(view, 0, p0, v0, 1.0);
    store(view, 1, p1, v1, 1.0);
}

/* presets.c:310-341: a unit-mass star at rest and n-1 planets of mass 1e-3
 * on circular orbits at radii 1 + 3 i / (n-1), random phases. Both variants
 * draw the phases from one stream. */
static void solar_system(NBodySimProParticleSystemView* view, uint64_t seed) {
    const size_t planets = view->particle_count - 1;
    nbx_rng g;
    nbx_rng_seed(&g, seed);
    const double zero[3] = {0.0, 0.0, 0.0};
    store(view, 0, zero, zero, 1.0);
    for (size_t i = 0; i < planets; ++i) {
        const double r = 1.0 + 3.0 * (double)i / (double)planets;
        const double vc = sqrt(1.0 / r);
        const double angle = 2.0 * M_PI * nbx_rng_unit(&g);
        const double p[3] = {r * cos(angle), r * sin(angle), 0.0};
        const double v[3] = {-vc * sin(angle), vc * cos(angle), 0.0};
        store(view, i + 1, p, v, 0.001);
    }
}

typedef struct ParallelBlock {
    const Block* b;
    NBodySimProParticleSystemView* view;
} ParallelBlock;

static void parallel_range(void* ctx, size_t lo, size_t hi) {
    const ParallelBlock* j = (const ParallelBlock*)ctx;
    const Block* b = j->b;
    for (size_t k = lo; k < hi; ++k) {
        const size_t i = b->start + k;
        /* presets_parallel.c:36-42: stream seed = mix64(seed ^ (index + golden)) */
        nbx_rng g;
        nbx_rng_seed(&g, nbx_mix64(b->stream ^ ((uint64_t)i + 0x9E3779B97F4A7C15ull)));
        double p[3], v[3];
        sample(b, &g, p, v);
        store(j->view, i, p, v, b->mass);
    }
}

static NBodySimProStatus generate(NBodySimProParticleSystemView* view, int preset, uint64_t seed,
                                  int parallel, NBodySimProError* error) {
    if (view == NULL || !view->positions_x || !view->positions_y || !view->positions_z ||
        !view->velocities_x || !view->velocities_y || !view->velocities_z || !view->masses) {
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "particle arrays must not be null");
    }
    const size_t n = view->particle_count;
    if (n == 0) {
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "particle count must be non-zero");
    }
    if (preset < 0 || preset >= NBX_PRESET_COUNT) {
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "unknown preset");
    }
    if (preset == NBX_PRESET_TWO_BODY) {
        if (n != 2) {
            FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "two_body preset requires exactly 2 particles");
        }
        two_body(view);
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    if (preset == NBX_PRESET_SOLAR_SYSTEM) {
        if (n < 2 && !parallel) {
            /* the reference reports this as a failed generation (presets.c:311-316, 460-464) */
            FAIL(N_BODY_SIM_PRO_STATUS_INVALID_STATE,
                 "preset generation failed: solar system preset needs at least one planet");
        }
        solar_system(view, seed);
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    Block blocks[3];
    const int count = preset_blocks(preset, n, seed, parallel, blocks);
    /* particles outside every block: zero-mass, at rest at the origin */
    size_t covered = 0;
    for (int k = 0; k < count; ++k) covered += blocks[k].count;
    const double zero[3] = {0.0, 0.0, 0.0};
    for (size_t i = covered; i < n; ++i) store(view, i, zero, zero, 0.0);
    if (parallel) {
        for (int k = 0; k < count; ++k) {
            ParallelBlock job = {&blocks[k], view};
            nbx_parallel_ranges(blocks[k].count, parallel_range, &job);
            remove_mean_velocity(view, blocks[k].start, blocks[k].count);
        }
    } else {
        nbx_rng g;
        nbx_rng_seed(&g, seed);
        for (int k = 0; k < count; ++k) {
            for (size_t i = blocks[k].start; i < blocks[k].start + blocks[k].count; ++i) {
                double p[3], v[3];
                sample(&blocks[k], &g, p, v);
                store(view, i, p, v, blocks[k].mass);
            }
            remove_mean_velocity(view, blocks[k].start, blocks[k].count);
        }
    }
    if (preset == NBX_PRESET_GALAXY_COLLISION) {
        /* approach speed 0.35 each, first half towards +x (presets.c:362-371) */
        const size_t half = n / 2;
        for (size_t i = 0; i < n; ++i) view->velocities_x[i] += i < half ? 0.35 : -0.35;
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_preset_generate(NBodySimProParticleSystemView* view, int preset,
                                      uint64_t seed, NBodySimProError* error) {
    return generate(view, preset, seed, 0, error);
}

NBodySimProStatus nbx_preset_generate_parallel(NBodySimProParticleSystemView* view, int preset,
                                               uint64_t seed, NBodySimProError* error) {
    return generate(view, preset, seed, 1, error);
}
