/*
 * nbo_random.c - oracle: xoshiro256** generator and the reference presets used
 * by the reference's own tests and benchmark (random_cloud, two_body,
 * globular_cluster, spiral_galaxy, galaxy_collision).
 *
 * TEST INFRASTRUCTURE ONLY (see nbo.h). Follows src/c/generation/random.c:9-58
 * and src/c/generation/presets.c:76-259, 272-288, 343-372, 387-468.
 */
#include "nbo.h"

#include <math.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* random.c:9-15 */
static uint64_t splitmix_next(uint64_t* s) {
    *s += 0x9E3779B97F4A7C15ull;
    uint64_t r = *s;
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
}

static uint64_t rotl(uint64_t v, int k) {
    return (v << k) | (v >> (64 - k));
}

void nbo_rng_init(nbo_rng* g, uint64_t seed) {
    uint64_t s = seed;
    for (int i = 0; i < 4; ++i) {
        g->s[i] = splitmix_next(&s);
    }
}

/* random.c:28-40 (xoshiro256**) */
uint64_t nbo_rng_next_u64(nbo_rng* g) {
    const uint64_t out = rotl(g->s[1] * 5, 7) * 9;
    const uint64_t t = g->s[1] << 17;
    g->s[2] ^= g->s[0];
    g->s[3] ^= g->s[1];
    g->s[1] ^= g->s[2];
    g->s[0] ^= g->s[3];
    g->s[2] ^= t;
    g->s[3] = rotl(g->s[3], 45);
    return out;
}

/* random.c:42-44 */
double nbo_rng_next_double(nbo_rng* g) {
    return (double)(nbo_rng_next_u64(g) >> 11) * (1.0 / 9007199254740992.0);
}

static double next_range(nbo_rng* g, double lo, double hi) {
    return lo + (hi - lo) * nbo_rng_next_double(g);
}

/* random.c:51-58 (Box-Muller, cosine branch) */
double nbo_rng_next_gaussian(nbo_rng* g) {
    double u1 = 0.0;
    while (u1 == 0.0) {
        u1 = nbo_rng_next_double(g);
    }
    const double u2 = nbo_rng_next_double(g);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

static void set_particle(NBodySimProParticleSystemView* v, size_t i, double x, double y,
                         double z, double vx, double vy, double vz, double m) {
    v->positions_x[i] = x;
    v->positions_y[i] = y;
    v->positions_z[i] = z;
    v->velocities_x[i] = vx;
    v->velocities_y[i] = vy;
    v->velocities_z[i] = vz;
    v->masses[i] = m;
}

/* presets.c:76-108 */
static void remove_mean_velocity(NBodySimProParticleSystemView* v, size_t start, size_t count) {
    double mx = 0.0, my = 0.0, mz = 0.0;
    for (size_t i = 0; i < count; ++i) {
        mx += v->velocities_x[start + i];
        my += v->velocities_y[start + i];
        mz += v->velocities_z[start + i];
    }
    mx /= (double)count;
    my /= (double)count;
    mz /= (double)count;
    for (size_t i = 0; i < count; ++i) {
        v->velocities_x[start + i] -= mx;
        v->velocities_y[start + i] -= my;
        v->velocities_z[start + i] -= mz;
    }
}

/* presets.c:113-127 */
static void random_direction(nbo_rng* g, double* x, double* y, double* z) {
    for (;;) {
        const double gx = nbo_rng_next_gaussian(g);
        const double gy = nbo_rng_next_gaussian(g);
        const double gz = nbo_rng_next_gaussian(g);
        const double norm = sqrt(gx * gx + gy * gy + gz * gz);
        if (norm > 1.0e-12) {
            *x = gx / norm;
            *y = gy / norm;
            *z = gz / norm;
            return;
        }
    }
}

/* presets.c:144-154 */
static double plummer_radius(nbo_rng* g, double a) {
    const double u = 1.0 - nbo_rng_next_double(g);
    return a / sqrt(pow(u, -2.0 / 3.0) - 1.0);
}

static double plummer_dispersion(double a, double r, double mass) {
    const double den = sqrt(a * a + r * r);
    return sqrt(1 * mass / (6.0 * den));
}

/* presets.c:173-202 */
static void plummer_sphere(nbo_rng* g, NBodySimProParticleSystemView* v, size_t start,
                           size_t count, double a, double mass) {
    const double pm = mass / (double)count;
    for (size_t i = 0; i < count; ++i) {
        const double r = plummer_radius(g, a);
        double dx, dy, dz;
        random_direction(g, &dx, &dy, &dz);
        const double sig = plummer_dispersion(a, r, mass);
        const double vx = nbo_rng_next_gaussian(g) * sig;
        const double vy = nbo_rng_next_gaussian(g) * sig;
        const double vz = nbo_rng_next_gaussian(g) * sig;
        set_particle(v, start + i, r * dx, r * dy, r * dz, vx, vy, vz, pm);
    }
    remove_mean_velocity(v, start, count);
}

/* presets.c:211-259 */
static void disk_galaxy(nbo_rng* g, NBodySimProParticleSystemView* v, size_t start, size_t count,
                        double rd, double hz, double vc, double sig, double twist, int arms,
                        const double* center, double pm) {
    for (size_t i = 0; i < count; ++i) {
        const double e1 = -log(1.0 - nbo_rng_next_double(g));
        const double e2 = -log(1.0 - nbo_rng_next_double(g));
        const double r = rd * (e1 + e2);
        double ang = 2.0 * M_PI * nbo_rng_next_double(g);
        if (twist > 0.0 && arms > 0) {
            const double ai = nbo_rng_next_double(g) * (double)arms;
            const double aa = 2.0 * M_PI * ai / (double)arms + twist * r / rd;
            ang = aa + nbo_rng_next_gaussian(g) * 0.15;
        }
        const double z = nbo_rng_next_gaussian(g) * hz;
        const double x = r * cos(ang);
        const double y = r * sin(ang);
        const double tx = -sin(ang);
        const double ty = cos(ang);
        const double vx = vc * tx + nbo_rng_next_gaussian(g) * sig;
        const double vy = vc * ty + nbo_rng_next_gaussian(g) * sig;
        const double vz = nbo_rng_next_gaussian(g) * sig * 0.5;
        double px = x, py = y, pz = z;
        if (center != NULL) {
            px += center[0];
            py += center[1];
            pz += center[2];
        }
        set_particle(v, start + i, px, py, pz, vx, vy, vz, pm);
    }
    remove_mean_velocity(v, start, count);
}

int nbo_preset_generate(NBodySimProParticleSystemView* v, int preset, uint64_t seed) {
    if (v == NULL || v->particle_count == 0) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    const size_t n = v->particle_count;
    nbo_rng g;
    nbo_rng_init(&g, seed);
    switch (preset) {
        case NBO_PRESET_TWO_BODY: { /* presets.c:261-270 */
            if (n != 2) {
                return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
            }
            const double s = sqrt(2.0) / 2.0;
            set_particle(v, 0, -0.5, 0.0, 0.0, 0.0, s, 0.0, 1.0);
            set_particle(v, 1, 0.5, 0.0, 0.0, 0.0, -s, 0.0, 1.0);
            return N_BODY_SIM_PRO_STATUS_OK;
        }
        case NBO_PRESET_RANDOM_CLOUD: { /* presets.c:272-288, 130-141 */
            const double pm = 1.0 / (double)n;
            for (size_t i = 0; i < n; ++i) {
                const double cr = cbrt(nbo_rng_next_double(&g));
                const double dist = 1.0 * cr;
                double dx, dy, dz;
                random_direction(&g, &dx, &dy, &dz);
                const double x = dist * dx, y = dist * dy, z = dist * dz;
                const double vx = next_range(&g, -0.05, 0.05);
                const double vy = next_range(&g, -0.05, 0.05);
                const double vz = next_range(&g, -0.05, 0.05);
                set_particle(v, i, x, y, z, vx, vy, vz, pm);
            }
            remove_mean_velocity(v, 0, n);
            return N_BODY_SIM_PRO_STATUS_OK;
        }
        case NBO_PRESET_GLOBULAR_CLUSTER: /* presets.c:437-440 */
            plummer_sphere(&g, v, 0, n, 1.0, 1.0);
            return N_BODY_SIM_PRO_STATUS_OK;
        case NBO_PRESET_SPIRAL_GALAXY: /* presets.c:445-450 */
            disk_galaxy(&g, v, 0, n, 0.7, 0.1, 1.0, 0.12, 3.5, 2, NULL, 1.0 / (double)n);
            return N_BODY_SIM_PRO_STATUS_OK;
        case NBO_PRESET_GALAXY_COLLISION: { /* presets.c:343-372 */
            const size_t half = n / 2;
            const double pm = 0.5 / (double)half;
            const double left[3] = {-2.5, 0.0, 0.0};
            const double right[3] = {2.5, 0.0, 0.0};
            disk_galaxy(&g, v, 0, half, 0.7, 0.1, 1.0, 0.12, 0.0, 0, left, pm);
            disk_galaxy(&g, v, half, half, 0.7, 0.1, 1.0, 0.12, 0.0, 0, right, pm);
            for (size_t i = 0; i < n; ++i) {
                v->velocities_x[i] += i < half ? 0.35 : -0.35;
            }
            return N_BODY_SIM_PRO_STATUS_OK;
        }
        default:
            return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
}
