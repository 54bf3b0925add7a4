/*
 * nbx_presets.c - C17 initial-condition generators (include/nbx_presets.h).
 *
 * The random number generator is xoshiro256** seeded through splitmix64, the
 * published algorithms the reference also uses (src/c/generation/random.c:9-58),
 * so nbx_preset_random_cloud reproduces the reference preset bit for bit.
 * The Plummer-DF and galaxy-pair generators are new: the reference's
 * globular_cluster draws Gaussian velocities with the local dispersion
 * (presets.c:149-154, 190-194) and its galaxy_collision has no bulge
 * (presets.c:343-372).
 */
#include "nbx_presets.h"

#include <pthread.h>
#include <stdlib.h>
#include <unistd.h>

#include "nbx_host_internal.h"

typedef nbx_rng rng;

/* Per-particle stream: seeded from (seed, index). */
static void rng_particle(rng* g, uint64_t seed, uint64_t index) {
    uint64_t x = seed ^ (0xD1B54A32D192ED03ull * (index + 1));
    nbx_rng_seed(g, nbx_splitmix(&x));
}

static int view_ok(const NBodySimProParticleSystemView* v) {
    return v != NULL && v->particle_count > 0 && v->positions_x && v->positions_y &&
           v->positions_z && v->velocities_x && v->velocities_y && v->velocities_z && v->masses;
}

static void put(NBodySimProParticleSystemView* v, size_t i, const double p[3], const double u[3],
                double m) {
    v->positions_x[i] = p[0];
    v->positions_y[i] = p[1];
    v->positions_z[i] = p[2];
    v->velocities_x[i] = u[0];
    v->velocities_y[i] = u[1];
    v->velocities_z[i] = u[2];
    v->masses[i] = m;
}

/* Host worker threads for the generators; 0 means all online CPUs. */
static int host_threads = 0;

void nbx_set_host_threads(int threads) {
    host_threads = threads > 0 ? threads : 0;
}

int nbx_host_threads(void) {
    if (host_threads > 0) {
        return host_threads;
    }
    const long c = sysconf(_SC_NPROCESSORS_ONLN);
    return c > 0 ? (int)(c < 256 ? c : 256) : 1;
}

typedef struct RangeJob {
    nbx_range_fn fn;
    void* ctx;
    size_t lo, hi;
} RangeJob;

static void* range_thread(void* arg) {
    RangeJob* j = (RangeJob*)arg;
    j->fn(j->ctx, j->lo, j->hi);
    return NULL;
}

void nbx_parallel_ranges(size_t n, nbx_range_fn fn, void* ctx) {
    enum { kMinRange = 8192, kMaxThreads = 256 };
    size_t t = (size_t)nbx_host_threads();
    if (t > n / kMinRange) t = n / kMinRange;
    if (t > kMaxThreads) t = kMaxThreads;
    if (t <= 1) {
        fn(ctx, 0, n);
        return;
    }
    pthread_t tid[kMaxThreads];
    RangeJob job[kMaxThreads];
    int started[kMaxThreads];
    for (size_t k = 0; k < t; ++k) {
        job[k] = (RangeJob){fn, ctx, n * k / t, n * (k + 1) / t};
    }
    for (size_t k = 1; k < t; ++k) {
        started[k] = pthread_create(&tid[k], NULL, range_thread, &job[k]) == 0;
        if (!started[k]) {
            range_thread(&job[k]);   /* no thread: do the range here */
        }
    }
    range_thread(&job[0]);
    for (size_t k = 1; k < t; ++k) {
        if (started[k]) {
            pthread_join(tid[k], NULL);
        }
    }
}

NBodySimProStatus nbx_preset_random_cloud(NBodySimProParticleSystemView* v, uint64_t seed) {
    if (!view_ok(v)) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    const size_t n = v->particle_count;
    rng g;
    nbx_rng_seed(&g, seed);
    const double pm = 1.0 / (double)n;
    for (size_t i = 0; i < n; ++i) {
        /* uniform point in the unit ball, then U(-0.05, 0.05) velocities */
        const double dist = 1.0 * cbrt(nbx_rng_unit(&g));
        double d[3];
        nbx_rng_direction(&g, d);
        const double p[3] = {dist * d[0], dist * d[1], dist * d[2]};
        double u[3];
        for (int k = 0; k < 3; ++k) {
            u[k] = -0.05 + (0.05 - -0.05) * nbx_rng_unit(&g);
        }
        put(v, i, p, u, pm);
    }
    double mean[3] = {0.0, 0.0, 0.0};
    for (size_t i = 0; i < n; ++i) {
        mean[0] += v->velocities_x[i];
        mean[1] += v->velocities_y[i];
        mean[2] += v->velocities_z[i];
    }
    for (int k = 0; k < 3; ++k) {
        mean[k] /= (double)n;
    }
    for (size_t i = 0; i < n; ++i) {
        v->velocities_x[i] -= mean[0];
        v->velocities_y[i] -= mean[1];
        v->velocities_z[i] -= mean[2];
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

/* Shift [start, start+count) so its mass-weighted mean position and velocity vanish. */
static void recentre(NBodySimProParticleSystemView* v, size_t start, size_t count) {
    double mp[3] = {0, 0, 0}, mv[3] = {0, 0, 0}, mt = 0.0;
    for (size_t i = start; i < start + count; ++i) {
        const double m = v->masses[i];
        mt += m;
        mp[0] += m * v->positions_x[i];
        mp[1] += m * v->positions_y[i];
        mp[2] += m * v->positions_z[i];
        mv[0] += m * v->velocities_x[i];
        mv[1] += m * v->velocities_y[i];
        mv[2] += m * v->velocities_z[i];
    }
    if (mt == 0.0) {
        return;
    }
    for (int k = 0; k < 3; ++k) {
        mp[k] /= mt;
        mv[k] /= mt;
    }
    for (size_t i = start; i < start + count; ++i) {
        v->positions_x[i] -= mp[0];
        v->positions_y[i] -= mp[1];
        v->positions_z[i] -= mp[2];
        v->velocities_x[i] -= mv[0];
        v->velocities_y[i] -= mv[1];
        v->velocities_z[i] -= mv[2];
    }
}

/*
 * One Plummer particle (Aarseth, Henon & Wielen 1974): r from the inverse
 * cumulative mass M(<r)/M = r^3/(r^2+a^2)^{3/2}; speed q*v_esc with q drawn
 * by rejection from g(q) = q^2 (1-q^2)^{7/2} (max 0.0921 < 0.1).
 */
static void plummer_particle(rng* g, double a, double mass, double p[3], double u[3]) {
    double x;
    do {
        x = nbx_rng_unit(g);
    } while (x <= 1.0e-10 || x > 0.999);
    const double r = a / sqrt(pow(x, -2.0 / 3.0) - 1.0);
    double d[3];
    nbx_rng_direction(g, d);
    p[0] = r * d[0];
    p[1] = r * d[1];
    p[2] = r * d[2];
    double q;
    for (;;) {
        q = nbx_rng_unit(g);
        const double y = 0.1 * nbx_rng_unit(g);
        if (y < q * q * pow(1.0 - q * q, 3.5)) {
            break;
        }
    }
    const double vesc = sqrt(2.0 * mass / sqrt(r * r + a * a));
    const double speed = q * vesc;
    nbx_rng_direction(g, d);
    u[0] = speed * d[0];
    u[1] = speed * d[1];
    u[2] = speed * d[2];
}

typedef struct PlummerJob {
    NBodySimProParticleSystemView* v;
    double a, mass, pm;
    uint64_t seed;
} PlummerJob;

static void plummer_range(void* ctx, size_t lo, size_t hi) {
    const PlummerJob* j = (const PlummerJob*)ctx;
    for (size_t i = lo; i < hi; ++i) {
        rng g;
        rng_particle(&g, j->seed, i);
        double p[3], u[3];
        plummer_particle(&g, j->a, j->mass, p, u);
        put(j->v, i, p, u, j->pm);
    }
}

NBodySimProStatus nbx_generate_plummer(NBodySimProParticleSystemView* v, double a, double mass,
                                       uint64_t seed) {
    if (!view_ok(v) || !(a > 0.0) || !(mass > 0.0)) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    const size_t n = v->particle_count;
    PlummerJob job = {v, a, mass, mass / (double)n, seed};
    nbx_parallel_ranges(n, plummer_range, &job);
    recentre(v, 0, n);
    return N_BODY_SIM_PRO_STATUS_OK;
}

void nbx_galaxy_pair_defaults(NBXGalaxyPairParams* p) {
    if (p == NULL) {
        return;
    }
    p->disk_scale_length = 1.0;
    p->disk_scale_height = 0.1;
    p->bulge_fraction = 0.1;
    p->bulge_scale = 0.2;
    p->separation = 20.0;
    p->pericenter = 2.0;
    p->inclination = M_PI / 4.0;
    p->velocity_dispersion = 0.1;
}

/* Enclosed mass (spherical approximation) of a disk+bulge galaxy of mass M. */
static double enclosed(const NBXGalaxyPairParams* p, double mass, double r) {
    const double md = mass * (1.0 - p->bulge_fraction);
    const double mb = mass * p->bulge_fraction;
    const double x = r / p->disk_scale_length;
    const double disk = md * (1.0 - (1.0 + x) * exp(-x));
    const double b = p->bulge_scale;
    const double bulge = mb * r * r * r / pow(r * r + b * b, 1.5);
    return disk + bulge;
}

/*
 * Writes one galaxy of total mass `mass` into [start, start+count), centred at
 * the origin, disk in the xy plane. Bulge particles come first.
 */
typedef struct GalaxyJob {
    NBodySimProParticleSystemView* v;
    const NBXGalaxyPairParams* p;
    size_t start, nb;
    double mass, mb, pm;
    uint64_t seed;
} GalaxyJob;

static void galaxy_range(void* ctx, size_t lo, size_t hi) {
    const GalaxyJob* j = (const GalaxyJob*)ctx;
    const NBXGalaxyPairParams* p = j->p;
    for (size_t k = lo; k < hi; ++k) {
        rng g;
        rng_particle(&g, j->seed, j->start + k);
        double pos[3], vel[3];
        if (k < j->nb) {
            plummer_particle(&g, p->bulge_scale, j->mb, pos, vel);
        } else {
            /* surface density ~ exp(-R/Rd): R ~ Gamma(2, Rd); z ~ sech^2(z/z0) */
            double e1, e2;
            do {
                e1 = nbx_rng_unit(&g);
                e2 = nbx_rng_unit(&g);
            } while (e1 <= 0.0 || e2 <= 0.0);
            const double R = p->disk_scale_length * (-log(e1) - log(e2));
            const double phi = 2.0 * M_PI * nbx_rng_unit(&g);
            double w;
            do {
                w = nbx_rng_unit(&g);
            } while (w <= 0.0 || w >= 1.0);
            const double z = p->disk_scale_height * atanh(2.0 * w - 1.0);
            pos[0] = R * cos(phi);
            pos[1] = R * sin(phi);
            pos[2] = z;
            const double rr = sqrt(R * R + z * z);
            const double vc = rr > 0.0 ? sqrt(enclosed(p, j->mass, rr) / rr) : 0.0;
            const double sig = p->velocity_dispersion * vc;
            vel[0] = -vc * sin(phi) + sig * nbx_rng_gauss(&g);
            vel[1] = vc * cos(phi) + sig * nbx_rng_gauss(&g);
            vel[2] = 0.5 * sig * nbx_rng_gauss(&g);
        }
        put(j->v, j->start + k, pos, vel, j->pm);
    }
}

/*
 * Writes one galaxy of total mass `mass` into [start, start+count), centred at
 * the origin, disk in the xy plane. Bulge particles come first. Each particle
 * has its own stream, so the worker split does not change the result.
 */
static void galaxy(NBodySimProParticleSystemView* v, size_t start, size_t count, double mass,
                   const NBXGalaxyPairParams* p, uint64_t seed) {
    size_t nb = (size_t)floor(p->bulge_fraction * (double)count + 0.5);
    if (nb > count) nb = count;
    GalaxyJob job = {v, p, start, nb, mass, mass * p->bulge_fraction, mass / (double)count, seed};
    nbx_parallel_ranges(count, galaxy_range, &job);
    recentre(v, start, count);
}

static const NBXGalaxyPairParams* params_or_default(const NBXGalaxyPairParams* p,
                                                    NBXGalaxyPairParams* store) {
    if (p != NULL) {
        return p;
    }
    nbx_galaxy_pair_defaults(store);
    return store;
}

NBodySimProStatus nbx_generate_disk_galaxy(NBodySimProParticleSystemView* v,
                                           const NBXGalaxyPairParams* params, uint64_t seed) {
    if (!view_ok(v)) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    NBXGalaxyPairParams store;
    const NBXGalaxyPairParams* p = params_or_default(params, &store);
    galaxy(v, 0, v->particle_count, 1.0, p, seed);
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_generate_two_galaxy(NBodySimProParticleSystemView* v,
                                          const NBXGalaxyPairParams* params, uint64_t seed) {
    if (!view_ok(v) || v->particle_count < 2) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    NBXGalaxyPairParams store;
    const NBXGalaxyPairParams* p = params_or_default(params, &store);
    const size_t n = v->particle_count;
    const size_t n1 = n / 2, n2 = n - n1;
    galaxy(v, 0, n1, (double)n1 / (double)n, p, seed);
    galaxy(v, n1, n2, (double)n2 / (double)n, p, seed ^ 0x5DEECE66Dull);
    /* tilt the second disk about the x axis */
    const double ci = cos(p->inclination), si = sin(p->inclination);
    for (size_t i = n1; i < n; ++i) {
        const double y = v->positions_y[i], z = v->positions_z[i];
        v->positions_y[i] = ci * y - si * z;
        v->positions_z[i] = si * y + ci * z;
        const double vy = v->velocities_y[i], vz = v->velocities_z[i];
        v->velocities_y[i] = ci * vy - si * vz;
        v->velocities_z[i] = si * vy + ci * vz;
    }
    /* parabolic relative orbit: |v_rel| = sqrt(2 G M / d), L = sqrt(2 G M r_p) */
    const double mt = 1.0, d = p->separation;
    const double vrel = sqrt(2.0 * mt / d);
    const double vt = sqrt(2.0 * mt * p->pericenter) / d;
    const double vr = -sqrt(vrel * vrel - vt * vt);
    const double f1 = (double)n2 / (double)n, f2 = (double)n1 / (double)n;
    /* galaxy 1 at -x, galaxy 2 at +x; relative vector r2 - r1 = (d, 0, 0) */
    for (size_t i = 0; i < n; ++i) {
        const int second = i >= n1;
        const double f = second ? f2 : -f1;
        v->positions_x[i] += f * d;
        v->velocities_x[i] += f * vr;
        v->velocities_y[i] += f * vt;
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_ensemble_sizes(size_t clusters, uint32_t lo, uint32_t hi, uint64_t seed,
                                     uint32_t* out) {
    if (out == NULL || lo == 0 || hi < lo) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    rng g;
    nbx_rng_seed(&g, seed);
    const double a = log((double)lo), b = log((double)hi + 1.0);
    for (size_t k = 0; k < clusters; ++k) {
        double s = floor(exp(a + (b - a) * nbx_rng_unit(&g)));
        if (s < lo) s = lo;
        if (s > hi) s = hi;
        out[k] = (uint32_t)s;
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}
