/*
 * nbx_host_internal.h - helpers shared by the C17 host generators (not installed).
 *
 * The random number generator is xoshiro256** (Blackman & Vigna, 2018) seeded
 * through the splitmix64 output function, the published pair the reference
 * also uses (src/c/generation/random.c:9-58): the same seed gives the same
 * stream, so presets restated from the reference reproduce it bit for bit.
 * Uniform doubles take the top 53 bits; Gaussians are Box-Muller's cosine
 * branch with u1 redrawn while it is zero.
 */
#ifndef NBX_HOST_INTERNAL_H
#define NBX_HOST_INTERNAL_H

#include <math.h>
#include <stddef.h>
#include <stdint.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct nbx_rng {
    uint64_t s[4];
} nbx_rng;

/* splitmix64's output function (its mixing of one 64-bit word). */
static inline uint64_t nbx_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* One splitmix64 step: advance the Weyl state, mix it. */
static inline uint64_t nbx_splitmix(uint64_t* x) {
    *x += 0x9E3779B97F4A7C15ull;
    return nbx_mix64(*x);
}

static inline void nbx_rng_seed(nbx_rng* g, uint64_t seed) {
    uint64_t x = seed;
    for (int k = 0; k < 4; ++k) {
        g->s[k] = nbx_splitmix(&x);
    }
}

static inline uint64_t nbx_rotl64(uint64_t v, int k) {
    return (v << k) | (v >> (64 - k));
}

static inline uint64_t nbx_rng_u64(nbx_rng* g) {
    const uint64_t r = nbx_rotl64(g->s[1] * 5, 7) * 9;
    const uint64_t t = g->s[1] << 17;
    g->s[2] ^= g->s[0];
    g->s[3] ^= g->s[1];
    g->s[1] ^= g->s[2];
    g->s[0] ^= g->s[3];
    g->s[2] ^= t;
    g->s[3] = nbx_rotl64(g->s[3], 45);
    return r;
}

/* Uniform in [0, 1) with 53 random bits. */
static inline double nbx_rng_unit(nbx_rng* g) {
    return (double)(nbx_rng_u64(g) >> 11) * (1.0 / 9007199254740992.0);
}

static inline double nbx_rng_gauss(nbx_rng* g) {
    double u1 = 0.0;
    while (u1 == 0.0) {
        u1 = nbx_rng_unit(g);
    }
    const double u2 = nbx_rng_unit(g);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* Uniform direction: three Gaussians normalised (redrawn if all vanish). */
static inline void nbx_rng_direction(nbx_rng* g, double d[3]) {
    for (;;) {
        const double a = nbx_rng_gauss(g), b = nbx_rng_gauss(g), c = nbx_rng_gauss(g);
        const double r = sqrt(a * a + b * b + c * c);
        if (r > 1.0e-12) {
            d[0] = a / r;
            d[1] = b / r;
            d[2] = c / r;
            return;
        }
    }
}

/*
 * Calls fn(ctx, lo, hi) over a partition of [0, n) into contiguous ranges, on
 * the host worker threads (nbx_set_host_threads; all online CPUs by default).
 * Small n runs on the calling thread. Every range is done when it returns.
 */
typedef void (*nbx_range_fn)(void* ctx, size_t lo, size_t hi);
__attribute__((visibility("hidden"))) void nbx_parallel_ranges(size_t n, nbx_range_fn fn,
                                                               void* ctx);

#endif /* NBX_HOST_INTERNAL_H */
