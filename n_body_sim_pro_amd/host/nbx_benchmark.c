/*
 * nbx_benchmark.c - force-evaluation benchmark for the HIP engine (C17).
 *
 * Mirrors the reference's force benchmark (benchmarks/benchmark_force.c:53-234):
 * same --particles/--steps/--algorithm/--theta options, the same default input
 * (random_cloud preset, seed 42, G = 1, eps = 0.01; benchmark_force.c:164-173),
 * the same wall-clock timing (CLOCK_MONOTONIC, ms per evaluation) and a CSV
 * summary. It adds the hip_* algorithm names, a warm-up followed by the median
 * (the reference's mean-without-warm-up is printed too), the device-resident
 * mode next to the drop-in (PCIe-inclusive) mode, and roofline columns from
 * the engine's HIP-event timings.
 *
 *   nbx_benchmark --particles 1000000 --steps 5 --algorithm hip_barnes_hut
 *                 [--theta 0.7] [--preset random_cloud|plummer|two_galaxy]
 *                 [--seed 42] [--eps 0.01] [--warmup 1] [--mode resident|dropin|both]
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "nbx_engine.h"
#include "nbx_presets.h"

#define FP32_PEAK_TFLOPS 157.3
#define FP64_PEAK_TFLOPS 78.6

typedef struct Options {
    size_t particle_count;
    int steps;
    int warmup;
    const char* algorithm;
    const char* preset;
    const char* mode;
    double theta;
    double eps;
    uint64_t seed;
} Options;

static double wall_time_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + (double)ts.tv_nsec * 1.0e-9;
}

static void print_usage(const char* program) {
    fprintf(stderr,
            "Usage: %s --particles N --steps S [--algorithm "
            "hip_direct_fp64|hip_direct_fp32|hip_barnes_hut|hip_barnes_hut_fp64|hip_bh_learned] "
            "[--theta T] [--preset random_cloud|plummer|two_galaxy] [--seed S] [--eps E] "
            "[--warmup W] [--mode resident|dropin|both]\n",
            program);
}

static int parse_arguments(int argc, char** argv, Options* o) {
    memset(o, 0, sizeof(*o));
    o->particle_count = 10000;
    o->steps = 5;
    o->warmup = 1;
    o->algorithm = "hip_direct_fp32";
    o->preset = "random_cloud";
    o->mode = "both";
    o->theta = 0.7;
    o->eps = 0.01;
    o->seed = 42;
    for (int i = 1; i < argc; ++i) {
        const int has = i + 1 < argc;
        if (strcmp(argv[i], "--particles") == 0 && has) {
            o->particle_count = (size_t)strtoull(argv[++i], NULL, 10);
        } else if (strcmp(argv[i], "--steps") == 0 && has) {
            o->steps = atoi(argv[++i]);
        } else if (strcmp(argv[i], "--warmup") == 0 && has) {
            o->warmup = atoi(argv[++i]);
        } else if (strcmp(argv[i], "--algorithm") == 0 && has) {
            o->algorithm = argv[++i];
        } else if (strcmp(argv[i], "--preset") == 0 && has) {
            o->preset = argv[++i];
        } else if (strcmp(argv[i], "--mode") == 0 && has) {
            o->mode = argv[++i];
        } else if (strcmp(argv[i], "--theta") == 0 && has) {
            o->theta = atof(argv[++i]);
        } else if (strcmp(argv[i], "--eps") == 0 && has) {
            o->eps = atof(argv[++i]);
        } else if (strcmp(argv[i], "--seed") == 0 && has) {
            o->seed = (uint64_t)strtoull(argv[++i], NULL, 10);
        } else if (strcmp(argv[i], "--help") == 0 || strcmp(argv[i], "-h") == 0) {
            print_usage(argv[0]);
            exit(0);
        } else {
            fprintf(stderr, "Unknown or incomplete argument: %s\n", argv[i]);
            print_usage(argv[0]);
            return 0;
        }
    }
    if (o->particle_count == 0 || o->steps <= 0 || o->warmup < 0) {
        fprintf(stderr, "--particles and --steps must be positive, --warmup >= 0\n");
        return 0;
    }
    if (nbx_algorithm_from_name(o->algorithm) < 0) {
        fprintf(stderr, "unknown algorithm: %s\n", o->algorithm);
        return 0;
    }
    return 1;
}

/* 64-byte aligned SoA storage, as the reference's particle system
 * (include/n_body_sim_pro/core/particle_system.h:64). */
typedef struct Storage {
    double* a[10];
    NBodySimProParticleSystemView view;
} Storage;

static int storage_create(Storage* s, size_t n) {
    memset(s, 0, sizeof(*s));
    const size_t bytes = ((n * sizeof(double) + 63) / 64) * 64;
    for (int k = 0; k < 10; ++k) {
        s->a[k] = (double*)aligned_alloc(64, bytes);
        if (s->a[k] == NULL) {
            return 0;
        }
        memset(s->a[k], 0, bytes);
    }
    s->view.particle_count = n;
    s->view.positions_x = s->a[0];
    s->view.positions_y = s->a[1];
    s->view.positions_z = s->a[2];
    s->view.velocities_x = s->a[3];
    s->view.velocities_y = s->a[4];
    s->view.velocities_z = s->a[5];
    s->view.accelerations_x = s->a[6];
    s->view.accelerations_y = s->a[7];
    s->view.accelerations_z = s->a[8];
    s->view.masses = s->a[9];
    return 1;
}

static void storage_destroy(Storage* s) {
    for (int k = 0; k < 10; ++k) {
        free(s->a[k]);
    }
}

static int cmp_double(const void* a, const void* b) {
    const double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

static double median(double* v, int n) {
    qsort(v, (size_t)n, sizeof(double), cmp_double);
    return n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

typedef NBodySimProStatus (*ForceFn)(const NBodySimProParticleSystemView*,
                                     const NBodySimProGravity*, void*, NBodySimProError*);

static ForceFn dropin_function(int algorithm) {
    switch (algorithm) {
        case NBX_ALGORITHM_DIRECT_FP64: return nbx_hip_direct_fp64;
        case NBX_ALGORITHM_DIRECT_FP32: return nbx_hip_direct_fp32;
        case NBX_ALGORITHM_BARNES_HUT: return nbx_hip_barnes_hut;
        case NBX_ALGORITHM_BARNES_HUT_FP64: return nbx_hip_barnes_hut_fp64;
        case NBX_ALGORITHM_BH_LEARNED: return nbx_hip_bh_learned;
        default: return NULL;
    }
}

typedef struct Result {
    double first_ms, mean_ms, median_ms, kernel_ms;
    double interactions;
} Result;

/* One timed series. resident: the state stays on the device between
 * evaluations; dropin: every evaluation goes through the
 * NBodySimProForceFunction with the host view (copies included). */
static int run_series(const Options* o, int algorithm, int resident, NBXEngine* engine,
                      Storage* s, const NBodySimProGravity* g, Result* r) {
    NBodySimProError error;
    memset(&error, 0, sizeof(error));
    const int total = o->warmup + o->steps;
    double* times = (double*)malloc(sizeof(double) * (size_t)total);
    double* kernel = (double*)malloc(sizeof(double) * (size_t)total);
    if (times == NULL || kernel == NULL) {
        free(times);
        free(kernel);
        return 0;
    }
    ForceFn fn = dropin_function(algorithm);
    if (resident && nbx_engine_upload(engine, &s->view, &error) != N_BODY_SIM_PRO_STATUS_OK) {
        fprintf(stderr, "upload failed: %s\n", error.message);
        free(times);
        free(kernel);
        return 0;
    }
    for (int k = 0; k < total; ++k) {
        const double t0 = wall_time_seconds();
        NBodySimProStatus st =
            resident ? nbx_engine_compute_force(engine, (NBXAlgorithm)algorithm, g, &error)
                     : fn(&s->view, g, engine, &error);
        const double t1 = wall_time_seconds();
        if (st != N_BODY_SIM_PRO_STATUS_OK) {
            fprintf(stderr, "force evaluation failed: %s\n", error.message);
            free(times);
            free(kernel);
            return 0;
        }
        NBXTimings tm;
        nbx_engine_timings(engine, &tm);
        times[k] = (t1 - t0) * 1e3;
        kernel[k] = tm.kernel_ms;
    }
    double sum = 0.0;
    for (int k = 0; k < total; ++k) {
        sum += times[k];
    }
    r->first_ms = times[0];
    r->mean_ms = sum / (double)total;   /* the reference's mean over every evaluation */
    r->median_ms = median(times + o->warmup, o->steps);
    r->kernel_ms = median(kernel + o->warmup, o->steps);
    const double n = (double)s->view.particle_count;
    r->interactions = n * (n - 1.0);
    if (algorithm == NBX_ALGORITHM_BARNES_HUT || algorithm == NBX_ALGORITHM_BARNES_HUT_FP64 ||
        algorithm == NBX_ALGORITHM_BH_LEARNED) {
        NBodySimProBarnesHutStats bs;
        nbx_engine_stats(engine, &bs);
        r->interactions = (double)(bs.accepted_approximations + bs.exact_interactions);
    }
    free(times);
    free(kernel);
    return 1;
}

int main(int argc, char** argv) {
    Options o;
    if (!parse_arguments(argc, argv, &o)) {
        return 1;
    }
    const int algorithm = nbx_algorithm_from_name(o.algorithm);
    Storage s;
    if (!storage_create(&s, o.particle_count)) {
        fprintf(stderr, "failed to allocate particle storage\n");
        storage_destroy(&s);
        return 1;
    }
    NBodySimProStatus st;
    if (strcmp(o.preset, "random_cloud") == 0) {
        st = nbx_preset_random_cloud(&s.view, o.seed);
    } else if (strcmp(o.preset, "plummer") == 0) {
        st = nbx_generate_plummer(&s.view, 1.0, 1.0, o.seed);
    } else if (strcmp(o.preset, "two_galaxy") == 0) {
        st = nbx_generate_two_galaxy(&s.view, NULL, o.seed);
    } else {
        fprintf(stderr, "unknown preset: %s\n", o.preset);
        storage_destroy(&s);
        return 1;
    }
    if (st != N_BODY_SIM_PRO_STATUS_OK) {
        fprintf(stderr, "failed to generate particles\n");
        storage_destroy(&s);
        return 1;
    }
    NBodySimProGravity g;
    g.gravitational_constant = 1.0;
    g.softening_length = o.eps;
    g.softening_squared = o.eps * o.eps;

    NBodySimProError error;
    memset(&error, 0, sizeof(error));
    NBXConfig cfg = {o.particle_count, 0, o.theta, 0u};
    NBXEngine* engine = nbx_engine_create(&cfg, &error);
    if (engine == NULL) {
        fprintf(stderr, "failed to create the engine: %s\n", error.message);
        storage_destroy(&s);
        return 1;
    }
    printf("particles=%zu steps=%d warmup=%d algorithm=%s preset=%s theta=%g\n",
           o.particle_count, o.steps, o.warmup, o.algorithm, o.preset, o.theta);
    const int fp64 = algorithm == NBX_ALGORITHM_DIRECT_FP64 ||
                     algorithm == NBX_ALGORITHM_BARNES_HUT_FP64;
    const double peak = fp64 ? FP64_PEAK_TFLOPS : FP32_PEAK_TFLOPS;
    char csv[1024];
    size_t off = (size_t)snprintf(csv, sizeof(csv),
                                  "mode,ms_first,ms_mean,ms_median,kernel_ms,"
                                  "ginteractions_s,kernel_tflops,frac_peak\n");
    int ok = 1;
    for (int pass = 0; pass < 2 && ok; ++pass) {
        const int resident = pass == 0;
        if ((resident && strcmp(o.mode, "dropin") == 0) ||
            (!resident && strcmp(o.mode, "resident") == 0)) {
            continue;
        }
        Result r;
        ok = run_series(&o, algorithm, resident, engine, &s, &g, &r);
        if (!ok) {
            break;
        }
        const double gint = r.interactions / (r.median_ms * 1e-3) / 1e9;
        const double tflops = r.kernel_ms > 0.0 ? 20.0 * r.interactions / (r.kernel_ms * 1e-3) / 1e12
                                                : 0.0;
        printf("mode=%-8s ms/evaluation(median)=%-10.3f first=%-10.3f mean=%-10.3f "
               "kernel_ms=%-10.3f Ginteractions/s=%-9.1f kernel TFLOP/s=%.2f (%.1f%% of %.1f)\n",
               resident ? "resident" : "dropin", r.median_ms, r.first_ms, r.mean_ms, r.kernel_ms,
               gint, tflops, 100.0 * tflops / peak, peak);
        off += (size_t)snprintf(csv + off, sizeof(csv) - off, "%s,%.6f,%.6f,%.6f,%.6f,%.3f,%.4f,%.5f\n",
                                resident ? "resident" : "dropin", r.first_ms, r.mean_ms,
                                r.median_ms, r.kernel_ms, gint, tflops, tflops / peak);
    }
    if (ok) {
        printf("%s", csv);
    }
    nbx_engine_destroy(engine);
    storage_destroy(&s);
    return ok ? 0 : 1;
}
