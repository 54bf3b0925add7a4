/*
 * nbx_headless.c - the command-line runner on the device-resident engine (C17).
 *
 * The reference's CLI subcommands (src/main.cpp:50-72 usage, 83-169
 * benchmark, 291-380 hardware / resume / save) with its headless runner
 * (src/cpp/benchmark/HeadlessRunner.cpp:17-103) and the parts of
 * SimulationController they use (src/cpp/simulation/SimulationController.cpp:
 * 27-65 apply_preset, 153-191 step, 213-272 checkpoints, 276-312 diagnostics
 * reference). The particles stay in HBM for the whole run: one
 * nbx_engine_advance call runs all timed steps, with no per-step copies.
 *
 *   nbx_headless hardware
 *   nbx_headless benchmark [--particles N] [--steps S] [--threads T]
 *                          [--algorithm all_pairs|barnes_hut|hip_*] [--theta T]
 *                          [--preset P] [--seed S] [--dt DT] [--eps E]
 *                          [--device D] [--save FILE]
 *   nbx_headless resume FILE [--steps S] [--threads T] [--algorithm A] [--device D]
 *                            [--save FILE]
 *   nbx_headless save FILE [--particles N] [--preset P] [--seed S] [--device D]
 *
 * Semantics kept from the reference:
 *   - G = 1, softening 0 (SimulationController.hpp:126-127; --eps overrides);
 *   - `barnes_hut` / `all_pairs` are the reference's words; the device runs
 *     them as hip_barnes_hut and hip_direct_fp64, and without --algorithm the
 *     choice is Barnes-Hut above 20000 particles (main.cpp:157-159);
 *   - presets by the reference's names and generators, the per-particle
 *     (parallel) generator from 100000 particles (HeadlessRunner.cpp:28);
 *   - accelerations are computed once after generation (the KDK contract,
 *     SimulationController.cpp:59), one first step is timed apart, then
 *     `steps` steps are timed together (HeadlessRunner.cpp:38-58);
 *   - momentum error |P - P0| / (1 + sum m|v|) and energy drift
 *     |E - E0| / |E0| (only up to 20000 particles) against the state after
 *     generation or loading (SimulationController.cpp:276-312), here from the
 *     device's FP64 reductions and pair sum;
 *   - `resume` continues a .hpcs checkpoint with its timestep, theta,
 *     integrator and Barnes-Hut flag (SimulationController.cpp:242-272);
 *     `save` writes a fresh preset as the reference's controller would (its
 *     defaults: dt 0.005, leapfrog, theta 0.7, all-pairs accelerations).
 * Exit status 0 on success, 1 on any error (as main.cpp).
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "nbx_checkpoint.h"
#include "nbx_engine.h"
#include "nbx_presets.h"

enum { ENERGY_TRACK_MAX_PARTICLES = 20000, PARALLEL_GENERATION_MIN = 100000 };

/* The reference's own generators plus this library's two galaxy-pair /
 * Plummer-DF models under names of their own. */
enum { PRESET_NBX_TWO_GALAXY = 100, PRESET_NBX_PLUMMER = 101 };

typedef struct Run {
    int preset;
    size_t particle_count;
    uint64_t seed;
    int steps;
    int threads;
    int algorithm;          /* NBXAlgorithm, -1 = by particle count */
    double theta;
    double dt;
    double eps;
    int integrator;
    int device;
    int parallel_generation;  /* per-particle preset generator */
    const char* save;
    const char* resume;
} Run;

typedef struct Report {
    double first_step_ms, average_step_ms, tree_build_ms, force_evaluation_ms;
    double energy_drift, momentum_error;
    int energy_available;
    size_t device_bytes;
    double simulation_time;
} Report;

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

static void usage(const char* p) {
    printf("N-Body Sim Pro - HIP engine (MI355X)\n");
    printf("Usage:\n");
    printf("  %s hardware           print the HIP device and host CPU\n", p);
    printf("  %s benchmark [options] run a headless benchmark on the device\n", p);
    printf("    --particles N       particle count (65536)\n");
    printf("    --steps S           number of simulated steps (100)\n");
    printf("    --threads T         host generator threads (0 = all CPUs)\n");
    printf("    --algorithm A       all_pairs | barnes_hut | hip_direct_fp64 | hip_direct_fp32 |\n");
    printf("                        hip_barnes_hut | hip_barnes_hut_fp64\n");
    printf("    --theta T           Barnes-Hut opening angle (0.7)\n");
    printf("    --preset P          two_body, random_cloud, solar_system,\n");
    printf("                        open_cluster, globular_cluster, spiral_galaxy,\n");
    printf("                        elliptical_galaxy, galaxy_collision,\n");
    printf("                        triple_galaxy, nbx_two_galaxy, nbx_plummer\n");
    printf("    --seed S  --dt DT  --eps E  --device D  --save FILE\n");
    printf("  %s resume FILE [--steps S] [--threads T] [--algorithm A] [--device D] [--save FILE]\n", p);
    printf("                        continue a checkpoint headless\n");
    printf("  %s save FILE [--particles N] [--preset P] [--seed S] [--device D]\n", p);
    printf("                        write a fresh preset as a checkpoint\n");
}

static int parse_unsigned(const char* text, unsigned long long* value) {
    if (text == NULL || *text == '\0' || *text == '-') {
        return 0;
    }
    char* end = NULL;
    *value = strtoull(text, &end, 10);
    return end != text && *end == '\0';
}

static int parse_double(const char* text, double* value) {
    if (text == NULL || *text == '\0') {
        return 0;
    }
    char* end = NULL;
    *value = strtod(text, &end);
    return end != text && *end == '\0';
}

static int preset_from_name(const char* name) {
    if (name == NULL) return -1;
    if (strcmp(name, "nbx_two_galaxy") == 0) return PRESET_NBX_TWO_GALAXY;
    if (strcmp(name, "nbx_plummer") == 0) return PRESET_NBX_PLUMMER;
    return nbx_preset_from_name(name);
}

static const char* preset_name(int preset) {
    if (preset == PRESET_NBX_TWO_GALAXY) return "nbx_two_galaxy";
    if (preset == PRESET_NBX_PLUMMER) return "nbx_plummer";
    return nbx_preset_name(preset);
}

/* main.cpp:120-134 plus the device names. */
static int algorithm_from_name(const char* name) {
    if (name == NULL) return -2;
    if (strcmp(name, "barnes_hut") == 0) return NBX_ALGORITHM_BARNES_HUT;
    if (strcmp(name, "all_pairs") == 0) return NBX_ALGORITHM_DIRECT_FP64;
    const int a = nbx_algorithm_from_name(name);
    return a >= 0 ? a : -2;
}

static int is_tree(int algorithm) {
    return algorithm == NBX_ALGORITHM_BARNES_HUT || algorithm == NBX_ALGORITHM_BARNES_HUT_FP64 ||
           algorithm == NBX_ALGORITHM_BH_LEARNED;
}

/* True when `arg` is one of the space-separated option names in `allowed`. */
static int allowed_option(const char* allowed, const char* arg) {
    const size_t len = strlen(arg);
    for (const char* p = allowed; (p = strstr(p, arg)) != NULL; p += len) {
        if ((p == allowed || p[-1] == ' ') && (p[len] == ' ' || p[len] == '\0')) return 1;
    }
    return 0;
}

/*
 * Options shared by the subcommands; `allowed` names the ones this
 * subcommand takes. Returns 0, or 1 after printing the problem.
 */
static int parse_options(const char* cmd, int argc, char** argv, int first, const char* allowed,
                         Run* r) {
    for (int i = first; i < argc; ++i) {
        const char* a = argv[i];
        if (strncmp(a, "--", 2) != 0 || !allowed_option(allowed, a)) {
            fprintf(stderr, "%s: unknown argument %s\n", cmd, a);
            return 1;
        }
        if (i + 1 >= argc) {
            fprintf(stderr, "%s: missing value for %s\n", cmd, a);
            return 1;
        }
        const char* v = argv[++i];
        unsigned long long u = 0;
        double d = 0.0;
        int ok = 1;
        if (strcmp(a, "--particles") == 0) {
            ok = parse_unsigned(v, &u) && u > 0;
            r->particle_count = (size_t)u;
        } else if (strcmp(a, "--steps") == 0) {
            ok = parse_unsigned(v, &u) && u > 0 && u < 1000000000ull;
            r->steps = (int)u;
        } else if (strcmp(a, "--threads") == 0) {
            ok = parse_unsigned(v, &u) && u < 4096;
            r->threads = (int)u;
        } else if (strcmp(a, "--seed") == 0) {
            ok = parse_unsigned(v, &u);
            r->seed = (uint64_t)u;
        } else if (strcmp(a, "--device") == 0) {
            ok = parse_unsigned(v, &u) && u < 1024;
            r->device = (int)u;
        } else if (strcmp(a, "--theta") == 0) {
            ok = parse_double(v, &d) && d >= 0.0;
            r->theta = d;
        } else if (strcmp(a, "--dt") == 0) {
            ok = parse_double(v, &d) && d > 0.0;
            r->dt = d;
        } else if (strcmp(a, "--eps") == 0) {
            ok = parse_double(v, &d) && d >= 0.0;
            r->eps = d;
        } else if (strcmp(a, "--algorithm") == 0) {
            r->algorithm = algorithm_from_name(v);
            if (r->algorithm < 0) {
                fprintf(stderr, "%s: unknown algorithm %s\n", cmd, v);
                return 1;
            }
            if (r->algorithm == NBX_ALGORITHM_BH_LEARNED) {
                /* its network weights are installed through nbx_engine_set_mlp */
                fprintf(stderr, "%s: hip_bh_learned needs a trained model; run it through the "
                                "library (nbx_engine_set_mlp)\n", cmd);
                return 1;
            }
        } else if (strcmp(a, "--preset") == 0) {
            r->preset = preset_from_name(v);
            if (r->preset < 0) {
                fprintf(stderr, "%s: unknown preset %s\n", cmd, v);
                return 1;
            }
        } else if (strcmp(a, "--save") == 0) {
            r->save = v;
        }
        if (!ok) {
            fprintf(stderr, "%s: invalid value %s for %s\n", cmd, v, a);
            return 1;
        }
    }
    return 0;
}

static int fail(const char* cmd, const char* what, const NBodySimProError* e) {
    fprintf(stderr, "%s: %s failed: %s\n", cmd, what, e->message);
    return 1;
}

/* Ten host arrays of n doubles in one block, as a view. */
static double* host_view(size_t n, NBodySimProParticleSystemView* v) {
    double* b = (double*)calloc(n * 10, sizeof(double));
    if (b == NULL) return NULL;
    double** f[10] = {&v->positions_x,     &v->positions_y,     &v->positions_z,
                      &v->velocities_x,    &v->velocities_y,    &v->velocities_z,
                      &v->accelerations_x, &v->accelerations_y, &v->accelerations_z,
                      &v->masses};
    for (int k = 0; k < 10; ++k) *f[k] = b + (size_t)k * n;
    v->particle_count = n;
    return b;
}

static NBodySimProStatus generate(const Run* r, NBodySimProParticleSystemView* v,
                                  NBodySimProError* err) {
    if (r->preset == PRESET_NBX_TWO_GALAXY) return nbx_generate_two_galaxy(v, NULL, r->seed);
    if (r->preset == PRESET_NBX_PLUMMER) return nbx_generate_plummer(v, 1.0, 1.0, r->seed);
    if (r->parallel_generation) {
        return nbx_preset_generate_parallel(v, r->preset, r->seed, err);
    }
    return nbx_preset_generate(v, r->preset, r->seed, err);
}

static NBodySimProGravity gravity_of(const Run* r) {
    NBodySimProGravity g;
    g.gravitational_constant = 1.0;
    g.softening_length = r->eps;
    g.softening_squared = r->eps * r->eps;
    return g;
}

/* 1 + sum m |v| over the device state (SimulationController.cpp:288-300). */
static int momentum_scale(NBXEngine* e, size_t n, double* scale, NBodySimProError* err) {
    NBodySimProParticleSystemView v;
    double* block = host_view(n, &v);
    if (block == NULL) return 1;
    if (nbx_engine_download(e, &v, err) != N_BODY_SIM_PRO_STATUS_OK) {
        free(block);
        return 1;
    }
    double s = 1.0;
    for (size_t i = 0; i < n; ++i) {
        s += v.masses[i] * sqrt(v.velocities_x[i] * v.velocities_x[i] +
                                v.velocities_y[i] * v.velocities_y[i] +
                                v.velocities_z[i] * v.velocities_z[i]);
    }
    *scale = s == 0.0 ? 1.0 : s;
    free(block);
    return 0;
}

/* The first step, then `steps` timed steps, then the diagnostics. */
static int simulate(const char* cmd, NBXEngine* e, const Run* r, size_t n, Report* rep) {
    NBodySimProError err;
    memset(&err, 0, sizeof(err));
    const NBodySimProGravity g = gravity_of(r);
    double scale = 1.0;
    if (momentum_scale(e, n, &scale, &err) != 0) return fail(cmd, "download", &err);
    rep->energy_available = n <= ENERGY_TRACK_MAX_PARTICLES;
    NBodySimProDiagnosticsQuantities q0, q1;
    double u0 = 0.0, u1 = 0.0;
    if (nbx_engine_diagnostics(e, &g, &q0, rep->energy_available ? &u0 : NULL, &err) !=
        N_BODY_SIM_PRO_STATUS_OK) {
        return fail(cmd, "diagnostics", &err);
    }
    const NBXAlgorithm algo = (NBXAlgorithm)r->algorithm;
    const NBodySimProIntegratorType integ = (NBodySimProIntegratorType)r->integrator;
    double t0 = now_ms();
    if (nbx_engine_advance(e, algo, &g, integ, r->dt, 1, &err) != N_BODY_SIM_PRO_STATUS_OK) {
        return fail(cmd, "first step", &err);
    }
    rep->first_step_ms = now_ms() - t0;
    NBXTimings tm;
    nbx_engine_timings(e, &tm);
    if (is_tree(r->algorithm)) {
        rep->tree_build_ms = tm.bbox_keys_ms + tm.sort_ms + tm.tree_ms + tm.com_ms;
        rep->force_evaluation_ms = tm.walk_ms;
    } else {
        rep->tree_build_ms = 0.0;
        rep->force_evaluation_ms = tm.force_ms;
    }
    t0 = now_ms();
    if (nbx_engine_advance(e, algo, &g, integ, r->dt, r->steps, &err) != N_BODY_SIM_PRO_STATUS_OK) {
        return fail(cmd, "steps", &err);
    }
    rep->average_step_ms = (now_ms() - t0) / (double)r->steps;
    if (nbx_engine_diagnostics(e, &g, &q1, rep->energy_available ? &u1 : NULL, &err) !=
        N_BODY_SIM_PRO_STATUS_OK) {
        return fail(cmd, "diagnostics", &err);
    }
    const double dx = q1.total_momentum_x - q0.total_momentum_x;
    const double dy = q1.total_momentum_y - q0.total_momentum_y;
    const double dz = q1.total_momentum_z - q0.total_momentum_z;
    rep->momentum_error = sqrt(dx * dx + dy * dy + dz * dz) / scale;
    if (rep->energy_available) {
        const double e0 = q0.kinetic_energy + u0, e1 = q1.kinetic_energy + u1;
        rep->energy_drift = fabs(e1 - e0) / (fabs(e0) > 0.0 ? fabs(e0) : 1.0);
    }
    rep->simulation_time += r->dt * (double)(r->steps + 1);
    rep->device_bytes = nbx_engine_device_bytes(e);
    return 0;
}

static int save_state(const char* cmd, NBXEngine* e, const Run* r, double simulation_time,
                      int barnes_hut) {
    NBodySimProCheckpointHeader h;
    memset(&h, 0, sizeof(h));
    h.simulation_time = simulation_time;
    h.timestep = r->dt;
    h.integrator = r->integrator;
    h.theta = r->theta;
    h.barnes_hut_enabled = barnes_hut;
    h.random_seed = r->seed;
    /* the format has no code for this library's own two models */
    h.preset = r->preset == PRESET_NBX_TWO_GALAXY ? NBX_PRESET_GALAXY_COLLISION
               : r->preset == PRESET_NBX_PLUMMER  ? NBX_PRESET_GLOBULAR_CLUSTER
                                                  : r->preset;
    NBodySimProError err;
    memset(&err, 0, sizeof(err));
    if (nbx_engine_checkpoint_save(e, r->save, &h, &err) != N_BODY_SIM_PRO_STATUS_OK) {
        return fail(cmd, "checkpoint save", &err);
    }
    return 0;
}

static void cpu_brand(char* out, size_t len, int* avx2, int* avx512) {
    snprintf(out, len, "unknown");
    *avx2 = *avx512 = 0;
    FILE* f = fopen("/proc/cpuinfo", "r");
    if (f == NULL) return;
    char line[8192];
    int have_name = 0;
    while (fgets(line, sizeof(line), f) != NULL) {
        if (!have_name && strncmp(line, "model name", 10) == 0) {
            const char* c = strchr(line, ':');
            if (c != NULL) {
                snprintf(out, len, "%s", c + 2);
                out[strcspn(out, "\n")] = '\0';
                have_name = 1;
            }
        }
        if (strncmp(line, "flags", 5) == 0) {
            *avx2 = strstr(line, " avx2") != NULL;
            *avx512 = strstr(line, " avx512f") != NULL;
            break;
        }
    }
    fclose(f);
}

static void json_string(const char* s) {
    putchar('"');
    for (; *s; ++s) {
        if (*s == '"' || *s == '\\') putchar('\\');
        if ((unsigned char)*s >= 0x20) putchar(*s);
    }
    putchar('"');
}

/* HeadlessRunner.cpp:106-164, with the device and the device-side numbers. */
static void print_report(const Run* r, size_t n, const Report* rep) {
    char device[256] = "unknown", cpu[256];
    int cus = 0, avx2 = 0, avx512 = 0;
    nbx_device_info(r->device, device, sizeof(device), &cus);
    cpu_brand(cpu, sizeof(cpu), &avx2, &avx512);
    const int bh = is_tree(r->algorithm);
    const char* family = bh ? "barnes_hut" : "all-pairs";
    printf("--- N-Body Sim Pro headless benchmark (HIP engine) ---\n");
    printf("Device         : %s, %d CUs\n", device, cus);
    printf("CPU            : %s\n", cpu);
    printf("SIMD           : AVX2 %s, AVX-512 %s, NEON no\n", avx2 ? "yes" : "no",
           avx512 ? "yes" : "no");
    printf("Host threads   : %d\n", nbx_host_threads());
    printf("Particles      : %zu\n", n);
    printf("Algorithm      : %s (theta %.2f) on %s\n", family, r->theta,
           nbx_algorithm_name(r->algorithm));
    printf("Steps          : %d\n", r->steps);
    printf("Estimated mem  : %.2f MiB on the device\n",
           (double)rep->device_bytes / (1024.0 * 1024.0));
    printf("Step time      : avg %.3f ms, first %.3f ms\n", rep->average_step_ms,
           rep->first_step_ms);
    if (bh) {
        printf("Tree build     : %.3f ms\n", rep->tree_build_ms);
        printf("Force eval     : %.3f ms\n", rep->force_evaluation_ms);
    }
    if (rep->energy_available) printf("Energy drift   : %.6e\n", rep->energy_drift);
    else printf("Energy drift   : N/A\n");
    printf("Momentum error : %.6e\n", rep->momentum_error);
    printf("--- machine-readable ---\n");
    printf("{\"cpu\":");
    json_string(cpu);
    printf(",\"simd\":{\"avx2\":%s,\"avx512\":%s,\"neon\":false},\"threads\":%d,"
           "\"particles\":%zu,\"algorithm\":\"%s\",\"theta\":%.2f,\"steps\":%d,"
           "\"estimated_bytes\":%zu,\"avg_step_ms\":%.6f,\"first_step_ms\":%.6f,\"energy_drift\":",
           avx2 ? "true" : "false", avx512 ? "true" : "false", nbx_host_threads(), n, family,
           r->theta, r->steps, rep->device_bytes, rep->average_step_ms, rep->first_step_ms);
    if (rep->energy_available) printf("%.6e", rep->energy_drift);
    else printf("null");
    printf(",\"momentum_error\":%.6e,\"device\":", rep->momentum_error);
    json_string(device);
    printf(",\"engine_algorithm\":\"%s\",\"tree_build_ms\":%.6f,\"force_evaluation_ms\":%.6f,"
           "\"preset\":\"%s\",\"dt\":%.17g,\"eps\":%.17g,\"resumed\":%s}\n",
           nbx_algorithm_name(r->algorithm), rep->tree_build_ms, rep->force_evaluation_ms,
           preset_name(r->preset), r->dt, r->eps, r->resume ? "true" : "false");
}

static NBXEngine* create_engine(const char* cmd, const Run* r, size_t n) {
    NBodySimProError err;
    memset(&err, 0, sizeof(err));
    NBXConfig cfg = {n, r->device, r->theta, 0u};
    NBXEngine* e = nbx_engine_create(&cfg, &err);
    if (e == NULL) fail(cmd, "engine create", &err);
    return e;
}

/* A preset on the device with its initial accelerations (apply_preset). */
static NBXEngine* prepare(const char* cmd, const Run* r) {
    const size_t n = r->particle_count;
    NBodySimProParticleSystemView v;
    double* block = host_view(n, &v);
    if (block == NULL) {
        fprintf(stderr, "%s: host allocation failed\n", cmd);
        return NULL;
    }
    NBodySimProError err;
    memset(&err, 0, sizeof(err));
    if (generate(r, &v, &err) != N_BODY_SIM_PRO_STATUS_OK) {
        fprintf(stderr, "%s: failed to initialize preset: %s\n", cmd,
                err.message[0] ? err.message : "invalid particle count for this preset");
        free(block);
        return NULL;
    }
    NBXEngine* e = create_engine(cmd, r, n);
    if (e == NULL) {
        free(block);
        return NULL;
    }
    const NBodySimProGravity g = gravity_of(r);
    if (nbx_engine_upload(e, &v, &err) != N_BODY_SIM_PRO_STATUS_OK ||
        nbx_engine_compute_force(e, (NBXAlgorithm)r->algorithm, &g, &err) !=
            N_BODY_SIM_PRO_STATUS_OK) {
        fail(cmd, "initial forces", &err);
        nbx_engine_destroy(e);
        e = NULL;
    }
    free(block);
    return e;
}

static void defaults(Run* r) {
    memset(r, 0, sizeof(*r));
    r->preset = NBX_PRESET_GALAXY_COLLISION;
    r->particle_count = 65536;
    r->seed = 42;
    r->steps = 100;
    r->algorithm = -1;
    r->theta = 0.7;
    r->dt = 0.001;
    r->integrator = N_BODY_SIM_PRO_INTEGRATOR_LEAPFROG;
}

static int cmd_benchmark(int argc, char** argv) {
    Run r;
    defaults(&r);
    if (parse_options("benchmark", argc, argv, 2,
                      "--particles --steps --threads --algorithm --theta --preset --seed --dt "
                      "--eps --device --save",
                      &r) != 0) {
        return 1;
    }
    if (r.algorithm < 0) {
        r.algorithm = r.particle_count > ENERGY_TRACK_MAX_PARTICLES ? NBX_ALGORITHM_BARNES_HUT
                                                                    : NBX_ALGORITHM_DIRECT_FP64;
    }
    nbx_set_host_threads(r.threads);
    r.parallel_generation = r.particle_count >= PARALLEL_GENERATION_MIN;
    NBXEngine* e = prepare("headless", &r);
    if (e == NULL) return 1;
    Report rep;
    memset(&rep, 0, sizeof(rep));
    int status = simulate("headless", e, &r, r.particle_count, &rep);
    if (status == 0 && r.save != NULL) {
        status = save_state("headless", e, &r, rep.simulation_time, is_tree(r.algorithm));
    }
    if (status == 0) print_report(&r, r.particle_count, &rep);
    nbx_engine_destroy(e);
    return status;
}

static int cmd_resume(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "resume: missing checkpoint path\n");
        return 1;
    }
    Run r;
    defaults(&r);
    r.resume = argv[2];
    if (parse_options("resume", argc, argv, 3, "--steps --threads --algorithm --device --save",
                      &r) != 0) {
        return 1;
    }
    nbx_set_host_threads(r.threads);
    NBodySimProError err;
    memset(&err, 0, sizeof(err));
    NBodySimProCheckpointHeader h;
    if (nbx_checkpoint_peek(r.resume, &h, &err) != N_BODY_SIM_PRO_STATUS_OK) {
        fprintf(stderr, "resume: failed to load checkpoint: %s\n", err.message);
        return 1;
    }
    /* the checkpoint's own settings (SimulationController.cpp:260-266) */
    r.dt = h.timestep;
    r.theta = h.theta;
    r.integrator = h.integrator;
    r.seed = h.random_seed;
    r.preset = h.preset;
    if (r.algorithm < 0) {
        r.algorithm = h.barnes_hut_enabled ? NBX_ALGORITHM_BARNES_HUT : NBX_ALGORITHM_DIRECT_FP64;
    }
    if (h.particle_count == 0 || !(r.dt > 0.0) || r.integrator < 0 || r.integrator > 2) {
        fprintf(stderr, "resume: checkpoint holds no particles, or an invalid timestep or "
                        "integrator\n");
        return 1;
    }
    NBXEngine* e = create_engine("resume", &r, h.particle_count);
    if (e == NULL) return 1;
    if (nbx_engine_checkpoint_load(e, r.resume, &h, &err) != N_BODY_SIM_PRO_STATUS_OK) {
        fprintf(stderr, "resume: failed to load checkpoint: %s\n", err.message);
        nbx_engine_destroy(e);
        return 1;
    }
    Report rep;
    memset(&rep, 0, sizeof(rep));
    rep.simulation_time = h.simulation_time;
    int status = simulate("resume", e, &r, h.particle_count, &rep);
    if (status == 0 && r.save != NULL) {
        status = save_state("resume", e, &r, rep.simulation_time, is_tree(r.algorithm));
    }
    if (status == 0) {
        /* main.cpp:318-323, then the full report */
        printf("Resumed %zu steps from %s\n", (size_t)r.steps, r.resume);
        printf("Avg step : %.3f ms\n", rep.average_step_ms);
        printf("Energy drift : %.6e\n", rep.energy_drift);
        printf("Momentum error : %.6e\n", rep.momentum_error);
        print_report(&r, h.particle_count, &rep);
    }
    nbx_engine_destroy(e);
    return status;
}

static int cmd_save(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "save: missing checkpoint path\n");
        return 1;
    }
    Run r;
    defaults(&r);
    r.save = argv[2];
    if (parse_options("save", argc, argv, 3, "--particles --preset --seed --device", &r) != 0) {
        return 1;
    }
    /* the reference controller's defaults (SimulationController.hpp:57-75):
     * timestep 0.005, leapfrog, theta 0.7, Barnes-Hut off, sequential
     * generation; its initial accelerations are the all-pairs FP64 sum */
    r.dt = 0.005;
    r.algorithm = NBX_ALGORITHM_DIRECT_FP64;
    NBXEngine* e = prepare("save", &r);
    if (e == NULL) return 1;
    const int status = save_state("save", e, &r, 0.0, 0);
    nbx_engine_destroy(e);
    return status;
}

static int cmd_hardware(void) {
    char device[256] = "none", cpu[256];
    int cus = 0, avx2 = 0, avx512 = 0;
    const int devices = nbx_device_count();
    cpu_brand(cpu, sizeof(cpu), &avx2, &avx512);
    printf("CPU             : %s\n", cpu);
    printf("SIMD            : AVX2 %s, AVX-512 %s\n", avx2 ? "yes" : "no", avx512 ? "yes" : "no");
    printf("Host threads    : %d\n", nbx_host_threads());
    printf("HIP devices     : %d\n", devices);
    for (int d = 0; d < devices; ++d) {
        if (nbx_device_info(d, device, sizeof(device), &cus) == 0) {
            printf("  device %d      : %s, %d CUs\n", d, device, cus);
        }
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && strcmp(argv[1], "hardware") == 0) return cmd_hardware();
    if (argc >= 2 && strcmp(argv[1], "benchmark") == 0) return cmd_benchmark(argc, argv);
    if (argc >= 2 && strcmp(argv[1], "resume") == 0) return cmd_resume(argc, argv);
    if (argc >= 2 && strcmp(argv[1], "save") == 0) return cmd_save(argc, argv);
    if (argc >= 2 && (strcmp(argv[1], "--help") == 0 || strcmp(argv[1], "-h") == 0)) {
        usage(argv[0]);
        return 0;
    }
    usage(argv[0]);
    return 1;
}
