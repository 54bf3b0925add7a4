/*
 * nbx_checkpoint.c - ".hpcs" checkpoint files (include/nbx_checkpoint.h).
 *
 * The reference writes and reads its header as one raw struct
 * (src/c/checkpoint/checkpoint.c:42-49, 186-193), so the file holds the
 * compiler's padding. Here each field is encoded little-endian at the offset
 * that struct has on x86-64 (padding written as zero), which keeps every v1
 * file readable in both directions and makes the format independent of the
 * host compiler. Reads also check that the file holds all ten arrays.
 */
#include "nbx_checkpoint.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { ARRAYS = 10 };

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

static void put_u32(unsigned char* b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b[i] = (unsigned char)(v >> (8 * i));
}

static void put_u64(unsigned char* b, uint64_t v) {
    for (int i = 0; i < 8; ++i) b[i] = (unsigned char)(v >> (8 * i));
}

static void put_f64(unsigned char* b, double v) {
    uint64_t u;
    memcpy(&u, &v, sizeof(u));
    put_u64(b, u);
}

static uint32_t get_u32(const unsigned char* b) {
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) v |= (uint32_t)b[i] << (8 * i);
    return v;
}

static uint64_t get_u64(const unsigned char* b) {
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= (uint64_t)b[i] << (8 * i);
    return v;
}

static double get_f64(const unsigned char* b) {
    const uint64_t u = get_u64(b);
    double v;
    memcpy(&v, &u, sizeof(v));
    return v;
}

/* Field offsets of the 80-byte prefix (the reference's magic, version and
 * NBodySimProCheckpointHeader as laid out by an LP64 compiler). */
enum {
    OFF_MAGIC = 0,
    OFF_VERSION = 4,
    OFF_H_MAGIC = 8,
    OFF_H_VERSION = 12,
    OFF_COUNT = 16,
    OFF_TIME = 24,
    OFF_STEP = 32,
    OFF_INTEGRATOR = 40,
    OFF_THETA = 48,
    OFF_BH = 56,
    OFF_SEED = 64,
    OFF_PRESET = 72
};

static void encode(unsigned char out[NBX_CHECKPOINT_DATA_OFFSET],
                   const NBodySimProCheckpointHeader* h) {
    memset(out, 0, NBX_CHECKPOINT_DATA_OFFSET);
    put_u32(out + OFF_MAGIC, N_BODY_SIM_PRO_CHECKPOINT_MAGIC);
    put_u32(out + OFF_VERSION, N_BODY_SIM_PRO_CHECKPOINT_VERSION);
    put_u32(out + OFF_H_MAGIC, h->magic);
    put_u32(out + OFF_H_VERSION, h->version);
    put_u64(out + OFF_COUNT, (uint64_t)h->particle_count);
    put_f64(out + OFF_TIME, h->simulation_time);
    put_f64(out + OFF_STEP, h->timestep);
    put_u32(out + OFF_INTEGRATOR, (uint32_t)h->integrator);
    put_f64(out + OFF_THETA, h->theta);
    put_u32(out + OFF_BH, (uint32_t)h->barnes_hut_enabled);
    put_u64(out + OFF_SEED, h->random_seed);
    put_u32(out + OFF_PRESET, (uint32_t)h->preset);
}

static void decode(const unsigned char in[NBX_CHECKPOINT_DATA_OFFSET],
                   NBodySimProCheckpointHeader* h) {
    memset(h, 0, sizeof(*h));
    h->magic = get_u32(in + OFF_H_MAGIC);
    h->version = get_u32(in + OFF_H_VERSION);
    h->particle_count = (size_t)get_u64(in + OFF_COUNT);
    h->simulation_time = get_f64(in + OFF_TIME);
    h->timestep = get_f64(in + OFF_STEP);
    h->integrator = (int32_t)get_u32(in + OFF_INTEGRATOR);
    h->theta = get_f64(in + OFF_THETA);
    h->barnes_hut_enabled = (int32_t)get_u32(in + OFF_BH);
    h->random_seed = get_u64(in + OFF_SEED);
    h->preset = (int32_t)get_u32(in + OFF_PRESET);
}

static double* array_of(const NBodySimProParticleSystemView* v, int k) {
    double* const a[ARRAYS] = {v->positions_x,     v->positions_y,     v->positions_z,
                               v->velocities_x,    v->velocities_y,    v->velocities_z,
                               v->accelerations_x, v->accelerations_y, v->accelerations_z,
                               v->masses};
    return a[k];
}

NBodySimProStatus nbx_checkpoint_write(const char* path, const NBodySimProParticleSystemView* view,
                                       const NBodySimProCheckpointHeader* header,
                                       NBodySimProError* error) {
    if (path == NULL || view == NULL || header == NULL) {
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "path, view, and header must not be null");
    }
    if (view->particle_count != header->particle_count) {
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
             "checkpoint particle count does not match the view");
    }
    for (int k = 0; k < ARRAYS; ++k) {
        if (array_of(view, k) == NULL) {
            FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "checkpoint array is null");
        }
    }
    FILE* f = fopen(path, "wb");
    if (f == NULL) {
        FAIL(N_BODY_SIM_PRO_STATUS_IO_ERROR, "failed to open checkpoint file for writing");
    }
    unsigned char head[NBX_CHECKPOINT_DATA_OFFSET];
    encode(head, header);
    int ok = fwrite(head, sizeof(head), 1, f) == 1;
    const size_t n = view->particle_count;
    for (int k = 0; ok && k < ARRAYS && n > 0; ++k) {
        ok = fwrite(array_of(view, k), sizeof(double), n, f) == n;
    }
    if (fclose(f) != 0) {
        ok = 0;
    }
    if (!ok) {
        FAIL(N_BODY_SIM_PRO_STATUS_IO_ERROR, "failed to write checkpoint file");
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

/* Opens `path`, checks magic and version (checkpoint.c:80-102) and decodes the
 * header; on success the stream is left at the first array. */
static NBodySimProStatus open_checked(const char* path, NBodySimProCheckpointHeader* header,
                                      FILE** out, long* length, NBodySimProError* error) {
    FILE* f = fopen(path, "rb");
    if (f == NULL) {
        FAIL(N_BODY_SIM_PRO_STATUS_IO_ERROR, "failed to open checkpoint file for reading");
    }
    unsigned char head[NBX_CHECKPOINT_DATA_OFFSET];
    if (fread(head, 1, 8, f) != 8) {
        fclose(f);
        FAIL(N_BODY_SIM_PRO_STATUS_IO_ERROR, "failed to read checkpoint magic");
    }
    if (get_u32(head + OFF_MAGIC) != N_BODY_SIM_PRO_CHECKPOINT_MAGIC) {
        fclose(f);
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_STATE, "not an N-Body Sim Pro checkpoint file");
    }
    if (get_u32(head + OFF_VERSION) != N_BODY_SIM_PRO_CHECKPOINT_VERSION) {
        fclose(f);
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_STATE, "unsupported checkpoint version");
    }
    if (fread(head + 8, 1, sizeof(head) - 8, f) != sizeof(head) - 8) {
        fclose(f);
        FAIL(N_BODY_SIM_PRO_STATUS_IO_ERROR, "failed to read checkpoint metadata");
    }
    decode(head, header);
    if (length != NULL) {
        if (fseek(f, 0, SEEK_END) != 0 || (*length = ftell(f)) < 0 ||
            fseek(f, NBX_CHECKPOINT_DATA_OFFSET, SEEK_SET) != 0) {
            fclose(f);
            FAIL(N_BODY_SIM_PRO_STATUS_IO_ERROR, "failed to size checkpoint file");
        }
    }
    *out = f;
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_checkpoint_peek(const char* path, NBodySimProCheckpointHeader* header,
                                      NBodySimProError* error) {
    if (path == NULL || header == NULL) {
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "path and header must not be null");
    }
    FILE* f = NULL;
    NBodySimProStatus st = open_checked(path, header, &f, NULL, error);
    if (st != N_BODY_SIM_PRO_STATUS_OK) {
        return st;
    }
    fclose(f);
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_checkpoint_read(const char* path, NBodySimProCheckpointHeader* header,
                                      NBodySimProParticleSystemView* view, size_t capacity,
                                      NBodySimProError* error) {
    if (path == NULL || header == NULL || view == NULL) {
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
             "path, header, and particle view must not be null");
    }
    NBodySimProCheckpointHeader stored;
    FILE* f = NULL;
    long length = 0;
    NBodySimProStatus st = open_checked(path, &stored, &f, &length, error);
    if (st != N_BODY_SIM_PRO_STATUS_OK) {
        return st;
    }
    const size_t n = stored.particle_count;
    if (n > capacity) {
        fclose(f);
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
             "particle system capacity is smaller than checkpoint count");
    }
    if ((unsigned long)length - NBX_CHECKPOINT_DATA_OFFSET < (unsigned long)ARRAYS * sizeof(double) * n) {
        fclose(f);
        FAIL(N_BODY_SIM_PRO_STATUS_IO_ERROR, "checkpoint file is truncated");
    }
    for (int k = 0; k < ARRAYS; ++k) {
        if (array_of(view, k) == NULL) {
            fclose(f);
            FAIL(N_BODY_SIM_PRO_STATUS_INVALID_STATE, "particle system array is null");
        }
    }
    for (int k = 0; k < ARRAYS && n > 0; ++k) {
        if (fread(array_of(view, k), sizeof(double), n, f) != n) {
            fclose(f);
            FAIL(N_BODY_SIM_PRO_STATUS_IO_ERROR, "failed to read checkpoint array");
        }
    }
    if (fclose(f) != 0) {
        FAIL(N_BODY_SIM_PRO_STATUS_IO_ERROR, "failed to close checkpoint file");
    }
    view->particle_count = n;
    *header = stored;
    return N_BODY_SIM_PRO_STATUS_OK;
}

/* Ten host arrays of n doubles in one allocation, as a view. */
static double* host_view(size_t n, NBodySimProParticleSystemView* v) {
    double* block = (double*)malloc((n > 0 ? n : 1) * ARRAYS * sizeof(double));
    if (block == NULL) {
        return NULL;
    }
    double** slots[ARRAYS] = {&v->positions_x,     &v->positions_y,     &v->positions_z,
                              &v->velocities_x,    &v->velocities_y,    &v->velocities_z,
                              &v->accelerations_x, &v->accelerations_y, &v->accelerations_z,
                              &v->masses};
    for (int k = 0; k < ARRAYS; ++k) {
        *slots[k] = block + (size_t)k * n;
    }
    v->particle_count = n;
    return block;
}

NBodySimProStatus nbx_engine_checkpoint_save(NBXEngine* engine, const char* path,
                                             const NBodySimProCheckpointHeader* meta,
                                             NBodySimProError* error) {
    if (engine == NULL || path == NULL || meta == NULL) {
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "engine, path, and header must not be null");
    }
    const size_t n = nbx_engine_particle_count(engine);
    NBodySimProParticleSystemView v;
    double* block = host_view(n, &v);
    if (block == NULL) {
        FAIL(N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY, "checkpoint staging allocation failed");
    }
    NBodySimProStatus st = nbx_engine_download(engine, &v, error);
    if (st == N_BODY_SIM_PRO_STATUS_OK) {
        NBodySimProCheckpointHeader h = *meta;
        h.magic = N_BODY_SIM_PRO_CHECKPOINT_MAGIC;
        h.version = N_BODY_SIM_PRO_CHECKPOINT_VERSION;
        h.particle_count = n;
        st = nbx_checkpoint_write(path, &v, &h, error);
    }
    free(block);
    return st;
}

NBodySimProStatus nbx_engine_checkpoint_load(NBXEngine* engine, const char* path,
                                             NBodySimProCheckpointHeader* header,
                                             NBodySimProError* error) {
    if (engine == NULL || path == NULL || header == NULL) {
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "engine, path, and header must not be null");
    }
    NBodySimProCheckpointHeader peeked;
    NBodySimProStatus st = nbx_checkpoint_peek(path, &peeked, error);
    if (st != N_BODY_SIM_PRO_STATUS_OK) {
        return st;
    }
    if (peeked.particle_count == 0) {
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "checkpoint holds no particles");
    }
    if (peeked.particle_count > nbx_engine_capacity(engine)) {
        FAIL(N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
             "engine capacity is smaller than checkpoint count");
    }
    NBodySimProParticleSystemView v;
    double* block = host_view(peeked.particle_count, &v);
    if (block == NULL) {
        FAIL(N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY, "checkpoint staging allocation failed");
    }
    st = nbx_checkpoint_read(path, header, &v, peeked.particle_count, error);
    if (st == N_BODY_SIM_PRO_STATUS_OK) {
        st = nbx_engine_upload(engine, &v, error);
    }
    free(block);
    return st;
}
