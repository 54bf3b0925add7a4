// nbx_engine.hip - engine lifecycle, host<->device transfer, force dispatch,
// FP64 integrator and diagnostics, parity probes and the drop-in force
// functions of the C ABI (include/nbx_engine.h).
#include "nbx_internal.h"

#include <array>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <new>
#include <vector>

namespace nbx {

void set_error(NBodySimProError* error, NBodySimProStatus status, const char* file, int line,
               const char* message) {
    if (error == nullptr) {
        return;
    }
    error->status = status;
    error->source_file = file;
    error->source_line = line;
    if (message == nullptr) {
        error->message[0] = '\0';
        return;
    }
    const size_t len = std::strlen(message);
    const size_t copy = len < sizeof(error->message) ? len : sizeof(error->message) - 1;
    std::memcpy(error->message, message, copy);
    error->message[copy] = '\0';
}

NBodySimProStatus map_hip_error(hipError_t e) {
    switch (e) {
        case hipSuccess:
            return N_BODY_SIM_PRO_STATUS_OK;
        case hipErrorOutOfMemory:
            return N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY;
        case hipErrorNoDevice:
        case hipErrorInvalidDevice:
        case hipErrorInsufficientDriver:
            return N_BODY_SIM_PRO_STATUS_UNSUPPORTED_PLATFORM;
        default:
            return N_BODY_SIM_PRO_STATUS_INVALID_STATE;
    }
}

NBodySimProStatus ensure(void** ptr, size_t* cap, size_t bytes, NBodySimProError* error) {
    if (bytes <= *cap && *ptr != nullptr) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    if (*ptr != nullptr) {
        NBX_HIP(error, hipFree(*ptr));
        *ptr = nullptr;
        *cap = 0;
    }
    // One eighth of headroom: sizes that follow the tree (cells, nodes) move
    // by a fraction of a percent from step to step, and every new maximum
    // would otherwise cost a hipFree (which waits for the device) and a
    // hipMalloc inside the evaluation. (Measured on MI355X, 1M two-galaxy
    // leapfrog: one build in eight paid about 0.2 ms for it.)
    size_t want = bytes + bytes / 8;
    want = want < 256 ? 256 : want;
    NBX_HIP(error, hipMalloc(ptr, want));
    *cap = want;
    return N_BODY_SIM_PRO_STATUS_OK;
}

namespace {

constexpr int kThreads = 256;

// Every grow-on-demand device buffer of an engine (Barnes-Hut build and walk
// arrays, diagnostics scratch); the array's size follows the list.
auto work_buffers(Engine* e) {
    BHState& s = e->bh;
    return std::array{&s.b4, &s.partial, &s.flags, &s.mkey, &s.tkey, &s.sk0, &s.sk1, &s.sv0, &s.sv1,
            &s.hist, &s.mperm, &s.mkey_sorted, &s.tperm, &s.trank, &s.tword, &s.ext, &s.delta,
            &s.cnt, &s.scan, &s.scan_part, &s.sp4, &s.s4, &s.tile_hist, &s.tile_base, &s.n64,
            &s.n32, &s.icom, &s.irec, &s.idepth, &s.gnode, &s.im2, &s.gquad,
            &s.tmp32, &s.counters, &s.deferred, &e->scratch, &e->scratch2};
}

void free_buf(Buf& b) {
    if (b.p != nullptr) {
        (void)hipFree(b.p);
    }
    b.p = nullptr;
    b.cap = 0;
}

// ---------------------------------------------------------------- integrator
// src/c/physics/integrator.c:64-83 (compiled with contraction off).
__global__ void kick_drift_kernel(double* __restrict__ px, double* __restrict__ py,
                                  double* __restrict__ pz, double* __restrict__ vx,
                                  double* __restrict__ vy, double* __restrict__ vz,
                                  const double* __restrict__ ax, const double* __restrict__ ay,
                                  const double* __restrict__ az, int n, double half, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double nvx = vx[i] + ax[i] * half;
        const double nvy = vy[i] + ay[i] * half;
        const double nvz = vz[i] + az[i] * half;
        vx[i] = nvx;
        vy[i] = nvy;
        vz[i] = nvz;
        px[i] = px[i] + nvx * dt;
        py[i] = py[i] + nvy * dt;
        pz[i] = pz[i] + nvz * dt;
    }
}

__global__ void kick_kernel(double* __restrict__ vx, double* __restrict__ vy,
                            double* __restrict__ vz, const double* __restrict__ ax,
                            const double* __restrict__ ay, const double* __restrict__ az, int n,
                            double half) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        vx[i] = vx[i] + ax[i] * half;
        vy[i] = vy[i] + ay[i] * half;
        vz[i] = vz[i] + az[i] * half;
    }
}

// integrator.c:35-43
__global__ void euler_kernel(double* __restrict__ px, double* __restrict__ py,
                             double* __restrict__ pz, double* __restrict__ vx,
                             double* __restrict__ vy, double* __restrict__ vz,
                             const double* __restrict__ ax, const double* __restrict__ ay,
                             const double* __restrict__ az, int n, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        px[i] = px[i] + vx[i] * dt;
        py[i] = py[i] + vy[i] * dt;
        pz[i] = pz[i] + vz[i] * dt;
        vx[i] = vx[i] + ax[i] * dt;
        vy[i] = vy[i] + ay[i] * dt;
        vz[i] = vz[i] + az[i] * dt;
    }
}

// ---------------------------------------------------------------- diagnostics
constexpr int kDiag = 11;

// Per-block FP64 partial sums of the numerics.c:5-81 terms (fixed order).
__global__ __launch_bounds__(kThreads) void diag_partial_kernel(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
    const double* __restrict__ m, int n, double* __restrict__ partial) {
    __shared__ double s[kDiag][kThreads];
    double acc[kDiag];
#pragma unroll
    for (int k = 0; k < kDiag; ++k) {
        acc[k] = 0.0;
    }
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const double mass = m[i];
        const double x = px[i], y = py[i], z = pz[i];
        const double u = vx[i], v = vy[i], w = vz[i];
        acc[0] += 0.5 * mass * (u * u + v * v + w * w);
        acc[1] += mass;
        acc[2] += mass * u;
        acc[3] += mass * v;
        acc[4] += mass * w;
        acc[5] += mass * (y * w - z * v);
        acc[6] += mass * (z * u - x * w);
        acc[7] += mass * (x * v - y * u);
        acc[8] += mass * x;
        acc[9] += mass * y;
        acc[10] += mass * z;
    }
#pragma unroll
    for (int k = 0; k < kDiag; ++k) {
        s[k][threadIdx.x] = acc[k];
    }
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < static_cast<unsigned>(off)) {
#pragma unroll
            for (int k = 0; k < kDiag; ++k) {
                s[k][threadIdx.x] += s[k][threadIdx.x + off];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < kDiag) {
        partial[blockIdx.x * kDiag + threadIdx.x] = s[threadIdx.x][0];
    }
}

// Softened potential: per-target partial sums over j > i (numerics.c:83-116).
__global__ __launch_bounds__(kThreads) void potential_partial_kernel(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ m, int n, double eps2, double* __restrict__ partial) {
    __shared__ double s[kThreads];
    const int i = blockIdx.x * kThreads + threadIdx.x;
    double sum = 0.0;
    if (i < n) {
        const double x = px[i], y = py[i], z = pz[i], mi = m[i];
        for (int j = i + 1; j < n; ++j) {
            const double dx = px[j] - x;
            const double dy = py[j] - y;
            const double dz = pz[j] - z;
            const double d = sqrt(dx * dx + dy * dy + dz * dz + eps2);
            sum += mi * m[j] / d;
        }
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < static_cast<unsigned>(off)) {
            s[threadIdx.x] += s[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s[0];
    }
}

// Single-block fixed-order reduction of `count` rows of `width` partials.
__global__ __launch_bounds__(kThreads) void reduce_rows_kernel(const double* __restrict__ partial,
                                                              int count, int width,
                                                              double* __restrict__ out) {
    __shared__ double s[kThreads];
    for (int k = 0; k < width; ++k) {
        double acc = 0.0;
        for (int r = threadIdx.x; r < count; r += kThreads) {
            acc += partial[r * width + k];
        }
        s[threadIdx.x] = acc;
        __syncthreads();
        for (int off = kThreads / 2; off > 0; off >>= 1) {
            if (threadIdx.x < static_cast<unsigned>(off)) {
                s[threadIdx.x] += s[threadIdx.x + off];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            out[k] = s[0];
        }
        __syncthreads();
    }
}

NBodySimProStatus check_engine(NBXEngine* engine, NBodySimProError* error) {
    if (engine == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "engine must not be null");
    }
    NBX_HIP(error, hipSetDevice(engine->e.device));
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus copy_array(double* dst, const double* src, size_t n, hipMemcpyKind kind,
                             hipStream_t s, NBodySimProError* error) {
    if (src == nullptr || dst == nullptr || n == 0) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    NBX_HIP(error, hipMemcpyAsync(dst, src, n * sizeof(double), kind, s));
    return N_BODY_SIM_PRO_STATUS_OK;
}

std::mutex g_default_mutex;
NBXEngine* g_default = nullptr;

}  // namespace

NBodySimProStatus compute_force(Engine* e, NBXAlgorithm algorithm, const NBodySimProGravity* g,
                                NBodySimProError* error) {
    switch (algorithm) {
        case NBX_ALGORITHM_DIRECT_FP64:
            return direct_fp64(e, g, error);
        case NBX_ALGORITHM_DIRECT_FP32:
            return direct_fp32(e, g, error);
        case NBX_ALGORITHM_BARNES_HUT:
            return barnes_hut(e, g, false, error);
        case NBX_ALGORITHM_BARNES_HUT_FP64:
            return barnes_hut(e, g, true, error);
        case NBX_ALGORITHM_BH_LEARNED:
            return bh_learned(e, g, error);
    }
    NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "unknown force algorithm");
}

}  // namespace nbx

using nbx::Engine;

extern "C" {

int nbx_device_count(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        return 0;
    }
    return count;
}

int nbx_device_info(int device, char* name, size_t len, int* compute_units) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        return -1;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        return -1;
    }
    if (name != nullptr && len > 0) {
        std::snprintf(name, len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units != nullptr) {
        *compute_units = prop.multiProcessorCount;
    }
    return 0;
}

int nbx_algorithm_from_name(const char* name) {
    if (name == nullptr) {
        return -1;
    }
    static const char* names[] = {"hip_direct_fp64", "hip_direct_fp32", "hip_barnes_hut",
                                  "hip_barnes_hut_fp64", "hip_bh_learned"};
    for (int k = 0; k < 5; ++k) {
        if (std::strcmp(name, names[k]) == 0) {
            return k;
        }
    }
    return -1;
}

const char* nbx_algorithm_name(int algorithm) {
    static const char* names[] = {"hip_direct_fp64", "hip_direct_fp32", "hip_barnes_hut",
                                  "hip_barnes_hut_fp64", "hip_bh_learned"};
    return (algorithm >= 0 && algorithm < 5) ? names[algorithm] : nullptr;
}

NBXEngine* nbx_engine_create(const NBXConfig* config, NBodySimProError* error) {
    if (config == nullptr || config->capacity == 0) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, __FILE__, __LINE__,
                       "config must be non-null with capacity > 0");
        return nullptr;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_UNSUPPORTED_PLATFORM, __FILE__, __LINE__,
                       "no HIP device is visible");
        return nullptr;
    }
    if (config->device < 0 || config->device >= count) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, __FILE__, __LINE__,
                       "device ordinal out of range");
        return nullptr;
    }
    NBXEngine* engine = new (std::nothrow) NBXEngine();
    if (engine == nullptr) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY, __FILE__, __LINE__,
                       "engine allocation failed");
        return nullptr;
    }
    Engine* e = &engine->e;
    e->device = config->device;
    e->capacity = config->capacity;
    e->theta = config->theta > 0.0 ? config->theta : 0.7;
    e->flags = config->flags;
    bool ok = hipSetDevice(e->device) == hipSuccess &&
              hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) == hipSuccess;
    int cus = 0;
    if (ok && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device) ==
                  hipSuccess &&
        cus > 0) {
        e->cu_count = cus;
    }
    double** arrays[] = {&e->px, &e->py, &e->pz, &e->vx, &e->vy, &e->vz,
                         &e->ax, &e->ay, &e->az, &e->m};
    for (double** a : arrays) {
        ok = ok && hipMalloc(reinterpret_cast<void**>(a), e->capacity * sizeof(double)) == hipSuccess;
    }
    ok = ok && hipMalloc(reinterpret_cast<void**>(&e->pos4), e->capacity * sizeof(float4)) == hipSuccess;
    for (auto& ev : e->ev) {
        ok = ok && hipEventCreate(&ev) == hipSuccess;
    }
    if (!ok) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY, __FILE__, __LINE__,
                       "device allocation failed while creating the engine");
        nbx_engine_destroy(engine);
        return nullptr;
    }
    return engine;
}

void nbx_mlp_release(void* mlp);

void nbx_engine_destroy(NBXEngine* engine) {
    if (engine == nullptr) {
        return;
    }
    Engine* e = &engine->e;
    (void)hipSetDevice(e->device);
    if (e->stream != nullptr) {
        (void)hipStreamSynchronize(e->stream);
    }
    double* arrays[] = {e->px, e->py, e->pz, e->vx, e->vy, e->vz, e->ax, e->ay, e->az, e->m};
    for (double* a : arrays) {
        if (a != nullptr) (void)hipFree(a);
    }
    if (e->pos4 != nullptr) (void)hipFree(e->pos4);
    for (nbx::Buf* b : nbx::work_buffers(e)) {
        nbx::free_buf(*b);
    }
    if (e->bh.host_flags != nullptr) (void)hipHostFree(e->bh.host_flags);
    if (e->mlp != nullptr) {
        nbx_mlp_release(e->mlp);
    }
    for (auto& ev : e->ev) {
        if (ev != nullptr) (void)hipEventDestroy(ev);
    }
    if (e->stream != nullptr) (void)hipStreamDestroy(e->stream);
    delete engine;
}

size_t nbx_engine_device_bytes(const NBXEngine* engine) {
    if (engine == nullptr) {
        return 0;
    }
    Engine* e = const_cast<Engine*>(&engine->e);
    size_t bytes = e->capacity * (10 * sizeof(double) + sizeof(float4));
    for (const nbx::Buf* b : nbx::work_buffers(e)) {
        bytes += b->cap;
    }
    return bytes;
}

void nbx_engine_set_theta(NBXEngine* engine, double theta) {
    if (engine != nullptr) {
        engine->e.theta = theta;
    }
}

double nbx_engine_theta(const NBXEngine* engine) {
    return engine == nullptr ? 0.0 : engine->e.theta;
}

NBodySimProStatus nbx_engine_upload(NBXEngine* engine, const NBodySimProParticleSystemView* view,
                                    NBodySimProError* error) {
    NBX_TRY(nbx::check_engine(engine, error));
    if (view == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "view must not be null");
    }
    Engine* e = &engine->e;
    const size_t n = view->particle_count;
    if (n > e->capacity) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INSUFFICIENT_MEMORY,
                 "particle count exceeds the engine capacity");
    }
    if (n > 0 && (view->positions_x == nullptr || view->positions_y == nullptr ||
                  view->positions_z == nullptr || view->masses == nullptr)) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "view positions/masses are null");
    }
    const hipMemcpyKind k = hipMemcpyHostToDevice;
    NBX_TRY(nbx::copy_array(e->px, view->positions_x, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(e->py, view->positions_y, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(e->pz, view->positions_z, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(e->m, view->masses, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(e->vx, view->velocities_x, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(e->vy, view->velocities_y, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(e->vz, view->velocities_z, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(e->ax, view->accelerations_x, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(e->ay, view->accelerations_y, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(e->az, view->accelerations_z, n, k, e->stream, error));
    NBX_HIP(error, hipStreamSynchronize(e->stream));
    e->n = n;
    e->pos4_valid = false;
    return N_BODY_SIM_PRO_STATUS_OK;
}

size_t nbx_engine_particle_count(const NBXEngine* engine) {
    return engine != nullptr ? engine->e.n : 0;
}

size_t nbx_engine_capacity(const NBXEngine* engine) {
    return engine != nullptr ? engine->e.capacity : 0;
}

NBodySimProStatus nbx_engine_download(NBXEngine* engine, const NBodySimProParticleSystemView* view,
                                      NBodySimProError* error) {
    NBX_TRY(nbx::check_engine(engine, error));
    if (view == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "view must not be null");
    }
    Engine* e = &engine->e;
    if (view->particle_count != e->n) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                 "view particle count differs from the uploaded count");
    }
    const size_t n = e->n;
    const hipMemcpyKind k = hipMemcpyDeviceToHost;
    NBX_TRY(nbx::copy_array(view->positions_x, e->px, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(view->positions_y, e->py, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(view->positions_z, e->pz, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(view->masses, e->m, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(view->velocities_x, e->vx, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(view->velocities_y, e->vy, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(view->velocities_z, e->vz, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(view->accelerations_x, e->ax, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(view->accelerations_y, e->ay, n, k, e->stream, error));
    NBX_TRY(nbx::copy_array(view->accelerations_z, e->az, n, k, e->stream, error));
    NBX_HIP(error, hipStreamSynchronize(e->stream));
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_engine_compute_force(NBXEngine* engine, NBXAlgorithm algorithm,
                                           const NBodySimProGravity* gravity,
                                           NBodySimProError* error) {
    NBX_TRY(nbx::check_engine(engine, error));
    if (gravity == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "gravity must not be null");
    }
    return nbx::compute_force(&engine->e, algorithm, gravity, error);
}

NBodySimProStatus nbx_engine_advance(NBXEngine* engine, NBXAlgorithm algorithm,
                                     const NBodySimProGravity* gravity,
                                     NBodySimProIntegratorType integrator, double timestep,
                                     int steps, NBodySimProError* error) {
    NBX_TRY(nbx::check_engine(engine, error));
    if (gravity == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "gravity must not be null");
    }
    if (!(timestep > 0.0)) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "timestep must be positive");
    }
    if (steps < 0) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "steps must be non-negative");
    }
    if (integrator != N_BODY_SIM_PRO_INTEGRATOR_EULER &&
        integrator != N_BODY_SIM_PRO_INTEGRATOR_LEAPFROG &&
        integrator != N_BODY_SIM_PRO_INTEGRATOR_VELOCITY_VERLET) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "unknown integrator type");
    }
    Engine* e = &engine->e;
    const int n = static_cast<int>(e->n);
    if (n == 0) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    hipEvent_t t0, t1;
    NBX_HIP(error, hipEventCreate(&t0));
    NBX_HIP(error, hipEventCreate(&t1));
    NBX_HIP(error, hipEventRecord(t0, e->stream));
    const unsigned g = nbx::grid_for(e->n, nbx::kThreads);
    const double half = 0.5 * timestep;
    NBodySimProStatus st = N_BODY_SIM_PRO_STATUS_OK;
    for (int k = 0; k < steps && st == N_BODY_SIM_PRO_STATUS_OK; ++k) {
        if (integrator == N_BODY_SIM_PRO_INTEGRATOR_EULER) {
            st = nbx::compute_force(e, algorithm, gravity, error);
            if (st != N_BODY_SIM_PRO_STATUS_OK) break;
            nbx::euler_kernel<<<g, nbx::kThreads, 0, e->stream>>>(e->px, e->py, e->pz, e->vx, e->vy,
                                                                  e->vz, e->ax, e->ay, e->az, n,
                                                                  timestep);
            e->pos4_valid = false;
        } else {
            nbx::kick_drift_kernel<<<g, nbx::kThreads, 0, e->stream>>>(
                e->px, e->py, e->pz, e->vx, e->vy, e->vz, e->ax, e->ay, e->az, n, half, timestep);
            e->pos4_valid = false;
            st = nbx::compute_force(e, algorithm, gravity, error);
            if (st != N_BODY_SIM_PRO_STATUS_OK) break;
            nbx::kick_kernel<<<g, nbx::kThreads, 0, e->stream>>>(e->vx, e->vy, e->vz, e->ax, e->ay,
                                                                 e->az, n, half);
        }
        if (hipGetLastError() != hipSuccess) {
            nbx::set_error(error, N_BODY_SIM_PRO_STATUS_INVALID_STATE, __FILE__, __LINE__,
                           "integrator kernel launch failed");
            st = N_BODY_SIM_PRO_STATUS_INVALID_STATE;
        }
    }
    const NBXTimings keep = e->timings;
    (void)hipEventRecord(t1, e->stream);
    (void)hipEventSynchronize(t1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    e->timings = keep;
    e->timings.integrate_ms = ms;
    return st;
}

NBodySimProStatus nbx_engine_diagnostics(NBXEngine* engine, const NBodySimProGravity* gravity,
                                         NBodySimProDiagnosticsQuantities* q, double* potential,
                                         NBodySimProError* error) {
    NBX_TRY(nbx::check_engine(engine, error));
    if (q == nullptr || (potential != nullptr && gravity == nullptr)) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                 "quantities (and gravity when a potential is requested) must not be null");
    }
    Engine* e = &engine->e;
    const int n = static_cast<int>(e->n);
    const unsigned blocks = n == 0 ? 1u : (nbx::grid_for(e->n, nbx::kThreads) < 512u
                                               ? nbx::grid_for(e->n, nbx::kThreads) : 512u);
    const unsigned pblocks = n == 0 ? 1u : nbx::grid_for(e->n, nbx::kThreads);
    NBX_TRY(e->scratch.need<double>(blocks * nbx::kDiag + 16, error));
    NBX_TRY(e->scratch2.need<double>(pblocks + 16, error));
    double* part = e->scratch.as<double>();
    double* out = part + blocks * nbx::kDiag;
    double sums[nbx::kDiag];
    nbx::diag_partial_kernel<<<blocks, nbx::kThreads, 0, e->stream>>>(
        e->px, e->py, e->pz, e->vx, e->vy, e->vz, e->m, n, part);
    nbx::reduce_rows_kernel<<<1, nbx::kThreads, 0, e->stream>>>(part, static_cast<int>(blocks),
                                                                nbx::kDiag, out);
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipMemcpyAsync(sums, out, sizeof(sums), hipMemcpyDeviceToHost, e->stream));
    double u = 0.0;
    if (potential != nullptr) {
        double* pp = e->scratch2.as<double>();
        nbx::potential_partial_kernel<<<pblocks, nbx::kThreads, 0, e->stream>>>(
            e->px, e->py, e->pz, e->m, n, gravity->softening_squared, pp);
        nbx::reduce_rows_kernel<<<1, nbx::kThreads, 0, e->stream>>>(pp, static_cast<int>(pblocks), 1,
                                                                    pp + pblocks);
        NBX_HIP(error, hipGetLastError());
        NBX_HIP(error, hipMemcpyAsync(&u, pp + pblocks, sizeof(double), hipMemcpyDeviceToHost,
                                      e->stream));
    }
    NBX_HIP(error, hipStreamSynchronize(e->stream));
    q->kinetic_energy = sums[0];
    q->total_mass = sums[1];
    q->total_momentum_x = sums[2];
    q->total_momentum_y = sums[3];
    q->total_momentum_z = sums[4];
    q->total_angular_momentum_x = sums[5];
    q->total_angular_momentum_y = sums[6];
    q->total_angular_momentum_z = sums[7];
    if (sums[1] != 0.0) {
        q->center_of_mass_x = sums[8] / sums[1];
        q->center_of_mass_y = sums[9] / sums[1];
        q->center_of_mass_z = sums[10] / sums[1];
    } else {
        q->center_of_mass_x = q->center_of_mass_y = q->center_of_mass_z = 0.0;
    }
    if (potential != nullptr) {
        *potential = -gravity->gravitational_constant * u;
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

int nbx_engine_stats(const NBXEngine* engine, NBodySimProBarnesHutStats* stats) {
    if (engine == nullptr || stats == nullptr) {
        return 1;
    }
    *stats = engine->e.stats;
    return 0;
}

int nbx_engine_timings(const NBXEngine* engine, NBXTimings* timings) {
    if (engine == nullptr || timings == nullptr) {
        return 1;
    }
    *timings = engine->e.timings;
    return 0;
}

NBodySimProStatus nbx_engine_bh_morton(NBXEngine* engine, double* root, uint64_t* keys,
                                       int64_t* permutation, NBodySimProError* error) {
    NBX_TRY(nbx::check_engine(engine, error));
    Engine* e = &engine->e;
    const size_t n = e->bh.n;
    if (n == 0) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_STATE, "no Barnes-Hut build to report");
    }
    if (root != nullptr) {
        for (int k = 0; k < 4; ++k) root[k] = e->bh.root[k];
    }
    if (keys != nullptr) {
        NBX_HIP(error, hipMemcpy(keys, e->bh.mkey.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost));
    }
    if (permutation != nullptr) {
        std::vector<uint32_t> p(n);
        NBX_HIP(error, hipMemcpy(p.data(), e->bh.mperm.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) permutation[i] = static_cast<int64_t>(p[i]);
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_engine_bh_export_nodes(NBXEngine* engine, double* rows,
                                             NBodySimProError* error) {
    NBX_TRY(nbx::check_engine(engine, error));
    Engine* e = &engine->e;
    nbx::BHState& s = e->bh;
    if (s.n == 0 || rows == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_STATE, "no Barnes-Hut build to export");
    }
    const size_t n = s.n, nodes = s.nodes, internal = s.internal;
    std::vector<double> cx(nodes), cy(nodes), cz(nodes), m(nodes);
    std::vector<int32_t> depth(nodes), start(nodes), count(nodes), leaf(nodes);
    std::vector<uint64_t> word(n), ext;
    std::vector<uint32_t> trank(n);
    {
        // internal nodes: records by internal id, placed at their node index
        std::vector<double> com(4 * internal);
        std::vector<int32_t> rec(4 * internal), idepth(internal);
        NBX_HIP(error, hipMemcpy(com.data(), s.icom.p, internal * 32, hipMemcpyDeviceToHost));
        NBX_HIP(error, hipMemcpy(rec.data(), s.irec.p, internal * 16, hipMemcpyDeviceToHost));
        NBX_HIP(error, hipMemcpy(idepth.data(), s.idepth.p, internal * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < internal; ++i) {
            const size_t k = static_cast<size_t>(rec[4 * i]);
            cx[k] = com[4 * i];
            cy[k] = com[4 * i + 1];
            cz[k] = com[4 * i + 2];
            m[k] = com[4 * i + 3];
            depth[k] = idepth[i];
            start[k] = rec[4 * i + 2];
            count[k] = rec[4 * i + 3] - rec[4 * i + 2] + 1;   // last - first + 1
        }
    }
    {
        NBX_TRY(nbx::bh_node_records(e, error));
        NBX_HIP(error, hipStreamSynchronize(e->stream));
        std::vector<nbx::NodeF32> recs(nodes);
        NBX_HIP(error, hipMemcpy(recs.data(), s.n32.p, nodes * sizeof(nbx::NodeF32),
                                 hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nodes; ++k) leaf[k] = recs[k].leaf;
    }
    NBX_HIP(error, hipMemcpy(word.data(), s.tword.p, n * 8, hipMemcpyDeviceToHost));
    NBX_HIP(error, hipMemcpy(trank.data(), s.trank.p, n * 4, hipMemcpyDeviceToHost));
    // A leaf's row comes from its tree position t: depth one below the deeper
    // neighbour cell, the particle itself as mass and centre of mass.
    std::vector<int32_t> delta(n);
    std::vector<double> sp(4 * n);
    NBX_HIP(error, hipMemcpy(delta.data(), s.delta.p, n * 4, hipMemcpyDeviceToHost));
    if (s.inputs_fp64) {
        NBX_HIP(error, hipMemcpy(sp.data(), s.sp4.p, n * 32, hipMemcpyDeviceToHost));
    } else {
        // FP32 inputs: the tree-order float4 values, widened exactly
        std::vector<float> sf(4 * n);
        NBX_HIP(error, hipMemcpy(sf.data(), s.s4.p, n * 16, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < 4 * n; ++k) sp[k] = static_cast<double>(sf[k]);
    }
    for (size_t k = 0; k < nodes; ++k) {
        const int t = leaf[k];
        if (t < 0) {
            continue;
        }
        const int a = t > 0 ? delta[t - 1] : -1;
        depth[k] = 1 + (a > delta[t] ? a : delta[t]);
        start[k] = t;
        count[k] = 1;
        cx[k] = sp[4 * t];
        cy[k] = sp[4 * t + 1];
        cz[k] = sp[4 * t + 2];
        m[k] = sp[4 * t + 3];
    }
    if (s.tie_path) {
        ext.resize(n * 6);
        NBX_HIP(error, hipMemcpy(ext.data(), static_cast<uint64_t*>(s.ext.p) + 2 * n * 6,
                                 n * 6 * 8, hipMemcpyDeviceToHost));
    }
    for (size_t k = 0; k < nodes; ++k) {
        // Replay the cell centre from the root along the node's octant path
        // with the reference's FP64 operations (barnes_hut.c:202-211).
        double c[3] = {s.root[0], s.root[1], s.root[2]};
        double h = s.root[3];
        const int t = start[k];
        for (int d = 0; d < depth[k]; ++d) {
            uint64_t w = d < 21 ? word[t] : ext[static_cast<size_t>(t) * 6 + (d / 21 - 1)];
            const int o = static_cast<int>((w >> (60 - 3 * (d % 21))) & 7u);
            const double off = h * 0.5;
            c[0] = c[0] + ((o & 1) ? off : -off);
            c[1] = c[1] + ((o & 2) ? off : -off);
            c[2] = c[2] + ((o & 4) ? off : -off);
            h = 0.5 * h;
        }
        double* r = rows + 11 * k;
        r[0] = depth[k];
        r[1] = c[0];
        r[2] = c[1];
        r[3] = c[2];
        r[4] = h;
        r[5] = count[k];
        r[6] = m[k];
        r[7] = cx[k];
        r[8] = cy[k];
        r[9] = cz[k];
        r[10] = leaf[k] >= 0 ? static_cast<double>(trank[leaf[k]]) : -1.0;
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

static NBodySimProStatus dropin(const NBodySimProParticleSystemView* view,
                                const NBodySimProGravity* gravity, void* context,
                                NBodySimProError* error, NBXAlgorithm algorithm) {
    if (view == nullptr || gravity == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                 "view and gravity parameters must not be null");
    }
    const size_t n = view->particle_count;
    if (n == 0) {
        if (algorithm == NBX_ALGORITHM_BARNES_HUT || algorithm == NBX_ALGORITHM_BARNES_HUT_FP64 ||
            algorithm == NBX_ALGORITHM_BH_LEARNED) {
            NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                     "cannot build a Barnes-Hut tree for an empty system");
        }
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    std::unique_lock<std::mutex> lock(nbx::g_default_mutex, std::defer_lock);
    NBXEngine* engine = static_cast<NBXEngine*>(context);
    if (engine == nullptr) {
        lock.lock();
        if (nbx::g_default != nullptr && nbx::g_default->e.capacity < n) {
            nbx_engine_destroy(nbx::g_default);
            nbx::g_default = nullptr;
        }
        if (nbx::g_default == nullptr) {
            NBXConfig cfg{n, 0, 0.7, 0u};
            nbx::g_default = nbx_engine_create(&cfg, error);
            if (nbx::g_default == nullptr) {
                return error != nullptr ? error->status : N_BODY_SIM_PRO_STATUS_INVALID_STATE;
            }
        }
        engine = nbx::g_default;
    }
    NBodySimProParticleSystemView in{};
    in.particle_count = n;
    in.positions_x = view->positions_x;
    in.positions_y = view->positions_y;
    in.positions_z = view->positions_z;
    in.masses = view->masses;
    NBX_TRY(nbx_engine_upload(engine, &in, error));
    NBX_TRY(nbx::compute_force(&engine->e, algorithm, gravity, error));
    NBodySimProParticleSystemView out{};
    out.particle_count = n;
    out.accelerations_x = view->accelerations_x;
    out.accelerations_y = view->accelerations_y;
    out.accelerations_z = view->accelerations_z;
    return nbx_engine_download(engine, &out, error);
}

NBodySimProStatus nbx_hip_direct_fp64(const NBodySimProParticleSystemView* view,
                                      const NBodySimProGravity* gravity, void* context,
                                      NBodySimProError* error) {
    return dropin(view, gravity, context, error, NBX_ALGORITHM_DIRECT_FP64);
}

NBodySimProStatus nbx_hip_direct_fp32(const NBodySimProParticleSystemView* view,
                                      const NBodySimProGravity* gravity, void* context,
                                      NBodySimProError* error) {
    return dropin(view, gravity, context, error, NBX_ALGORITHM_DIRECT_FP32);
}

NBodySimProStatus nbx_hip_barnes_hut(const NBodySimProParticleSystemView* view,
                                     const NBodySimProGravity* gravity, void* context,
                                     NBodySimProError* error) {
    return dropin(view, gravity, context, error, NBX_ALGORITHM_BARNES_HUT);
}

NBodySimProStatus nbx_hip_barnes_hut_fp64(const NBodySimProParticleSystemView* view,
                                          const NBodySimProGravity* gravity, void* context,
                                          NBodySimProError* error) {
    return dropin(view, gravity, context, error, NBX_ALGORITHM_BARNES_HUT_FP64);
}

NBodySimProStatus nbx_hip_bh_learned(const NBodySimProParticleSystemView* view,
                                     const NBodySimProGravity* gravity, void* context,
                                     NBodySimProError* error) {
    return dropin(view, gravity, context, error, NBX_ALGORITHM_BH_LEARNED);
}

}  // extern "C"
