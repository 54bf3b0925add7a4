// nbx_direct.hip - all-pairs gravitational accelerations on gfx950.
//
// direct_fp64: bit-identical to the scalar reference kernel
//   (src/c/physics/gravity.c:37-67). Every pair term is formed with the
//   reference's operations in the reference's order (this file is compiled
//   with -ffp-contract=off, FP64 sqrt and division are IEEE correctly
//   rounded), and each target's three sums are accumulated in ascending j
//   with j == i skipped. Only the term formation is parallel: four waves
//   compute a 64-target x 32-source tile of terms into LDS, then three waves
//   (one per component) add them in order. This keeps the serial dependency
//   (one add per term) off the expensive part (div + sqrt per term).
//
// direct_fp32: FP32 pair arithmetic on FP32-rounded positions/masses with a
//   wave-uniform source index, so every source is fetched once per wave by
//   scalar loads and broadcast through SGPR operands (no LDS traffic). Each
//   lane owns one target; its sources are split over S waves of the block.
//   Partial sums of T sources are folded into a per-target compensated
//   (Kahan-Babuska) accumulator once per tile, and the S ranges' accumulators
//   are combined in LDS in range order. G multiplies each target's sum once,
//   in FP64, so a pair costs 12 full-rate VALU instructions and one
//   v_rsq_f32.
#include <mutex>

#include "nbx_internal.h"
#include "nbx_learned.h"

namespace nbx {
namespace {

constexpr int kExactTargets = 64;
constexpr int kExactSources = 16;

// One tile of terms: kExactSources sources (from j0) x the block's 64
// targets, each of the block's WAVES waves forming kExactSources / WAVES
// sources' terms with the reference's operations. STAGE: the tile's sources
// (x, y, z, m) are read from the LDS stage `src`; otherwise from memory (the
// last source repeated past n).
template <int WAVES, bool STAGE>
__device__ inline void exact_terms(double (*terms)[kExactSources][kExactTargets],
                                   const double (*src)[kExactSources],
                                   const double* __restrict__ px, const double* __restrict__ py,
                                   const double* __restrict__ pz, const double* __restrict__ m,
                                   int n, int j0, int wave, int lane, double xi, double yi,
                                   double zi, double G, double eps2) {
    constexpr int kPerWave = kExactSources / WAVES;
    static_assert(kExactSources % WAVES == 0, "whole sources per wave");
#pragma unroll
    for (int k = 0; k < kPerWave; ++k) {
        const int jj = wave * kPerWave + k;
        double sx, sy, sz, sm;
        if (STAGE) {
            sx = src[0][jj];
            sy = src[1][jj];
            sz = src[2][jj];
            sm = src[3][jj];
        } else {
            const int j = j0 + jj;
            const int jc = j < n ? j : n - 1;
            sx = px[jc];
            sy = py[jc];
            sz = pz[jc];
            sm = m[jc];
        }
        const double dx = sx - xi;
        const double dy = sy - yi;
        const double dz = sz - zi;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        const double inv = 1.0 / sqrt(r2);
        const double inv3 = inv * inv * inv;
        const double s = G * sm * inv3;
        terms[0][jj][lane] = s * dx;
        terms[1][jj][lane] = s * dy;
        terms[2][jj][lane] = s * dz;
    }
}

// Loads (x, y, z, m) of the tile's kExactSources sources starting at j0 (the
// last source repeated past n), one value per lane of the calling wave:
// lane = field * kExactSources + source.
static_assert(4 * kExactSources == kWave, "one staged source value per lane");
__device__ inline double stage_value(const double* __restrict__ px, const double* __restrict__ py,
                                     const double* __restrict__ pz, const double* __restrict__ m,
                                     int n, int j0, int lane) {
    const int f = lane / kExactSources;
    const int j = j0 + lane % kExactSources;
    const int jc = j < n ? j : n - 1;
    const double* a = f == 0 ? px : (f == 1 ? py : (f == 2 ? pz : m));
    return a[jc];
}

// SUBSET: the block's targets are targets[k] for k = blockIdx.x * 64 + lane < count,
// and the accelerations go to out[3k .. 3k+2].
//
// WAVES waves form the terms of a 16-source tile while waves 0-2 add the
// previous tile's terms of one component each, in source order; two term
// buffers alternate, one barrier per tile. (The serial sums are one add per
// term; the terms, an IEEE divide and square root each, are spread over the
// block's waves, which also hides their FP64 latency.) Sixteen waves suit
// grids of under two blocks per CU, eight the larger ones; measured on
// MI355X, ms per evaluation at N = 16384 / 65536 / 262144:
//   16 waves: 1.18 / 11.96 / 179.9;  8 waves: 1.35 / 11.93 / 156.8;
//   the earlier single-buffered 4-wave kernel: 1.69 / - / 192.
// Loading a tile's 16 terms before its 16 adds (instead of one LDS wait per
// add): 1.16 -> 0.82 / 11.9 -> 9.9 / 157 -> 136 ms with the dispatch above.
// STAGE (from 32768 sources on): the sources come through an LDS stage that
// the last wave fills one tile ahead (9.9 -> 9.4 ms at 65536, 136 -> 131 ms
// at 262144, the 8192-target checker at 10M 0.57 -> 0.45 s; at 16384, where
// the sources stay in cache, 0.82 -> 0.85 ms, so it is off there).
template <bool SUBSET, int WAVES, bool STAGE>
__global__ __launch_bounds__(WAVES * 64) void direct_fp64_kernel(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ m, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, int n, double G, double eps2, const int64_t* __restrict__ targets,
    int count, double* __restrict__ out) {
    __shared__ double terms[2][3][kExactSources][kExactTargets];
    // sources of the next two tiles: the last wave loads tile t + 2's while
    // the block forms tile t + 1's terms, so no wave waits on a global load
    // after a barrier
    __shared__ double src[2][4][kExactSources];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
    const int k = blockIdx.x * kExactTargets + lane;
    const bool valid = SUBSET ? k < count : k < n;
    const int i = !valid ? -1 : (SUBSET ? static_cast<int>(targets[k]) : k);
    const int ic = valid ? i : n - 1;
    const double xi = px[ic], yi = py[ic], zi = pz[ic];
    double acc = 0.0;
    const int tiles = (n + kExactSources - 1) / kExactSources;
    constexpr int kLoader = WAVES - 1;
    if (STAGE && wave == kLoader) {
        (&src[0][0][0])[lane] = stage_value(px, py, pz, m, n, 0, lane);
        (&src[1][0][0])[lane] = stage_value(px, py, pz, m, n, kExactSources, lane);
    }
    __syncthreads();
    exact_terms<WAVES, STAGE>(terms[0], src[0], px, py, pz, m, n, 0, wave, lane, xi, yi, zi, G,
                              eps2);
    __syncthreads();
    for (int t = 0; t < tiles; ++t) {
        const int j0 = t * kExactSources;
        if (wave < 3) {
            // in-order sum of tile t, one component per wave
            const int jn = (n - j0) < kExactSources ? (n - j0) : kExactSources;
            const double (*row)[kExactTargets] = terms[t & 1][wave];
            if (jn == kExactSources) {
                // all of the tile's loads first, then the adds: one LDS
                // latency per tile instead of one per term. The own term
                // adds +0.0, which leaves the sum bit for bit as skipping it
                // would (the sum starts at +0.0, and +0.0 + -0.0 is +0.0,
                // so it is never -0.0).
                double v[kExactSources];
#pragma unroll
                for (int jj = 0; jj < kExactSources; ++jj) {
                    v[jj] = row[jj][lane];
                }
#pragma unroll
                for (int jj = 0; jj < kExactSources; ++jj) {
                    acc += (j0 + jj != i) ? v[jj] : 0.0;
                }
            } else {
                for (int jj = 0; jj < jn; ++jj) {
                    if (j0 + jj != i) {
                        acc += row[jj][lane];
                    }
                }
            }
        }
        if (t + 1 < tiles) {
            // tile t + 2's sources go to the stage tile t used
            double next = 0.0;
            if (STAGE && wave == kLoader) {
                next = stage_value(px, py, pz, m, n, j0 + 2 * kExactSources, lane);
            }
            exact_terms<WAVES, STAGE>(terms[(t + 1) & 1], src[(t + 1) & 1], px, py, pz, m, n,
                                      j0 + kExactSources, wave, lane, xi, yi, zi, G, eps2);
            if (STAGE && wave == kLoader) {
                (&src[t & 1][0][0])[lane] = next;
            }
        }
        __syncthreads();
    }
    if (valid && wave < 3) {
        if (SUBSET) {
            out[3 * static_cast<size_t>(k) + wave] = acc;
        } else {
            double* o = wave == 0 ? ax : (wave == 1 ? ay : az);
            o[i] = acc;
        }
    }
}

// Launches the FP64 all-pairs kernel over `count` targets (all n particles,
// or the SUBSET `targets`) with the wave count that suits the grid.
template <bool SUBSET>
void launch_direct_fp64(hipStream_t st, int cu_count, const double* px, const double* py,
                        const double* pz, const double* m, double* ax, double* ay, double* az,
                        int n, double G, double eps2, const int64_t* targets, int count,
                        double* out) {
    const unsigned blocks = grid_for(static_cast<size_t>(count), kExactTargets);
    const bool stage = n >= 32768;
    if (blocks < 2u * static_cast<unsigned>(cu_count)) {
        if (stage) {
            direct_fp64_kernel<SUBSET, 16, true><<<blocks, 16 * 64, 0, st>>>(
                px, py, pz, m, ax, ay, az, n, G, eps2, targets, count, out);
        } else {
            direct_fp64_kernel<SUBSET, 16, false><<<blocks, 16 * 64, 0, st>>>(
                px, py, pz, m, ax, ay, az, n, G, eps2, targets, count, out);
        }
    } else {
        if (stage) {
            direct_fp64_kernel<SUBSET, 8, true><<<blocks, 8 * 64, 0, st>>>(
                px, py, pz, m, ax, ay, az, n, G, eps2, targets, count, out);
        } else {
            direct_fp64_kernel<SUBSET, 8, false><<<blocks, 8 * 64, 0, st>>>(
                px, py, pz, m, ax, ay, az, n, G, eps2, targets, count, out);
        }
    }
}

__device__ inline void two_sum_add(float& s, float& c, float v) {
    // Kahan-Babuska (Neumaier) step: exact error term of s + v goes to c.
    const float t = s + v;
    const float e = (fabsf(s) >= fabsf(v)) ? ((s - t) + v) : ((v - t) + s);
    s = t;
    c += e;
}

// One target per lane. A block's four waves cover 256 / S targets: each
// target's sources are split into S ranges of whole tiles, one per wave, and
// the partial sums meet in LDS in range order (deterministic).
template <int S, bool ZERO_EPS>
__global__ __launch_bounds__(256) void direct_fp32_kernel(const float4* __restrict__ pos,
                                                          double* __restrict__ ax,
                                                          double* __restrict__ ay,
                                                          double* __restrict__ az, int n,
                                                          double G, float eps2, int tile) {
    static_assert(S == 1 || S == 2 || S == 4, "source ranges per target");
    constexpr int kTargetWaves = 4 / S;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int part = wave / kTargetWaves;                  // source range of this wave
    const int t = (wave % kTargetWaves) * kWave + lane;    // target within the block
    const int i = blockIdx.x * (kWave * kTargetWaves) + t;
    const float4 p = pos[i < n ? i : n - 1];
    const float xi = p.x, yi = p.y, zi = p.z;
    const int tiles = (n + tile - 1) / tile;
    const int per = (tiles + S - 1) / S;
    const int jb = min(n, part * per * tile);
    const int je = min(n, (part + 1) * per * tile);
    float sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    for (int j0 = jb; j0 < je; j0 += tile) {
        const int j1 = (j0 + tile) < je ? (j0 + tile) : je;
        float px = 0.f, py = 0.f, pz = 0.f;
#pragma unroll 4
        for (int j = j0; j < j1; ++j) {
            const float4 q = pos[j];
            const float dx = q.x - xi;
            const float dy = q.y - yi;
            const float dz = q.z - zi;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
            float inv = __builtin_amdgcn_rsqf(r2);
            if (ZERO_EPS) {
                inv = r2 > 0.f ? inv : 0.f;
            }
            const float inv2 = inv * inv;
            const float s = (q.w * inv) * inv2;
            px = fmaf(s, dx, px);
            py = fmaf(s, dy, py);
            pz = fmaf(s, dz, pz);
        }
        two_sum_add(sx, cx, px);
        two_sum_add(sy, cy, py);
        two_sum_add(sz, cz, pz);
    }
    if (S > 1) {
        __shared__ float part_sums[S > 1 ? S - 1 : 1][6][kWave * kTargetWaves];
        if (part > 0) {
            float* r = &part_sums[part - 1][0][0];
            constexpr int kStride = kWave * kTargetWaves;
            r[0 * kStride + t] = sx;
            r[1 * kStride + t] = sy;
            r[2 * kStride + t] = sz;
            r[3 * kStride + t] = cx;
            r[4 * kStride + t] = cy;
            r[5 * kStride + t] = cz;
        }
        __syncthreads();
        if (part == 0) {
#pragma unroll
            for (int k = 0; k < S - 1; ++k) {
                two_sum_add(sx, cx, part_sums[k][0][t]);
                two_sum_add(sy, cy, part_sums[k][1][t]);
                two_sum_add(sz, cz, part_sums[k][2][t]);
                cx += part_sums[k][3][t];
                cy += part_sums[k][4][t];
                cz += part_sums[k][5][t];
            }
        }
    }
    if (part == 0 && i < n) {
        // G applied once per target, in FP64 (exact for G = 1)
        ax[i] = G * (static_cast<double>(sx) + static_cast<double>(cx));
        ay[i] = G * (static_cast<double>(sy) + static_cast<double>(cy));
        az[i] = G * (static_cast<double>(sz) + static_cast<double>(cz));
    }
}

__global__ void refresh_pos4_kernel(const double* __restrict__ px, const double* __restrict__ py,
                                    const double* __restrict__ pz, const double* __restrict__ m,
                                    float4* __restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out[i] = make_float4(static_cast<float>(px[i]), static_cast<float>(py[i]),
                             static_cast<float>(pz[i]), static_cast<float>(m[i]));
    }
}

}  // namespace

NBodySimProStatus refresh_pos4(Engine* e, NBodySimProError* error) {
    if (e->pos4_valid || e->n == 0) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    const int n = static_cast<int>(e->n);
    refresh_pos4_kernel<<<grid_for(e->n, 256), 256, 0, e->stream>>>(e->px, e->py, e->pz, e->m,
                                                                     e->pos4, n);
    NBX_HIP(error, hipGetLastError());
    e->pos4_valid = true;
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus direct_fp64(Engine* e, const NBodySimProGravity* g, NBodySimProError* error) {
    const size_t n = e->n;
    if (n == 0) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    NBX_HIP(error, hipEventRecord(e->ev[0], e->stream));
    launch_direct_fp64<false>(e->stream, e->cu_count, e->px, e->py, e->pz, e->m, e->ax, e->ay,
                              e->az, static_cast<int>(n), g->gravitational_constant,
                              g->softening_squared, nullptr, static_cast<int>(n), nullptr);
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[1], e->stream));
    NBX_HIP(error, hipEventSynchronize(e->ev[1]));
    float ms = 0.f;
    NBX_HIP(error, hipEventElapsedTime(&ms, e->ev[0], e->ev[1]));
    e->timings = NBXTimings{};
    e->timings.force_ms = ms;
    e->timings.kernel_ms = ms;
    const double pairs = static_cast<double>(n) * static_cast<double>(n - 1);
    e->timings.kernel_flops = 20.0 * pairs;
    e->timings.kernel_bytes = 4.0 * 8.0 * static_cast<double>(n) + 3.0 * 8.0 * static_cast<double>(n);
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus direct_fp32(Engine* e, const NBodySimProGravity* g, NBodySimProError* error) {
    const size_t n = e->n;
    if (n == 0) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    NBX_HIP(error, hipEventRecord(e->ev[0], e->stream));
    NBX_TRY(refresh_pos4(e, error));
    NBX_HIP(error, hipEventRecord(e->ev[1], e->stream));
    // Four source ranges per target: 16 waves per SIMD at N = 262144.
    // Measured on MI355X at that size: 3,311 Ginteractions/s, against 3,206
    // with two ranges and 2,935 with one (one target per lane throughout;
    // two targets per lane and one range: 2,640).
    constexpr int S = 4;
    const unsigned grid = grid_for(n, 256 / S);
    const double G = g->gravitational_constant;
    const float eps2 = static_cast<float>(g->softening_squared);
    if (eps2 > 0.f) {
        direct_fp32_kernel<S, false><<<grid, 256, 0, e->stream>>>(e->pos4, e->ax, e->ay, e->az,
                                                                  static_cast<int>(n), G, eps2, 256);
    } else {
        direct_fp32_kernel<S, true><<<grid, 256, 0, e->stream>>>(e->pos4, e->ax, e->ay, e->az,
                                                                 static_cast<int>(n), G, eps2, 256);
    }
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[2], e->stream));
    NBX_HIP(error, hipEventSynchronize(e->ev[2]));
    float total = 0.f, kern = 0.f;
    NBX_HIP(error, hipEventElapsedTime(&total, e->ev[0], e->ev[2]));
    NBX_HIP(error, hipEventElapsedTime(&kern, e->ev[1], e->ev[2]));
    e->timings = NBXTimings{};
    e->timings.force_ms = total;
    e->timings.kernel_ms = kern;
    const double pairs = static_cast<double>(n) * static_cast<double>(n - 1);
    e->timings.kernel_flops = 20.0 * pairs;
    e->timings.kernel_bytes = 16.0 * static_cast<double>(n) + 24.0 * static_cast<double>(n);
    return N_BODY_SIM_PRO_STATUS_OK;
}

}  // namespace nbx

extern "C" NBodySimProStatus nbx_engine_direct_fp64_targets(NBXEngine* engine,
                                                            const NBodySimProGravity* gravity,
                                                            const int64_t* targets, size_t count,
                                                            double* out, NBodySimProError* error) {
    if (engine == nullptr || gravity == nullptr || (count > 0 && (targets == nullptr || out == nullptr))) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                 "engine, gravity, targets and out are required");
    }
    nbx::Engine* e = &engine->e;
    if (count == 0) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    if (count > 0x7FFFFFFFull) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "too many targets");
    }
    for (size_t k = 0; k < count; ++k) {
        if (targets[k] < 0 || static_cast<size_t>(targets[k]) >= e->n) {
            NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "target index out of range");
        }
    }
    NBX_HIP(error, hipSetDevice(e->device));
    const size_t tbytes = (count * sizeof(int64_t) + 255) / 256 * 256;
    NBX_TRY(e->scratch.need<unsigned char>(tbytes + 3 * count * sizeof(double), error));
    int64_t* dt = e->scratch.as<int64_t>();
    double* dout = reinterpret_cast<double*>(e->scratch.as<unsigned char>() + tbytes);
    NBX_HIP(error, hipMemcpyAsync(dt, targets, count * sizeof(int64_t), hipMemcpyHostToDevice,
                                  e->stream));
    nbx::launch_direct_fp64<true>(e->stream, e->cu_count, e->px, e->py, e->pz, e->m, nullptr,
                                  nullptr, nullptr, static_cast<int>(e->n),
                                  gravity->gravitational_constant, gravity->softening_squared, dt,
                                  static_cast<int>(count), dout);
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipMemcpyAsync(out, dout, 3 * count * sizeof(double), hipMemcpyDeviceToHost,
                                  e->stream));
    NBX_HIP(error, hipStreamSynchronize(e->stream));
    return N_BODY_SIM_PRO_STATUS_OK;
}

// ---- FP32 all-pairs without allocation (nbx_direct_fp32_noalloc) ----
//
// The FP32 (x, y, z, m) staging that nbx_hip_direct_fp32 keeps in its engine
// lives here in a fixed device array instead: each GPU the library runs on
// holds its own copy (16 MiB) until the process ends. Every call on a GPU
// shares that GPU's copy, so calls that overlap on one GPU (on different
// streams, or from different host threads) race. The staging is fully
// rewritten for [0, n) by each call, so nothing needs clearing.
namespace nbx {
namespace {
constexpr size_t kNoAllocMaxParticles = size_t(1) << 20;
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mutex;
void* g_noalloc_staging[kNoAllocMaxDevices] = {};
}  // namespace
}  // namespace nbx

static __device__ __align__(256) float4 scratch[nbx::kNoAllocMaxParticles];

extern "C" NBodySimProStatus nbx_direct_fp32_noalloc(
    const double* positions_x, const double* positions_y, const double* positions_z,
    const double* masses, double* accelerations_x, double* accelerations_y,
    double* accelerations_z, size_t count, const NBodySimProGravity* gravity, void* stream,
    NBodySimProError* error) {
    if (gravity == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "gravity is required");
    }
    if (count == 0) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    if (positions_x == nullptr || positions_y == nullptr || positions_z == nullptr ||
        masses == nullptr || accelerations_x == nullptr || accelerations_y == nullptr ||
        accelerations_z == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                 "positions, masses and accelerations must be device pointers, not null");
    }
    if (count > nbx::kNoAllocMaxParticles) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                 "nbx_direct_fp32_noalloc takes at most 1048576 particles");
    }
    int device = -1;
    NBX_HIP(error, hipGetDevice(&device));
    if (device < 0 || device >= nbx::kNoAllocMaxDevices) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_UNSUPPORTED_PLATFORM, "device ordinal above 63");
    }
    float4* staging = nullptr;
    {
        // first call on this device (current): the device address of its copy
        std::lock_guard<std::mutex> lock(nbx::g_noalloc_mutex);
        if (nbx::g_noalloc_staging[device] == nullptr) {
            void* p = nullptr;
            NBX_HIP(error, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            nbx::g_noalloc_staging[device] = p;
        }
        staging = static_cast<float4*>(nbx::g_noalloc_staging[device]);
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int n = static_cast<int>(count);
    nbx::refresh_pos4_kernel<<<nbx::grid_for(count, 256), 256, 0, st>>>(
        positions_x, positions_y, positions_z, masses, staging, n);
    // same kernel, source split and tile as nbx_hip_direct_fp32 (identical results)
    constexpr int S = 4;
    const unsigned grid = nbx::grid_for(count, 256 / S);
    const double G = gravity->gravitational_constant;
    const float eps2 = static_cast<float>(gravity->softening_squared);
    if (eps2 > 0.f) {
        nbx::direct_fp32_kernel<S, false><<<grid, 256, 0, st>>>(
            staging, accelerations_x, accelerations_y, accelerations_z, n, G, eps2, 256);
    } else {
        nbx::direct_fp32_kernel<S, true><<<grid, 256, 0, st>>>(
            staging, accelerations_x, accelerations_y, accelerations_z, n, G, eps2, 256);
    }
    NBX_HIP(error, hipGetLastError());
    return N_BODY_SIM_PRO_STATUS_OK;
}
