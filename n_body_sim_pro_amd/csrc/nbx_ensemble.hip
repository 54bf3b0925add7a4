// nbx_ensemble.hip - batched ensemble of independent clusters (include/nbx_ensemble.h).
//
// Work decomposition. A work item is (cluster, 64*R consecutive targets of it)
// with R = targets per lane chosen per cluster-size bucket. Items are sorted
// by cost (targets x cluster size, longest first) and dealt round-robin to 8
// per-XCD queues; a persistent grid (blocks_per_cu x CUs) pulls items wave by
// wave from its XCD's queue, claiming the next item before computing the
// current one so the atomic's latency is hidden. The source loop of an item
// is wave-uniform (scalar loads, SGPR operands) exactly as in hip_direct_fp32,
// with the same tile-compensated FP32 accumulation; results are identical
// for any R or blocks-per-CU choice (only source_tile moves the fold points).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "nbx_ensemble.h"
#include "nbx_internal.h"

namespace nbx {
namespace {

constexpr int kThreads = 256;
constexpr int kQueues = 8;      // one work queue per XCD
constexpr int kBuckets = 8;

struct WorkItem {
    int off;     // first particle of the cluster
    int n;       // cluster size
    int start;   // first target, relative to the cluster
    int r;       // targets per lane (1, 2 or 4)
};

struct Ensemble {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t total = 0;
    size_t clusters = 0;
    std::vector<uint32_t> sizes;
    std::vector<size_t> offsets;
    double *px = nullptr, *py = nullptr, *pz = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *ax = nullptr, *ay = nullptr, *az = nullptr;
    double* m = nullptr;
    float4* pos4 = nullptr;
    Buf items;       // LPT-ordered, dealt to queues
    int nitems = 0;
    Buf eitems;      // cluster-ordered items of 64 targets (energies)
    std::vector<int> eitem_first;  // first energy item of each cluster (+ sentinel)
    Buf epart;       // energy partials per energy item
    Buf eout;
    Buf eidx;        // first energy item per cluster (device)
    unsigned* queue = nullptr;
    NBXTileConfig tile{};
    int cu_count = 256;
    std::string arch;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;
    double graph_dt = 0.0;
    double graph_g = 0.0, graph_eps2 = -1.0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

__device__ inline void kb_add(float& s, float& c, float v) {
    const float t = s + v;
    const float e = (fabsf(s) >= fabsf(v)) ? ((s - t) + v) : ((v - t) + s);
    s = t;
    c += e;
}

// ZERO_EPS: unsoftened sums; the self pair (r2 == 0) and coincident pairs
// contribute nothing instead of rsq(0) = inf times 0 = NaN.
template <int R, bool ZERO_EPS>
__device__ inline void tile_force(const float4* __restrict__ pos, double* __restrict__ ax,
                                  double* __restrict__ ay, double* __restrict__ az,
                                  const WorkItem& w, int lane, float G, float eps2, int tile) {
    const int end = w.off + w.n;
    float xi[R], yi[R], zi[R], sx[R], sy[R], sz[R], cx[R], cy[R], cz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = w.off + w.start + r * kWave + lane;
        const float4 p = pos[i < end ? i : end - 1];
        xi[r] = p.x;
        yi[r] = p.y;
        zi[r] = p.z;
        sx[r] = sy[r] = sz[r] = cx[r] = cy[r] = cz[r] = 0.f;
    }
    for (int j0 = w.off; j0 < end; j0 += tile) {
        const int j1 = (j0 + tile) < end ? (j0 + tile) : end;
        float px[R], py[R], pz[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            px[r] = py[r] = pz[r] = 0.f;
        }
#pragma unroll 4
        for (int j = j0; j < j1; ++j) {
            const float4 q = pos[j];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float dx = q.x - xi[r];
                const float dy = q.y - yi[r];
                const float dz = q.z - zi[r];
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
                float inv = __builtin_amdgcn_rsqf(r2);
                if (ZERO_EPS) {
                    inv = r2 > 0.f ? inv : 0.f;
                }
                const float s = (q.w * inv) * (inv * inv);
                px[r] = fmaf(s, dx, px[r]);
                py[r] = fmaf(s, dy, py[r]);
                pz[r] = fmaf(s, dz, pz[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            kb_add(sx[r], cx[r], px[r]);
            kb_add(sy[r], cy[r], py[r]);
            kb_add(sz[r], cz[r], pz[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = w.off + w.start + r * kWave + lane;
        if (i < end) {
            ax[i] = static_cast<double>(G) * (static_cast<double>(sx[r]) + static_cast<double>(cx[r]));
            ay[i] = static_cast<double>(G) * (static_cast<double>(sy[r]) + static_cast<double>(cy[r]));
            az[i] = static_cast<double>(G) * (static_cast<double>(sz[r]) + static_cast<double>(cz[r]));
        }
    }
}

// Persistent launch: waves pull items from their XCD's queue until it is empty.
template <bool ZERO_EPS>
__global__ __launch_bounds__(kThreads) void ens_force_kernel(
    const float4* __restrict__ pos, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, const WorkItem* __restrict__ items, int nitems,
    unsigned* __restrict__ queues, float G, float eps2, int tile) {
    const int lane = threadIdx.x & (kWave - 1);
    const int qid = blockIdx.x % kQueues;
    unsigned* q = queues + qid;
    unsigned k = 0;
    if (lane == 0) {
        k = atomicAdd(q, 1u);
    }
    int item = __builtin_amdgcn_readfirstlane(static_cast<int>(k)) * kQueues + qid;
    while (item < nitems) {
        unsigned kn = 0;
        if (lane == 0) {
            kn = atomicAdd(q, 1u);  // claim the next item before computing this one
        }
        const WorkItem w = items[item];
        if (w.r == 4) {
            tile_force<4, ZERO_EPS>(pos, ax, ay, az, w, lane, G, eps2, tile);
        } else if (w.r == 2) {
            tile_force<2, ZERO_EPS>(pos, ax, ay, az, w, lane, G, eps2, tile);
        } else {
            tile_force<1, ZERO_EPS>(pos, ax, ay, az, w, lane, G, eps2, tile);
        }
        item = __builtin_amdgcn_readfirstlane(static_cast<int>(kn)) * kQueues + qid;
    }
}

__global__ void ens_reset_kernel(unsigned* __restrict__ queues) {
    if (threadIdx.x < kQueues) {
        queues[threadIdx.x] = 0u;
    }
}

// Leapfrog pieces (src/c/physics/integrator.c:64-83, contraction off) with the
// FP32 mirror refreshed right after the drift.
__global__ void ens_kick_drift_kernel(double* __restrict__ px, double* __restrict__ py,
                                      double* __restrict__ pz, double* __restrict__ vx,
                                      double* __restrict__ vy, double* __restrict__ vz,
                                      const double* __restrict__ ax, const double* __restrict__ ay,
                                      const double* __restrict__ az, const double* __restrict__ m,
                                      float4* __restrict__ pos4, int n, double half, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double nvx = vx[i] + ax[i] * half;
        const double nvy = vy[i] + ay[i] * half;
        const double nvz = vz[i] + az[i] * half;
        vx[i] = nvx;
        vy[i] = nvy;
        vz[i] = nvz;
        const double x = px[i] + nvx * dt;
        const double y = py[i] + nvy * dt;
        const double z = pz[i] + nvz * dt;
        px[i] = x;
        py[i] = y;
        pz[i] = z;
        pos4[i] = make_float4(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z),
                              static_cast<float>(m[i]));
    }
}

__global__ void ens_kick_kernel(double* __restrict__ vx, double* __restrict__ vy,
                                double* __restrict__ vz, const double* __restrict__ ax,
                                const double* __restrict__ ay, const double* __restrict__ az,
                                int n, double half) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        vx[i] = vx[i] + ax[i] * half;
        vy[i] = vy[i] + ay[i] * half;
        vz[i] = vz[i] + az[i] * half;
    }
}

__global__ void ens_pos4_kernel(const double* __restrict__ px, const double* __restrict__ py,
                                const double* __restrict__ pz, const double* __restrict__ m,
                                float4* __restrict__ pos4, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        pos4[i] = make_float4(static_cast<float>(px[i]), static_cast<float>(py[i]),
                              static_cast<float>(pz[i]), static_cast<float>(m[i]));
    }
}

// FP64 kinetic energy and softened pair potential of one 64-target slice of a
// cluster (numerics.c:5-81 and :83-116 restricted to the cluster): one wave
// per energy item, fixed-order wave reduction.
__global__ __launch_bounds__(kWave) void ens_energy_kernel(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
    const double* __restrict__ m, const WorkItem* __restrict__ items, int nitems, double eps2,
    double* __restrict__ partial) {
    const int item = blockIdx.x;
    if (item >= nitems) {
        return;
    }
    const WorkItem w = items[item];
    const int lane = threadIdx.x;
    const int end = w.off + w.n;
    const int i = w.off + w.start + lane;
    double ke = 0.0, pot = 0.0;
    if (i < end) {
        const double mi = m[i];
        ke = 0.5 * mi * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
        const double x = px[i], y = py[i], z = pz[i];
        for (int j = i + 1; j < end; ++j) {
            const double dx = px[j] - x;
            const double dy = py[j] - y;
            const double dz = pz[j] - z;
            const double d = sqrt(dx * dx + dy * dy + dz * dz + eps2);
            pot += mi * m[j] / d;
        }
    }
    __shared__ double s[2][kWave];
    s[0][lane] = ke;
    s[1][lane] = pot;
    __syncthreads();
    for (int off = kWave / 2; off > 0; off >>= 1) {
        if (lane < off) {
            s[0][lane] += s[0][lane + off];
            s[1][lane] += s[1][lane + off];
        }
        __syncthreads();
    }
    if (lane == 0) {
        partial[2 * item] = s[0][0];
        partial[2 * item + 1] = s[1][0];
    }
}

__global__ void ens_energy_reduce_kernel(const double* __restrict__ partial,
                                         const int* __restrict__ first, int clusters, double G,
                                         double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= clusters) {
        return;
    }
    double ke = 0.0, pot = 0.0;
    for (int k = first[c]; k < first[c + 1]; ++k) {
        ke += partial[2 * k];
        pot += partial[2 * k + 1];
    }
    out[2 * c] = ke;
    out[2 * c + 1] = -G * pot;
}

int bucket_of(uint32_t n) {
    int b = 0;
    while (b < kBuckets - 1 && n >= (128u << b)) {
        ++b;
    }
    return b;
}

void default_tile(NBXTileConfig* t) {
    for (int b = 0; b < kBuckets; ++b) {
        t->targets_per_lane[b] = b < 2 ? 1 : (b < 5 ? 2 : 4);
    }
    t->blocks_per_cu = 4;
    t->source_tile = 256;
}

bool tile_ok(const NBXTileConfig* t) {
    for (int b = 0; b < kBuckets; ++b) {
        const int r = t->targets_per_lane[b];
        if (r != 1 && r != 2 && r != 4) return false;
    }
    return t->blocks_per_cu >= 1 && t->blocks_per_cu <= 8 && t->source_tile >= 4 &&
           t->source_tile % 4 == 0;
}

// LPT-ordered items for the clusters selected by `use` (all when empty),
// dealt round-robin to the per-XCD queues (item k goes to queue k % 8).
std::vector<WorkItem> build_items(const Ensemble& s, const NBXTileConfig& t,
                                  const std::vector<char>& use) {
    std::vector<WorkItem> v;
    for (size_t c = 0; c < s.clusters; ++c) {
        if (!use.empty() && !use[c]) continue;
        const int n = static_cast<int>(s.sizes[c]);
        const int r = t.targets_per_lane[bucket_of(s.sizes[c])];
        for (int st = 0; st < n; st += kWave * r) {
            v.push_back(WorkItem{static_cast<int>(s.offsets[c]), n, st, r});
        }
    }
    std::stable_sort(v.begin(), v.end(), [](const WorkItem& a, const WorkItem& b) {
        const long long ca = static_cast<long long>(std::min(a.n - a.start, 64 * a.r)) * a.n;
        const long long cb = static_cast<long long>(std::min(b.n - b.start, 64 * b.r)) * b.n;
        return ca > cb;
    });
    return v;
}

NBodySimProStatus upload_items(Ensemble& s, const std::vector<WorkItem>& v,
                               NBodySimProError* error) {
    NBX_TRY(s.items.need<WorkItem>(v.size() + 1, error));
    if (!v.empty()) {
        NBX_HIP(error, hipMemcpy(s.items.p, v.data(), v.size() * sizeof(WorkItem),
                                 hipMemcpyHostToDevice));
    }
    s.nitems = static_cast<int>(v.size());
    return N_BODY_SIM_PRO_STATUS_OK;
}

void launch_force(Ensemble& s, float G, float eps2) {
    ens_reset_kernel<<<1, kQueues, 0, s.stream>>>(s.queue);
    const unsigned grid = static_cast<unsigned>(s.cu_count * s.tile.blocks_per_cu);
    if (eps2 > 0.f) {
        ens_force_kernel<false><<<grid, kThreads, 0, s.stream>>>(
            s.pos4, s.ax, s.ay, s.az, s.items.as<WorkItem>(), s.nitems, s.queue, G, eps2,
            s.tile.source_tile);
    } else {
        ens_force_kernel<true><<<grid, kThreads, 0, s.stream>>>(
            s.pos4, s.ax, s.ay, s.az, s.items.as<WorkItem>(), s.nitems, s.queue, G, eps2,
            s.tile.source_tile);
    }
}

void drop_graph(Ensemble& s) {
    if (s.graph_exec != nullptr) (void)hipGraphExecDestroy(s.graph_exec);
    if (s.graph != nullptr) (void)hipGraphDestroy(s.graph);
    s.graph_exec = nullptr;
    s.graph = nullptr;
    s.graph_steps = 0;
}

void enqueue_steps(Ensemble& s, float G, float eps2, double dt, int steps) {
    const int n = static_cast<int>(s.total);
    const unsigned g = grid_for(s.total, kThreads);
    const double half = 0.5 * dt;
    for (int k = 0; k < steps; ++k) {
        ens_kick_drift_kernel<<<g, kThreads, 0, s.stream>>>(s.px, s.py, s.pz, s.vx, s.vy, s.vz,
                                                            s.ax, s.ay, s.az, s.m, s.pos4, n, half,
                                                            dt);
        launch_force(s, G, eps2);
        ens_kick_kernel<<<g, kThreads, 0, s.stream>>>(s.vx, s.vy, s.vz, s.ax, s.ay, s.az, n, half);
    }
}

float time_force(Ensemble& s, float G, float eps2, int reps) {
    std::vector<float> t;
    for (int k = 0; k < reps; ++k) {
        (void)hipEventRecord(s.ev0, s.stream);
        launch_force(s, G, eps2);
        (void)hipEventRecord(s.ev1, s.stream);
        (void)hipEventSynchronize(s.ev1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, s.ev0, s.ev1);
        t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

}  // namespace
}  // namespace nbx

struct NBXEnsemble {
    nbx::Ensemble s;
};

using nbx::Ensemble;

extern "C" {

void nbx_ensemble_destroy(NBXEnsemble* e) {
    if (e == nullptr) {
        return;
    }
    Ensemble& s = e->s;
    (void)hipSetDevice(s.device);
    if (s.stream != nullptr) (void)hipStreamSynchronize(s.stream);
    nbx::drop_graph(s);
    double* arrays[] = {s.px, s.py, s.pz, s.vx, s.vy, s.vz, s.ax, s.ay, s.az, s.m};
    for (double* a : arrays) {
        if (a != nullptr) (void)hipFree(a);
    }
    if (s.pos4 != nullptr) (void)hipFree(s.pos4);
    if (s.queue != nullptr) (void)hipFree(s.queue);
    nbx::Buf* bufs[] = {&s.items, &s.eitems, &s.epart, &s.eout, &s.eidx};
    for (nbx::Buf* b : bufs) {
        if (b->p != nullptr) (void)hipFree(b->p);
    }
    if (s.ev0 != nullptr) (void)hipEventDestroy(s.ev0);
    if (s.ev1 != nullptr) (void)hipEventDestroy(s.ev1);
    if (s.stream != nullptr) (void)hipStreamDestroy(s.stream);
    delete e;
}

NBXEnsemble* nbx_ensemble_create(int device, const uint32_t* sizes, size_t clusters,
                                 NBodySimProError* error) {
    if (sizes == nullptr || clusters == 0) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, __FILE__, __LINE__,
                       "cluster sizes must be given");
        return nullptr;
    }
    size_t total = 0;
    for (size_t c = 0; c < clusters; ++c) {
        if (sizes[c] == 0) {
            nbx::set_error(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, __FILE__, __LINE__,
                           "every cluster needs at least one particle");
            return nullptr;
        }
        total += sizes[c];
    }
    if (total > 0x7FFFFFFFull) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, __FILE__, __LINE__,
                       "ensemble exceeds 2^31 particles");
        return nullptr;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0 || device < 0 || device >= count) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_UNSUPPORTED_PLATFORM, __FILE__, __LINE__,
                       "no such HIP device");
        return nullptr;
    }
    NBXEnsemble* e = new (std::nothrow) NBXEnsemble();
    if (e == nullptr) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY, __FILE__, __LINE__,
                       "ensemble allocation failed");
        return nullptr;
    }
    Ensemble& s = e->s;
    s.device = device;
    s.clusters = clusters;
    s.total = total;
    s.sizes.assign(sizes, sizes + clusters);
    s.offsets.resize(clusters);
    for (size_t c = 0, o = 0; c < clusters; ++c) {
        s.offsets[c] = o;
        o += sizes[c];
    }
    hipDeviceProp_t prop;
    bool ok = hipSetDevice(device) == hipSuccess &&
              hipGetDeviceProperties(&prop, device) == hipSuccess &&
              hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&s.ev0) == hipSuccess && hipEventCreate(&s.ev1) == hipSuccess;
    if (ok) {
        s.cu_count = prop.multiProcessorCount;
        s.arch = prop.gcnArchName;
    }
    double** arrays[] = {&s.px, &s.py, &s.pz, &s.vx, &s.vy, &s.vz, &s.ax, &s.ay, &s.az, &s.m};
    for (double** a : arrays) {
        ok = ok && hipMalloc(reinterpret_cast<void**>(a), total * sizeof(double)) == hipSuccess;
    }
    ok = ok && hipMalloc(reinterpret_cast<void**>(&s.pos4), total * sizeof(float4)) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&s.queue), 64 * sizeof(unsigned)) == hipSuccess;
    if (!ok) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY, __FILE__, __LINE__,
                       "device allocation failed while creating the ensemble");
        nbx_ensemble_destroy(e);
        return nullptr;
    }
    nbx::default_tile(&s.tile);
    // energy items: cluster order, 64 targets each
    std::vector<nbx::WorkItem> ev;
    s.eitem_first.assign(clusters + 1, 0);
    for (size_t c = 0; c < clusters; ++c) {
        s.eitem_first[c] = static_cast<int>(ev.size());
        for (uint32_t st = 0; st < sizes[c]; st += nbx::kWave) {
            ev.push_back(nbx::WorkItem{static_cast<int>(s.offsets[c]), static_cast<int>(sizes[c]),
                                       static_cast<int>(st), 1});
        }
    }
    s.eitem_first[clusters] = static_cast<int>(ev.size());
    NBodySimProError local{};
    NBodySimProError* err = error != nullptr ? error : &local;
    if (s.eitems.need<nbx::WorkItem>(ev.size(), err) != N_BODY_SIM_PRO_STATUS_OK ||
        s.epart.need<double>(2 * ev.size(), err) != N_BODY_SIM_PRO_STATUS_OK ||
        s.eout.need<double>(2 * clusters, err) != N_BODY_SIM_PRO_STATUS_OK ||
        s.eidx.need<int>(clusters + 1, err) != N_BODY_SIM_PRO_STATUS_OK ||
        hipMemcpy(s.eitems.p, ev.data(), ev.size() * sizeof(nbx::WorkItem),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s.eidx.p, s.eitem_first.data(), (clusters + 1) * sizeof(int),
                  hipMemcpyHostToDevice) != hipSuccess ||
        nbx::upload_items(s, nbx::build_items(s, s.tile, {}), err) != N_BODY_SIM_PRO_STATUS_OK) {
        nbx_ensemble_destroy(e);
        return nullptr;
    }
    return e;
}

size_t nbx_ensemble_particle_count(const NBXEnsemble* e) {
    return e == nullptr ? 0 : e->s.total;
}

NBodySimProStatus nbx_ensemble_upload(NBXEnsemble* e, const NBodySimProParticleSystemView* v,
                                      NBodySimProError* error) {
    if (e == nullptr || v == nullptr || v->particle_count != e->s.total) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                 "view must cover exactly the ensemble's particles");
    }
    Ensemble& s = e->s;
    NBX_HIP(error, hipSetDevice(s.device));
    const size_t bytes = s.total * sizeof(double);
    const double* src[] = {v->positions_x, v->positions_y, v->positions_z, v->velocities_x,
                           v->velocities_y, v->velocities_z, v->accelerations_x,
                           v->accelerations_y, v->accelerations_z, v->masses};
    double* dst[] = {s.px, s.py, s.pz, s.vx, s.vy, s.vz, s.ax, s.ay, s.az, s.m};
    for (int k = 0; k < 10; ++k) {
        if (src[k] != nullptr) {
            NBX_HIP(error, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyHostToDevice, s.stream));
        } else if (k < 6 || k == 9) {
            NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                     "view positions, velocities and masses must not be null");
        }
    }
    nbx::ens_pos4_kernel<<<nbx::grid_for(s.total, nbx::kThreads), nbx::kThreads, 0, s.stream>>>(
        s.px, s.py, s.pz, s.m, s.pos4, static_cast<int>(s.total));
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipStreamSynchronize(s.stream));
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_ensemble_download(NBXEnsemble* e, const NBodySimProParticleSystemView* v,
                                        NBodySimProError* error) {
    if (e == nullptr || v == nullptr || v->particle_count != e->s.total) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                 "view must cover exactly the ensemble's particles");
    }
    Ensemble& s = e->s;
    NBX_HIP(error, hipSetDevice(s.device));
    const size_t bytes = s.total * sizeof(double);
    double* dst[] = {v->positions_x, v->positions_y, v->positions_z, v->velocities_x,
                     v->velocities_y, v->velocities_z, v->accelerations_x, v->accelerations_y,
                     v->accelerations_z, v->masses};
    const double* src[] = {s.px, s.py, s.pz, s.vx, s.vy, s.vz, s.ax, s.ay, s.az, s.m};
    for (int k = 0; k < 10; ++k) {
        if (dst[k] != nullptr) {
            NBX_HIP(error, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, s.stream));
        }
    }
    NBX_HIP(error, hipStreamSynchronize(s.stream));
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_ensemble_force(NBXEnsemble* e, const NBodySimProGravity* g,
                                     NBodySimProError* error) {
    if (e == nullptr || g == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "ensemble and gravity required");
    }
    Ensemble& s = e->s;
    NBX_HIP(error, hipSetDevice(s.device));
    nbx::launch_force(s, static_cast<float>(g->gravitational_constant),
                      static_cast<float>(g->softening_squared));
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipStreamSynchronize(s.stream));
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_ensemble_advance(NBXEnsemble* e, const NBodySimProGravity* g,
                                       double dt, int steps, int graph_steps, double* elapsed_ms,
                                       NBodySimProError* error) {
    if (e == nullptr || g == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "ensemble and gravity required");
    }
    if (!(dt > 0.0)) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "timestep must be positive");
    }
    if (steps < 0 || graph_steps < 0) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "steps must be non-negative");
    }
    Ensemble& s = e->s;
    NBX_HIP(error, hipSetDevice(s.device));
    const float G = static_cast<float>(g->gravitational_constant);
    const float eps2 = static_cast<float>(g->softening_squared);
    if (graph_steps > 0 && (s.graph_exec == nullptr || s.graph_steps != graph_steps ||
                            s.graph_dt != dt || s.graph_g != g->gravitational_constant ||
                            s.graph_eps2 != g->softening_squared)) {
        nbx::drop_graph(s);
        NBX_HIP(error, hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
        nbx::enqueue_steps(s, G, eps2, dt, graph_steps);
        // End the capture whatever the launches returned, so the stream never
        // stays in capture mode; a failed capture leaves no graph behind.
        const hipError_t launched = hipGetLastError();
        hipGraph_t graph = nullptr;
        const hipError_t ended = hipStreamEndCapture(s.stream, &graph);
        s.graph = graph;
        if (launched != hipSuccess || ended != hipSuccess) {
            nbx::drop_graph(s);
            NBX_HIP(error, launched != hipSuccess ? launched : ended);
        }
        const hipError_t inst = hipGraphInstantiate(&s.graph_exec, s.graph, nullptr, nullptr, 0);
        if (inst != hipSuccess) {
            s.graph_exec = nullptr;
            nbx::drop_graph(s);
            NBX_HIP(error, inst);
        }
        s.graph_steps = graph_steps;
        s.graph_dt = dt;
        s.graph_g = g->gravitational_constant;
        s.graph_eps2 = g->softening_squared;
    }
    NBX_HIP(error, hipEventRecord(s.ev0, s.stream));
    int left = steps;
    if (graph_steps > 0) {
        while (left >= graph_steps) {
            NBX_HIP(error, hipGraphLaunch(s.graph_exec, s.stream));
            left -= graph_steps;
        }
    }
    nbx::enqueue_steps(s, G, eps2, dt, left);
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(s.ev1, s.stream));
    NBX_HIP(error, hipEventSynchronize(s.ev1));
    if (elapsed_ms != nullptr) {
        float ms = 0.f;
        NBX_HIP(error, hipEventElapsedTime(&ms, s.ev0, s.ev1));
        *elapsed_ms = ms;
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_ensemble_energies(NBXEnsemble* e, const NBodySimProGravity* g, double* out,
                                        NBodySimProError* error) {
    if (e == nullptr || g == nullptr || out == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "ensemble, gravity, out required");
    }
    Ensemble& s = e->s;
    NBX_HIP(error, hipSetDevice(s.device));
    const int nitems = s.eitem_first[s.clusters];
    nbx::ens_energy_kernel<<<nitems, nbx::kWave, 0, s.stream>>>(
        s.px, s.py, s.pz, s.vx, s.vy, s.vz, s.m, s.eitems.as<nbx::WorkItem>(), nitems,
        g->softening_squared, s.epart.as<double>());
    nbx::ens_energy_reduce_kernel<<<nbx::grid_for(s.clusters, nbx::kThreads), nbx::kThreads, 0,
                                    s.stream>>>(s.epart.as<double>(), s.eidx.as<int>(),
                                                static_cast<int>(s.clusters),
                                                g->gravitational_constant, s.eout.as<double>());
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipMemcpyAsync(out, s.eout.p, 2 * s.clusters * sizeof(double),
                                  hipMemcpyDeviceToHost, s.stream));
    NBX_HIP(error, hipStreamSynchronize(s.stream));
    return N_BODY_SIM_PRO_STATUS_OK;
}

void nbx_ensemble_get_tile(const NBXEnsemble* e, NBXTileConfig* config) {
    if (e != nullptr && config != nullptr) {
        *config = e->s.tile;
    }
}

NBodySimProStatus nbx_ensemble_set_tile(NBXEnsemble* e, const NBXTileConfig* config,
                                        NBodySimProError* error) {
    if (e == nullptr || config == nullptr || !nbx::tile_ok(config)) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "invalid tile configuration");
    }
    Ensemble& s = e->s;
    NBX_HIP(error, hipSetDevice(s.device));
    s.tile = *config;
    nbx::drop_graph(s);
    return nbx::upload_items(s, nbx::build_items(s, s.tile, {}), error);
}

NBodySimProStatus nbx_ensemble_autotune(NBXEnsemble* e, const NBodySimProGravity* g,
                                        const char* cache_path, char* report, size_t report_size,
                                        NBodySimProError* error) {
    if (e == nullptr || g == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "ensemble and gravity required");
    }
    Ensemble& s = e->s;
    NBX_HIP(error, hipSetDevice(s.device));
    const float G = static_cast<float>(g->gravitational_constant);
    const float eps2 = static_cast<float>(g->softening_squared);
    // cache key: architecture + per-bucket cluster counts
    std::vector<size_t> hist(nbx::kBuckets, 0);
    for (uint32_t n : s.sizes) hist[nbx::bucket_of(n)]++;
    std::string key = s.arch;
    for (size_t h : hist) key += ":" + std::to_string(h);
    NBXTileConfig best = s.tile;
    bool cached = false;
    if (cache_path != nullptr) {
        if (FILE* f = std::fopen(cache_path, "r")) {
            char line[512];
            while (std::fgets(line, sizeof(line), f) != nullptr) {
                char k[400];
                NBXTileConfig t{};
                if (std::sscanf(line, "%399s %d %d %d %d %d %d %d %d %d %d", k,
                                &t.targets_per_lane[0], &t.targets_per_lane[1],
                                &t.targets_per_lane[2], &t.targets_per_lane[3],
                                &t.targets_per_lane[4], &t.targets_per_lane[5],
                                &t.targets_per_lane[6], &t.targets_per_lane[7], &t.blocks_per_cu,
                                &t.source_tile) == 11 &&
                    key == k && nbx::tile_ok(&t)) {
                    best = t;
                    cached = true;
                }
            }
            std::fclose(f);
        }
    }
    double best_ms = 0.0;
    if (!cached) {
        // 1) targets per lane, bucket by bucket, on that bucket's clusters alone
        for (int b = 0; b < nbx::kBuckets; ++b) {
            if (hist[b] == 0) continue;
            std::vector<char> use(s.clusters, 0);
            for (size_t c = 0; c < s.clusters; ++c) use[c] = nbx::bucket_of(s.sizes[c]) == b;
            float bt = 1e30f;
            int br = best.targets_per_lane[b];
            for (int r : {1, 2, 4}) {
                NBXTileConfig t = best;
                t.targets_per_lane[b] = r;
                s.tile = t;
                NBX_TRY(nbx::upload_items(s, nbx::build_items(s, t, use), error));
                nbx::time_force(s, G, eps2, 1);
                const float ms = nbx::time_force(s, G, eps2, 3);
                if (ms < bt) {
                    bt = ms;
                    br = r;
                }
            }
            best.targets_per_lane[b] = br;
        }
        // 2) persistent-grid width and compensation tile on the whole ensemble
        float bt = 1e30f;
        NBXTileConfig keep = best;
        for (int bpc : {1, 2, 4, 8}) {
            for (int st : {64, 128, 256, 512}) {
                NBXTileConfig t = keep;
                t.blocks_per_cu = bpc;
                t.source_tile = st;
                s.tile = t;
                NBX_TRY(nbx::upload_items(s, nbx::build_items(s, t, {}), error));
                nbx::time_force(s, G, eps2, 1);
                const float ms = nbx::time_force(s, G, eps2, 3);
                if (ms < bt) {
                    bt = ms;
                    best = t;
                }
            }
        }
        best_ms = bt;
        if (cache_path != nullptr) {
            if (FILE* f = std::fopen(cache_path, "a")) {
                std::fprintf(f, "%s %d %d %d %d %d %d %d %d %d %d\n", key.c_str(),
                             best.targets_per_lane[0], best.targets_per_lane[1],
                             best.targets_per_lane[2], best.targets_per_lane[3],
                             best.targets_per_lane[4], best.targets_per_lane[5],
                             best.targets_per_lane[6], best.targets_per_lane[7],
                             best.blocks_per_cu, best.source_tile);
                std::fclose(f);
            }
        }
    }
    s.tile = best;
    nbx::drop_graph(s);
    NBX_TRY(nbx::upload_items(s, nbx::build_items(s, best, {}), error));
    if (report != nullptr && report_size > 0) {
        std::snprintf(report, report_size,
                      "{\"cached\": %s, \"best_force_ms\": %.4f, \"targets_per_lane\": "
                      "[%d,%d,%d,%d,%d,%d,%d,%d], \"blocks_per_cu\": %d, \"source_tile\": %d}",
                      cached ? "true" : "false", best_ms, best.targets_per_lane[0],
                      best.targets_per_lane[1], best.targets_per_lane[2],
                      best.targets_per_lane[3], best.targets_per_lane[4],
                      best.targets_per_lane[5], best.targets_per_lane[6],
                      best.targets_per_lane[7], best.blocks_per_cu, best.source_tile);
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

}  // extern "C"
