// nbx_bh.hip - Barnes-Hut on gfx950: bounding cube, keys, sorts, octree,
// centre-of-mass pass and walk, reproducing src/c/barnes_hut/barnes_hut.c.
//
// Topology. The reference builds a point-region octree by inserting particles
// one at a time in Morton order (barnes_hut.c:503-520); the octant of a
// particle in a cell is decided by strict '>' tests against the FP64 cell
// centre, and child centres come from repeated c +/- h*0.5 (:129-142,
// :202-211). The resulting tree does not depend on insertion order: two
// particles share the depth-d cell iff their first d octant digits agree.
// Here every particle's first 21 octant digits are computed with exactly those
// FP64 operations ("true-path key", same bit layout as the Morton key). When
// the true-path key equals the Morton key for every particle (the usual case)
// the Morton order is already a depth-first order of the tree; otherwise the
// particles are re-sorted by true-path key (stable, so Morton order breaks
// ties). Particles that share all 21 digits get up to six more 21-digit words
// (depth <= 147 >= the reference's 128-level limit) and a multi-word LSD sort.
//
// Nodes are laid out in depth-first pre-order with children in DESCENDING
// octant order, built from the sorted keys with no per-level launches:
// delta[t] = common levels of neighbours t, t+1 (in descending key order);
// position t starts internal nodes at depths (delta[t-1], delta[t]]; a scan of
// those counts gives every node's pre-order index, a binary search on the keys
// gives each subtree's end (the "skip" pointer). That pre-order with children
// 7..0 is exactly the order in which the reference's LIFO walk pops nodes
// (children pushed 0..7, barnes_hut.c:617-634), so a walk that follows skip
// pointers visits, for every query, the same nodes in the same order.
//
// Centre of mass. Computed in FP64 from the children in octant order 0..7
// (barnes_hut.c:230-244), one thread per node chosen by an arrival counter
// (the last child to finish computes its parent), so results are
// deterministic and bit-identical to the reference.
//
// Walk. One wave = 64 consecutive queries in tree order; the node index is
// wave-uniform (node records come in through scalar loads). Each lane keeps
// "skip_until": it is active at node i iff i >= skip_until. Active lanes apply
// the reference's own test (d^2 > 0 and 4h^2 < theta^2 d^2, unsoftened) and
// either take the monopole/leaf term (then skip the subtree) or descend. The
// wave moves to i+1 if any active lane descends, else to skip[i]. Every lane
// thus sees exactly its reference interaction set, in reference order, with
// no stack. The FP64 variant uses the reference arithmetic and is
// bit-identical; the FP32 variant uses FP32 nodes and v_rsq_f32.
#include "nbx_internal.h"

#include <vector>

namespace nbx {
namespace {

constexpr int kThreads = 256;
constexpr int kExtWords = 6;
constexpr int kMaxLevels = 129;  // a shared depth-129 cell means depth > 128 (barnes_hut.c:149-153)
enum { FLAG_MISMATCH = 0, FLAG_TIE = 1, FLAG_OVERFLOW = 2, FLAG_MAXDEPTH = 3, FLAG_COUNT = 8 };

__device__ inline uint64_t spread3(uint64_t a) {
    uint64_t x = a & 0x1FFFFFull;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// barnes_hut.c:286-294
__device__ inline uint64_t axis_code(double v, double c, double h) {
    double u = (v - c) / (2.0 * h) + 0.5;
    if (u < 0.0) {
        u = 0.0;
    } else if (u >= 1.0) {
        u = 1.0 - 1.0e-15;
    }
    return static_cast<uint64_t>(u * 2097152.0);
}

// One 21-level word of octant digits starting at the given cell (advances it).
__device__ inline uint64_t path_word(double x, double y, double z, double& cx, double& cy,
                                     double& cz, double& h) {
    uint64_t key = 0;
    for (int d = 0; d < 21; ++d) {
        const int o = (x > cx ? 1 : 0) | (y > cy ? 2 : 0) | (z > cz ? 4 : 0);
        key |= static_cast<uint64_t>(o) << (60 - 3 * d);
        const double off = h * 0.5;
        cx = cx + ((o & 1) ? off : -off);
        cy = cy + ((o & 2) ? off : -off);
        cz = cz + ((o & 4) ? off : -off);
        h = 0.5 * h;
    }
    return key;
}

__device__ inline int word_levels(uint64_t x) {
    // x != 0, key bits 0..62: first differing bit 62-k lies in level k/3.
    return (__clzll(static_cast<long long>(x)) - 1) / 3;
}

__device__ inline int common_levels(const uint64_t* __restrict__ w,
                                    const uint64_t* __restrict__ ext, int a, int b) {
    const uint64_t x = w[a] ^ w[b];
    if (x != 0) {
        return word_levels(x);
    }
    if (ext == nullptr) {
        return 21;
    }
    for (int k = 0; k < kExtWords; ++k) {
        const uint64_t y = ext[static_cast<size_t>(a) * kExtWords + k] ^
                           ext[static_cast<size_t>(b) * kExtWords + k];
        if (y != 0) {
            const int l = 21 * (k + 1) + word_levels(y);
            return l < kMaxLevels ? l : kMaxLevels;
        }
    }
    return kMaxLevels;
}

// Last position e >= t with common_levels(t, e) >= d.
__device__ inline int run_end(const uint64_t* w, const uint64_t* ext, int n, int t, int d) {
    int lo = t, hi = n - 1;
    while (lo < hi) {
        const int mid = lo + (hi - lo + 1) / 2;
        if (common_levels(w, ext, t, mid) >= d) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}

// First position s <= t with common_levels(s, t) >= d.
__device__ inline int run_start(const uint64_t* w, const uint64_t* ext, int t, int d) {
    int lo = 0, hi = t;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (common_levels(w, ext, mid, t) >= d) {
            hi = mid;
        } else {
            lo = mid + 1;
        }
    }
    return lo;
}

// ---------------------------------------------------------------- inputs
__global__ void round_inputs_kernel(const double* __restrict__ px, const double* __restrict__ py,
                                    const double* __restrict__ pz, const double* __restrict__ m,
                                    double* __restrict__ bx, double* __restrict__ by,
                                    double* __restrict__ bz, double* __restrict__ bm, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        bx[i] = static_cast<double>(static_cast<float>(px[i]));
        by[i] = static_cast<double>(static_cast<float>(py[i]));
        bz[i] = static_cast<double>(static_cast<float>(pz[i]));
        bm[i] = static_cast<double>(static_cast<float>(m[i]));
    }
}

// ---------------------------------------------------------------- bounding cube
// barnes_hut.c:453-483. min/max are exact, so the reduction order is free.
__global__ __launch_bounds__(kThreads) void bbox_kernel(const double* __restrict__ bx,
                                                       const double* __restrict__ by,
                                                       const double* __restrict__ bz, int n,
                                                       double* __restrict__ partial,
                                                       unsigned* __restrict__ done,
                                                       double* __restrict__ root) {
    __shared__ double s[6][kThreads];
    double lo[3] = {bx[0], by[0], bz[0]};
    double hi[3] = {lo[0], lo[1], lo[2]};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const double p[3] = {bx[i], by[i], bz[i]};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (p[a] < lo[a]) lo[a] = p[a];
            if (p[a] > hi[a]) hi[a] = p[a];
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = lo[a];
        s[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < static_cast<unsigned>(off)) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double l = s[a][threadIdx.x + off];
                const double h = s[3 + a][threadIdx.x + off];
                if (l < s[a][threadIdx.x]) s[a][threadIdx.x] = l;
                if (h > s[3 + a][threadIdx.x]) s[3 + a][threadIdx.x] = h;
            }
        }
        __syncthreads();
    }
    __shared__ bool last;
    if (threadIdx.x == 0) {
        for (int a = 0; a < 6; ++a) {
            __hip_atomic_store(&partial[blockIdx.x * 6 + a], s[a][0], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        __threadfence();
        const unsigned prev = atomicAdd(done, 1u);
        last = (prev + 1 == gridDim.x);
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) {
        return;
    }
    __threadfence();
    double l[3], h[3];
    for (int a = 0; a < 3; ++a) {
        l[a] = __hip_atomic_load(&partial[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        h[a] = __hip_atomic_load(&partial[3 + a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    for (unsigned b = 1; b < gridDim.x; ++b) {
        for (int a = 0; a < 3; ++a) {
            const double pl = __hip_atomic_load(&partial[b * 6 + a], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            const double ph = __hip_atomic_load(&partial[b * 6 + 3 + a], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            if (pl < l[a]) l[a] = pl;
            if (ph > h[a]) h[a] = ph;
        }
    }
    double half = 0.5 * (h[0] - l[0]);
    const double hy = 0.5 * (h[1] - l[1]);
    const double hz = 0.5 * (h[2] - l[2]);
    if (hy > half) half = hy;
    if (hz > half) half = hz;
    if (half == 0.0) half = 0.5;
    half *= 1.0 + 1.0e-9;
    root[0] = 0.5 * (l[0] + h[0]);
    root[1] = 0.5 * (l[1] + h[1]);
    root[2] = 0.5 * (l[2] + h[2]);
    root[3] = half;
    *done = 0;
}

// ---------------------------------------------------------------- keys
__global__ void keys_kernel(const double* __restrict__ bx, const double* __restrict__ by,
                            const double* __restrict__ bz, int n, const double* __restrict__ root,
                            uint64_t* __restrict__ mkey, uint64_t* __restrict__ tkey,
                            uint64_t* __restrict__ sortkey, uint32_t* __restrict__ idx,
                            unsigned* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const double cx = root[0], cy = root[1], cz = root[2], h = root[3];
    const double x = bx[i], y = by[i], z = bz[i];
    const uint64_t mk = spread3(axis_code(x, cx, h)) | (spread3(axis_code(y, cy, h)) << 1) |
                        (spread3(axis_code(z, cz, h)) << 2);
    double ccx = cx, ccy = cy, ccz = cz, hh = h;
    const uint64_t tk = path_word(x, y, z, ccx, ccy, ccz, hh);
    mkey[i] = mk;
    tkey[i] = tk;
    sortkey[i] = mk;
    idx[i] = static_cast<uint32_t>(i);
    if (mk != tk) {
        atomicOr(&flags[FLAG_MISMATCH], 1u);
    }
}

// ext[pos*6+k] = octant digits of levels 21(k+1) .. 21(k+1)+20 of particle orig[pos].
__global__ void ext_words_kernel(const double* __restrict__ bx, const double* __restrict__ by,
                                 const double* __restrict__ bz, int n,
                                 const double* __restrict__ root,
                                 const uint32_t* __restrict__ orig, uint64_t* __restrict__ ext) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) {
        return;
    }
    const uint32_t i = orig[p];
    const double x = bx[i], y = by[i], z = bz[i];
    double cx = root[0], cy = root[1], cz = root[2], h = root[3];
    path_word(x, y, z, cx, cy, cz, h);
    for (int k = 0; k < kExtWords; ++k) {
        ext[static_cast<size_t>(p) * kExtWords + k] = path_word(x, y, z, cx, cy, cz, h);
    }
}

__global__ void gather_u64_kernel(const uint64_t* __restrict__ src, const uint32_t* __restrict__ idx,
                                  uint64_t* __restrict__ dst, uint32_t* __restrict__ iota, int n,
                                  int stride, int offset) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        dst[p] = src[static_cast<size_t>(idx[p]) * stride + offset];
        if (iota != nullptr) {
            iota[p] = static_cast<uint32_t>(p);
        }
    }
}

__global__ void compose_kernel(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                               uint32_t* __restrict__ out, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        out[p] = a[b[p]];
    }
}

__global__ void iota_kernel(uint32_t* __restrict__ out, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        out[p] = static_cast<uint32_t>(p);
    }
}

// Tree order = descending key order. asc_rank[k] = Morton rank of the k-th
// particle in ascending true order; mperm[r] = original index of Morton rank r.
__global__ void tree_order_kernel(const uint32_t* __restrict__ asc_rank,
                                  const uint32_t* __restrict__ mperm,
                                  const uint64_t* __restrict__ tkey_orig,
                                  const uint64_t* __restrict__ ext_asc,
                                  const double* __restrict__ bx, const double* __restrict__ by,
                                  const double* __restrict__ bz, const double* __restrict__ bm,
                                  int n, uint32_t* __restrict__ tperm, uint32_t* __restrict__ trank,
                                  uint64_t* __restrict__ tword, uint64_t* __restrict__ ext_tree,
                                  double* __restrict__ sx, double* __restrict__ sy,
                                  double* __restrict__ sz, double* __restrict__ sm,
                                  float4* __restrict__ s4) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) {
        return;
    }
    const int k = n - 1 - t;
    const uint32_t r = asc_rank != nullptr ? asc_rank[k] : static_cast<uint32_t>(k);
    const uint32_t o = mperm[r];
    tperm[t] = o;
    trank[t] = r;
    tword[t] = tkey_orig[o];
    if (ext_asc != nullptr) {
        for (int w = 0; w < kExtWords; ++w) {
            ext_tree[static_cast<size_t>(t) * kExtWords + w] =
                ext_asc[static_cast<size_t>(k) * kExtWords + w];
        }
    }
    const double x = bx[o], y = by[o], z = bz[o], m = bm[o];
    sx[t] = x;
    sy[t] = y;
    sz[t] = z;
    sm[t] = m;
    s4[t] = make_float4(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z),
                        static_cast<float>(m));
}

// ---------------------------------------------------------------- topology
__global__ void delta_kernel(const uint64_t* __restrict__ w, const uint64_t* __restrict__ ext,
                             int n, int32_t* __restrict__ delta, uint32_t* __restrict__ cnt,
                             unsigned* __restrict__ flags) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) {
        return;
    }
    const int dprev = t > 0 ? common_levels(w, ext, t - 1, t) : -1;
    const int dcur = t < n - 1 ? common_levels(w, ext, t, t + 1) : -1;
    delta[t] = dcur;
    const int c = t == 0 ? dcur + 1 : (dcur > dprev ? dcur - dprev : 0);
    cnt[t] = static_cast<uint32_t>(c > 0 ? c : 0);
    if (ext == nullptr && dcur >= 21) {
        atomicOr(&flags[FLAG_TIE], 1u);
    }
    if (dcur >= kMaxLevels) {
        atomicOr(&flags[FLAG_OVERFLOW], 1u);
    }
}

struct NodeSoA {
    double* cx;
    double* cy;
    double* cz;
    double* m;
    int32_t* skip;
    int32_t* leaf;
    int32_t* parent;
    uint32_t* nchild;
    uint32_t* arrive;
    int32_t* depth;
    int32_t* start;
    int32_t* count;
    int32_t* minpos;
};

__device__ inline int lo_of(const int32_t* delta, int t) {
    return t == 0 ? 0 : delta[t - 1] + 1;
}

__device__ inline int internal_index(const uint32_t* scan, const int32_t* delta, int s, int d) {
    return static_cast<int>(scan[s]) + s + (d - lo_of(delta, s));
}

__device__ inline int leaf_index(const uint32_t* scan, const uint32_t* cnt, int t) {
    return static_cast<int>(scan[t] + cnt[t]) + t;
}

__global__ void fill_kernel(const uint64_t* __restrict__ w, const uint64_t* __restrict__ ext,
                            const int32_t* __restrict__ delta, const uint32_t* __restrict__ cnt,
                            const uint32_t* __restrict__ scan, int n,
                            const double* __restrict__ sx, const double* __restrict__ sy,
                            const double* __restrict__ sz, const double* __restrict__ sm,
                            NodeSoA nd, unsigned* __restrict__ flags) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) {
        return;
    }
    const int lo = lo_of(delta, t);
    const int hi = delta[t];
    for (int d = lo; d <= hi; ++d) {
        const int idx = internal_index(scan, delta, t, d);
        const int e = run_end(w, ext, n, t, d);
        nd.skip[idx] = leaf_index(scan, cnt, e) + 1;
        nd.leaf[idx] = -1;
        nd.depth[idx] = d;
        nd.start[idx] = t;
        nd.count[idx] = e - t + 1;
        int parent = -1;
        if (d > 0) {
            parent = (d - 1 >= lo) ? idx - 1 : internal_index(scan, delta, run_start(w, ext, t, d - 1), d - 1);
        }
        nd.parent[idx] = parent;
        if (parent >= 0) {
            atomicAdd(&nd.nchild[parent], 1u);
        }
    }
    const int dprev = t > 0 ? delta[t - 1] : -1;
    const int depth = 1 + (dprev > hi ? dprev : hi);
    const int li = leaf_index(scan, cnt, t);
    nd.skip[li] = li + 1;
    nd.leaf[li] = t;
    nd.depth[li] = depth;
    nd.start[li] = t;
    nd.count[li] = 1;
    nd.cx[li] = sx[t];
    nd.cy[li] = sy[t];
    nd.cz[li] = sz[t];
    nd.m[li] = sm[t];
    nd.minpos[li] = t;
    int parent = -1;
    if (n > 1) {
        const int pd = depth - 1;
        parent = (pd >= lo && pd <= hi) ? li - 1 : internal_index(scan, delta, run_start(w, ext, t, pd), pd);
    }
    nd.parent[li] = parent;
    if (parent >= 0) {
        atomicAdd(&nd.nchild[parent], 1u);
    }
    atomicMax(&flags[FLAG_MAXDEPTH], static_cast<unsigned>(depth));
}

__device__ inline double ldc(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int32_t ldc(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void stc(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void stc(int32_t* p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// barnes_hut.c:214-253 (children summed in octant order 0..7; a zero-mass
// cell keeps the position of the particle that created it: the lowest Morton
// rank in the cell, or for the root the caller's particle 0, :504-512).
__global__ void com_kernel(const uint32_t* __restrict__ scan, const uint32_t* __restrict__ cnt,
                           const uint32_t* __restrict__ trank, int n, NodeSoA nd,
                           const double* __restrict__ sx, const double* __restrict__ sy,
                           const double* __restrict__ sz, const double* __restrict__ bx,
                           const double* __restrict__ by, const double* __restrict__ bz) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) {
        return;
    }
    int node = leaf_index(scan, cnt, t);
    int p = nd.parent[node];
    while (p >= 0) {
        __threadfence();
        const unsigned prev = atomicAdd(&nd.arrive[p], 1u);
        if (prev + 1 != nd.nchild[p]) {
            return;
        }
        __threadfence();
        int kids[8];
        int k = 0;
        const int end = nd.skip[p];
        for (int c = p + 1; c < end && k < 8; c = nd.skip[c]) {
            kids[k++] = c;
        }
        double mass = 0.0, x = 0.0, y = 0.0, z = 0.0;
        int minpos = -1;
        uint32_t minrank = 0xFFFFFFFFu;
        for (int q = k - 1; q >= 0; --q) {
            const int c = kids[q];
            const double cm = ldc(&nd.m[c]);
            mass += cm;
            x += cm * ldc(&nd.cx[c]);
            y += cm * ldc(&nd.cy[c]);
            z += cm * ldc(&nd.cz[c]);
            const int mp = ldc(&nd.minpos[c]);
            const uint32_t r = trank[mp];
            if (r < minrank) {
                minrank = r;
                minpos = mp;
            }
        }
        double ox, oy, oz;
        if (mass != 0.0) {
            ox = x / mass;
            oy = y / mass;
            oz = z / mass;
        } else if (p == 0) {
            ox = bx[0];
            oy = by[0];
            oz = bz[0];
        } else {
            ox = sx[minpos];
            oy = sy[minpos];
            oz = sz[minpos];
        }
        stc(&nd.m[p], mass);
        stc(&nd.cx[p], ox);
        stc(&nd.cy[p], oy);
        stc(&nd.cz[p], oz);
        stc(&nd.minpos[p], minpos);
        node = p;
        p = nd.parent[node];
    }
}

__global__ void pack_kernel(NodeSoA nd, int nodes, double root_h, NodeF64* __restrict__ n64,
                            NodeF32* __restrict__ n32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nodes) {
        return;
    }
    const double h = ldexp(root_h, -nd.depth[i]);
    NodeF64 a;
    a.cx = nd.cx[i];
    a.cy = nd.cy[i];
    a.cz = nd.cz[i];
    a.m = nd.m[i];
    a.h = h;
    a.skip = nd.skip[i];
    a.leaf = nd.leaf[i];
    n64[i] = a;
    NodeF32 b;
    b.cx = static_cast<float>(a.cx);
    b.cy = static_cast<float>(a.cy);
    b.cz = static_cast<float>(a.cz);
    b.m = static_cast<float>(a.m);
    b.size2 = static_cast<float>(4.0 * h * h);
    b.skip = a.skip;
    b.leaf = a.leaf;
    b.depth = nd.depth[i];
    n32[i] = b;
}

// ---------------------------------------------------------------- walk
template <bool FP64>
__global__ __launch_bounds__(kThreads) void walk_kernel(
    const NodeF64* __restrict__ n64, const NodeF32* __restrict__ n32, int nodes,
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const float4* __restrict__ s4, const uint32_t* __restrict__ tperm, int n, double G,
    double eps2, double theta2, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, unsigned long long* __restrict__ counters,
    float* __restrict__ features) {
    __shared__ unsigned long long block_counts[2];
    if (threadIdx.x < 2) {
        block_counts[threadIdx.x] = 0;
    }
    __syncthreads();
    const int t = blockIdx.x * kThreads + threadIdx.x;
    const bool valid = t < n;
    const int tc = valid ? t : n - 1;
    int skip_until = valid ? 0 : 0x7FFFFFFF;
    unsigned approx = 0, exact = 0;
    float ratio_sum = 0.f;

    if (FP64) {
        const double qx = sx[tc], qy = sy[tc], qz = sz[tc];
        double acx = 0.0, acy = 0.0, acz = 0.0;
        int i = 0;
        while (i < nodes) {
            const NodeF64 nd = n64[__builtin_amdgcn_readfirstlane(i)];
            const bool active = i >= skip_until;
            if (nd.leaf >= 0) {
                if (active && nd.leaf != t) {
                    const double dx = nd.cx - qx;
                    const double dy = nd.cy - qy;
                    const double dz = nd.cz - qz;
                    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                    const double inv3 = 1.0 / (r2 * sqrt(r2));
                    const double s = G * nd.m * inv3;
                    acx += s * dx;
                    acy += s * dy;
                    acz += s * dz;
                    ++exact;
                }
                i = i + 1;
            } else {
                bool descend = false;
                if (active) {
                    const double dx = nd.cx - qx;
                    const double dy = nd.cy - qy;
                    const double dz = nd.cz - qz;
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    const double size2 = 4.0 * nd.h * nd.h;
                    if (d2 > 0.0 && size2 < theta2 * d2) {
                        const double r2 = d2 + eps2;
                        const double inv3 = 1.0 / (r2 * sqrt(r2));
                        const double s = G * nd.m * inv3;
                        acx += s * dx;
                        acy += s * dy;
                        acz += s * dz;
                        ++approx;
                        skip_until = nd.skip;
                    } else {
                        descend = true;
                    }
                }
                i = __ballot(descend) != 0 ? i + 1 : nd.skip;
            }
        }
        if (valid) {
            const uint32_t o = tperm[t];
            ax[o] = acx;
            ay[o] = acy;
            az[o] = acz;
        }
    } else {
        const float4 q = s4[tc];
        const float Gf = static_cast<float>(G);
        const float e2 = static_cast<float>(eps2);
        const float th2 = static_cast<float>(theta2);
        float acx = 0.f, acy = 0.f, acz = 0.f;
        int i = 0;
        while (i < nodes) {
            const NodeF32 nd = n32[__builtin_amdgcn_readfirstlane(i)];
            const bool active = i >= skip_until;
            if (nd.leaf >= 0) {
                if (active && nd.leaf != t) {
                    const float dx = nd.cx - q.x;
                    const float dy = nd.cy - q.y;
                    const float dz = nd.cz - q.z;
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, e2)));
                    const float inv = __builtin_amdgcn_rsqf(r2);
                    const float s = (Gf * nd.m * inv) * (inv * inv);
                    acx = fmaf(s, dx, acx);
                    acy = fmaf(s, dy, acy);
                    acz = fmaf(s, dz, acz);
                    ++exact;
                }
                i = i + 1;
            } else {
                bool descend = false;
                if (active) {
                    const float dx = nd.cx - q.x;
                    const float dy = nd.cy - q.y;
                    const float dz = nd.cz - q.z;
                    const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    if (d2 > 0.f && nd.size2 < th2 * d2) {
                        const float r2 = d2 + e2;
                        const float inv = __builtin_amdgcn_rsqf(r2);
                        const float s = (Gf * nd.m * inv) * (inv * inv);
                        acx = fmaf(s, dx, acx);
                        acy = fmaf(s, dy, acy);
                        acz = fmaf(s, dz, acz);
                        ++approx;
                        if (features != nullptr) {
                            ratio_sum += sqrtf(d2 / nd.size2);
                        }
                        skip_until = nd.skip;
                    } else {
                        descend = true;
                    }
                }
                i = __ballot(descend) != 0 ? i + 1 : nd.skip;
            }
        }
        if (valid) {
            const uint32_t o = tperm[t];
            ax[o] = acx;
            ay[o] = acy;
            az[o] = acz;
            if (features != nullptr) {
                // per tree position: accepted cells, exact leaves, mean accepted d/size
                features[static_cast<size_t>(t) * 4 + 0] = static_cast<float>(approx);
                features[static_cast<size_t>(t) * 4 + 1] = static_cast<float>(exact);
                features[static_cast<size_t>(t) * 4 + 2] =
                    approx > 0 ? ratio_sum / static_cast<float>(approx) : 0.f;
                features[static_cast<size_t>(t) * 4 + 3] = static_cast<float>(o);
            }
        }
    }
    if (valid) {
        atomicAdd(&block_counts[0], static_cast<unsigned long long>(approx));
        atomicAdd(&block_counts[1], static_cast<unsigned long long>(exact));
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        atomicAdd(&counters[threadIdx.x], block_counts[threadIdx.x]);
    }
}

template <typename T>
T* B(Buf& b) {
    return b.as<T>();
}

}  // namespace


namespace {

__global__ void gather_rank_word0_kernel(const uint64_t* __restrict__ tkey,
                                         const uint32_t* __restrict__ mperm,
                                         const uint32_t* __restrict__ P, uint64_t* __restrict__ out,
                                         int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        out[p] = tkey[mperm[P[p]]];
    }
}

NodeSoA node_soa(BHState& s) {
    NodeSoA nd;
    nd.cx = B<double>(s.ncx);
    nd.cy = B<double>(s.ncy);
    nd.cz = B<double>(s.ncz);
    nd.m = B<double>(s.nm);
    nd.skip = B<int32_t>(s.nskip);
    nd.leaf = B<int32_t>(s.nleaf);
    nd.parent = B<int32_t>(s.nparent);
    nd.nchild = B<uint32_t>(s.nchild);
    nd.arrive = B<uint32_t>(s.narrive);
    nd.depth = B<int32_t>(s.ndepth);
    nd.start = B<int32_t>(s.nstart);
    nd.count = B<int32_t>(s.ncount);
    nd.minpos = B<int32_t>(s.nminpos);
    return nd;
}

}  // namespace

NBodySimProStatus bh_build(Engine* e, bool fp32_inputs, NBodySimProError* error) {
    BHState& s = e->bh;
    const size_t n = e->n;
    if (n == 0) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                 "cannot build a Barnes-Hut tree for an empty system");
    }
    if (n > 0x3FFFFFFFull) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "particle count exceeds 2^30");
    }
    const int ni = static_cast<int>(n);
    hipStream_t st = e->stream;
    const unsigned g = grid_for(n, kThreads);

    NBX_HIP(error, hipEventRecord(e->ev[0], st));
    const double *bx = e->px, *by = e->py, *bz = e->pz, *bm = e->m;
    if (fp32_inputs) {
        NBX_TRY(s.bx.need<double>(n, error));
        NBX_TRY(s.by.need<double>(n, error));
        NBX_TRY(s.bz.need<double>(n, error));
        NBX_TRY(s.bm.need<double>(n, error));
        round_inputs_kernel<<<g, kThreads, 0, st>>>(e->px, e->py, e->pz, e->m, B<double>(s.bx),
                                                    B<double>(s.by), B<double>(s.bz),
                                                    B<double>(s.bm), ni);
        bx = B<double>(s.bx);
        by = B<double>(s.by);
        bz = B<double>(s.bz);
        bm = B<double>(s.bm);
    }
    const unsigned bb_blocks = g < 1024u ? g : 1024u;
    NBX_TRY(s.partial.need<double>(6 * 1024 + 8, error));
    NBX_TRY(s.counter.need<unsigned>(4, error));
    NBX_TRY(s.flags.need<unsigned>(FLAG_COUNT, error));
    NBX_TRY(s.mkey.need<uint64_t>(n, error));
    NBX_TRY(s.tkey.need<uint64_t>(n, error));
    NBX_TRY(s.sk0.need<uint64_t>(n, error));
    NBX_TRY(s.sk1.need<uint64_t>(n, error));
    NBX_TRY(s.sv0.need<uint32_t>(n, error));
    NBX_TRY(s.sv1.need<uint32_t>(n, error));
    NBX_TRY(s.tmp32.need<uint32_t>(n, error));
    NBX_TRY(s.mperm.need<uint32_t>(n, error));
    NBX_TRY(s.mkey_sorted.need<uint64_t>(n, error));
    NBX_TRY(s.tperm.need<uint32_t>(n, error));
    NBX_TRY(s.trank.need<uint32_t>(n, error));
    NBX_TRY(s.tword.need<uint64_t>(n, error));
    NBX_TRY(s.delta.need<int32_t>(n, error));
    NBX_TRY(s.cnt.need<uint32_t>(n, error));
    NBX_TRY(s.scan.need<uint32_t>(n + 1, error));
    NBX_TRY(s.sx.need<double>(n, error));
    NBX_TRY(s.sy.need<double>(n, error));
    NBX_TRY(s.sz.need<double>(n, error));
    NBX_TRY(s.sm.need<double>(n, error));
    NBX_TRY(s.s4.need<float4>(n, error));
    NBX_TRY(s.counters.need<unsigned long long>(2, error));

    double* root = B<double>(s.partial) + 6 * 1024;
    unsigned* flags = B<unsigned>(s.flags);
    NBX_HIP(error, hipMemsetAsync(flags, 0, FLAG_COUNT * sizeof(unsigned), st));
    NBX_HIP(error, hipMemsetAsync(s.counter.p, 0, 4 * sizeof(unsigned), st));
    bbox_kernel<<<bb_blocks, kThreads, 0, st>>>(bx, by, bz, ni, B<double>(s.partial),
                                                B<unsigned>(s.counter), root);
    keys_kernel<<<g, kThreads, 0, st>>>(bx, by, bz, ni, root, B<uint64_t>(s.mkey),
                                        B<uint64_t>(s.tkey), B<uint64_t>(s.sk0),
                                        B<uint32_t>(s.sv0), flags);
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[1], st));

    // Morton order: stable sort of (key, original index) on the 63 key bits.
    NBX_TRY(radix_sort_pairs(st, B<uint64_t>(s.sk0), B<uint32_t>(s.sv0), B<uint64_t>(s.sk1),
                             B<uint32_t>(s.sv1), n, 63, s.hist, error));
    NBX_HIP(error, hipMemcpyAsync(s.mkey_sorted.p, s.sk0.p, n * sizeof(uint64_t),
                                  hipMemcpyDeviceToDevice, st));
    NBX_HIP(error, hipMemcpyAsync(s.mperm.p, s.sv0.p, n * sizeof(uint32_t),
                                  hipMemcpyDeviceToDevice, st));
    unsigned hflags[FLAG_COUNT];
    NBX_HIP(error, hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, st));
    NBX_HIP(error, hipMemcpyAsync(s.root, root, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    NBX_HIP(error, hipStreamSynchronize(st));
    s.mismatch_path = hflags[FLAG_MISMATCH] != 0;
    s.tie_path = false;

    // asc_rank[k]: Morton rank of the k-th particle in ascending true-path order.
    const uint32_t* asc_rank = nullptr;
    if (s.mismatch_path) {
        gather_u64_kernel<<<g, kThreads, 0, st>>>(B<uint64_t>(s.tkey), B<uint32_t>(s.mperm),
                                                  B<uint64_t>(s.sk0), B<uint32_t>(s.sv0), ni, 1, 0);
        NBX_TRY(radix_sort_pairs(st, B<uint64_t>(s.sk0), B<uint32_t>(s.sv0), B<uint64_t>(s.sk1),
                                 B<uint32_t>(s.sv1), n, 63, s.hist, error));
        NBX_HIP(error, hipMemcpyAsync(s.tmp32.p, s.sv0.p, n * sizeof(uint32_t),
                                      hipMemcpyDeviceToDevice, st));
        asc_rank = B<uint32_t>(s.tmp32);
    }
    tree_order_kernel<<<g, kThreads, 0, st>>>(
        asc_rank, B<uint32_t>(s.mperm), B<uint64_t>(s.tkey), nullptr, bx, by, bz, bm, ni,
        B<uint32_t>(s.tperm), B<uint32_t>(s.trank), B<uint64_t>(s.tword), nullptr, B<double>(s.sx),
        B<double>(s.sy), B<double>(s.sz), B<double>(s.sm), B<float4>(s.s4));
    delta_kernel<<<g, kThreads, 0, st>>>(B<uint64_t>(s.tword), nullptr, ni, B<int32_t>(s.delta),
                                         B<uint32_t>(s.cnt), flags);
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, st));
    NBX_HIP(error, hipStreamSynchronize(st));

    const uint64_t* ext_tree = nullptr;
    if (hflags[FLAG_TIE] != 0) {
        // Some particles share all 21 levels: extend every path by six more
        // 21-level words and LSD-sort by (word0, ext0, ..., ext5), least
        // significant word first, starting from Morton order (stable).
        s.tie_path = true;
        NBX_TRY(s.ext.need<uint64_t>(3 * n * kExtWords, error));
        uint64_t* ext_m = B<uint64_t>(s.ext);
        uint64_t* ext_a = ext_m + n * kExtWords;
        uint64_t* ext_t = ext_a + n * kExtWords;
        ext_words_kernel<<<g, kThreads, 0, st>>>(bx, by, bz, ni, root, B<uint32_t>(s.mperm), ext_m);
        uint32_t* P = B<uint32_t>(s.tmp32);
        iota_kernel<<<g, kThreads, 0, st>>>(P, ni);
        for (int k = kExtWords - 1; k >= 0; --k) {
            gather_u64_kernel<<<g, kThreads, 0, st>>>(ext_m, P, B<uint64_t>(s.sk0), nullptr, ni,
                                                      kExtWords, k);
            NBX_HIP(error, hipMemcpyAsync(s.sv0.p, P, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
            NBX_TRY(radix_sort_pairs(st, B<uint64_t>(s.sk0), B<uint32_t>(s.sv0), B<uint64_t>(s.sk1),
                                     B<uint32_t>(s.sv1), n, 63, s.hist, error));
            NBX_HIP(error, hipMemcpyAsync(P, s.sv0.p, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        }
        gather_rank_word0_kernel<<<g, kThreads, 0, st>>>(B<uint64_t>(s.tkey), B<uint32_t>(s.mperm),
                                                         P, B<uint64_t>(s.sk0), ni);
        NBX_HIP(error, hipMemcpyAsync(s.sv0.p, P, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        NBX_TRY(radix_sort_pairs(st, B<uint64_t>(s.sk0), B<uint32_t>(s.sv0), B<uint64_t>(s.sk1),
                                 B<uint32_t>(s.sv1), n, 63, s.hist, error));
        NBX_HIP(error, hipMemcpyAsync(P, s.sv0.p, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        for (int k = 0; k < kExtWords; ++k) {
            // ext_a[p*6+k] = ext_m[P[p]*6+k], via sk0 as staging
            gather_u64_kernel<<<g, kThreads, 0, st>>>(ext_m, P, B<uint64_t>(s.sk0), nullptr, ni,
                                                      kExtWords, k);
            NBX_HIP(error, hipMemcpy2DAsync(ext_a + k, kExtWords * sizeof(uint64_t), s.sk0.p,
                                            sizeof(uint64_t), sizeof(uint64_t), n,
                                            hipMemcpyDeviceToDevice, st));
        }
        NBX_HIP(error, hipMemsetAsync(flags, 0, FLAG_COUNT * sizeof(unsigned), st));
        tree_order_kernel<<<g, kThreads, 0, st>>>(
            P, B<uint32_t>(s.mperm), B<uint64_t>(s.tkey), ext_a, bx, by, bz, bm, ni,
            B<uint32_t>(s.tperm), B<uint32_t>(s.trank), B<uint64_t>(s.tword), ext_t,
            B<double>(s.sx), B<double>(s.sy), B<double>(s.sz), B<double>(s.sm), B<float4>(s.s4));
        delta_kernel<<<g, kThreads, 0, st>>>(B<uint64_t>(s.tword), ext_t, ni, B<int32_t>(s.delta),
                                             B<uint32_t>(s.cnt), flags);
        NBX_HIP(error, hipGetLastError());
        ext_tree = ext_t;
    }
    NBX_TRY(exclusive_scan_u32(st, B<uint32_t>(s.cnt), B<uint32_t>(s.scan), n, s.scan_part, error));
    uint32_t internal = 0;
    NBX_HIP(error, hipMemcpyAsync(&internal, B<uint32_t>(s.scan) + n, sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, st));
    NBX_HIP(error, hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, st));
    NBX_HIP(error, hipStreamSynchronize(st));
    if (hflags[FLAG_OVERFLOW] != 0) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_OVERFLOW,
                 "Barnes-Hut tree exceeded maximum depth (duplicate positions?)");
    }
    NBX_HIP(error, hipEventRecord(e->ev[2], st));

    const size_t nodes = n + internal;
    NBX_TRY(s.ncx.need<double>(nodes, error));
    NBX_TRY(s.ncy.need<double>(nodes, error));
    NBX_TRY(s.ncz.need<double>(nodes, error));
    NBX_TRY(s.nm.need<double>(nodes, error));
    NBX_TRY(s.nskip.need<int32_t>(nodes, error));
    NBX_TRY(s.nleaf.need<int32_t>(nodes, error));
    NBX_TRY(s.nparent.need<int32_t>(nodes, error));
    NBX_TRY(s.nchild.need<uint32_t>(nodes, error));
    NBX_TRY(s.narrive.need<uint32_t>(nodes, error));
    NBX_TRY(s.ndepth.need<int32_t>(nodes, error));
    NBX_TRY(s.nstart.need<int32_t>(nodes, error));
    NBX_TRY(s.ncount.need<int32_t>(nodes, error));
    NBX_TRY(s.nminpos.need<int32_t>(nodes, error));
    NBX_TRY(s.n64.need<NodeF64>(nodes, error));
    NBX_TRY(s.n32.need<NodeF32>(nodes, error));
    NBX_HIP(error, hipMemsetAsync(s.nchild.p, 0, nodes * sizeof(uint32_t), st));
    NBX_HIP(error, hipMemsetAsync(s.narrive.p, 0, nodes * sizeof(uint32_t), st));
    NodeSoA nd = node_soa(s);
    fill_kernel<<<g, kThreads, 0, st>>>(B<uint64_t>(s.tword), ext_tree, B<int32_t>(s.delta),
                                        B<uint32_t>(s.cnt), B<uint32_t>(s.scan), ni,
                                        B<double>(s.sx), B<double>(s.sy), B<double>(s.sz),
                                        B<double>(s.sm), nd, flags);
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[3], st));
    com_kernel<<<g, kThreads, 0, st>>>(B<uint32_t>(s.scan), B<uint32_t>(s.cnt),
                                       B<uint32_t>(s.trank), ni, nd, B<double>(s.sx),
                                       B<double>(s.sy), B<double>(s.sz), bx, by, bz);
    NBX_HIP(error, hipGetLastError());
    pack_kernel<<<grid_for(nodes, kThreads), kThreads, 0, st>>>(nd, static_cast<int>(nodes),
                                                                 s.root[3], B<NodeF64>(s.n64),
                                                                 B<NodeF32>(s.n32));
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[4], st));
    NBX_HIP(error, hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, st));
    NBX_HIP(error, hipStreamSynchronize(st));
    s.n = n;
    s.internal = internal;
    s.nodes = nodes;
    s.max_depth = static_cast<int>(hflags[FLAG_MAXDEPTH]);
    e->stats.node_count = nodes;
    e->stats.leaf_count = n;
    e->stats.internal_node_count = internal;
    e->stats.maximum_depth = static_cast<size_t>(s.max_depth) + 1;
    e->stats.accepted_approximations = 0;
    e->stats.exact_interactions = 0;
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus bh_walk(Engine* e, const NBodySimProGravity* g, bool fp64_exact, float* features,
                          NBodySimProError* error) {
    BHState& s = e->bh;
    hipStream_t st = e->stream;
    const int ni = static_cast<int>(s.n);
    NBX_HIP(error, hipMemsetAsync(s.counters.p, 0, 2 * sizeof(unsigned long long), st));
    NBX_HIP(error, hipEventRecord(e->ev[5], st));
    const double theta2 = e->theta * e->theta;
    const unsigned grid = grid_for(s.n, kThreads);
    if (fp64_exact) {
        walk_kernel<true><<<grid, kThreads, 0, st>>>(
            B<NodeF64>(s.n64), B<NodeF32>(s.n32), static_cast<int>(s.nodes), B<double>(s.sx),
            B<double>(s.sy), B<double>(s.sz), B<float4>(s.s4), B<uint32_t>(s.tperm), ni,
            g->gravitational_constant, g->softening_squared, theta2, e->ax, e->ay, e->az,
            B<unsigned long long>(s.counters), nullptr);
    } else {
        walk_kernel<false><<<grid, kThreads, 0, st>>>(
            B<NodeF64>(s.n64), B<NodeF32>(s.n32), static_cast<int>(s.nodes), B<double>(s.sx),
            B<double>(s.sy), B<double>(s.sz), B<float4>(s.s4), B<uint32_t>(s.tperm), ni,
            g->gravitational_constant, g->softening_squared, theta2, e->ax, e->ay, e->az,
            B<unsigned long long>(s.counters), features);
    }
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[6], st));
    unsigned long long counts[2] = {0, 0};
    NBX_HIP(error, hipMemcpyAsync(counts, s.counters.p, sizeof(counts), hipMemcpyDeviceToHost, st));
    NBX_HIP(error, hipStreamSynchronize(st));
    e->stats.accepted_approximations = counts[0];
    e->stats.exact_interactions = counts[1];
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus barnes_hut(Engine* e, const NBodySimProGravity* g, bool fp64_exact,
                             NBodySimProError* error) {
    NBX_TRY(bh_build(e, !fp64_exact, error));
    NBX_TRY(bh_walk(e, g, fp64_exact, nullptr, error));
    float t[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = 1; k <= 6; ++k) {
        NBX_HIP(error, hipEventElapsedTime(&t[k], e->ev[0], e->ev[k]));
    }
    NBXTimings tm{};
    tm.bbox_keys_ms = t[1];
    tm.sort_ms = t[2] - t[1];
    tm.tree_ms = t[3] - t[2];
    tm.com_ms = t[4] - t[3];
    tm.walk_ms = t[6] - t[5];
    tm.force_ms = t[6];
    tm.kernel_ms = tm.walk_ms;
    const double interactions = static_cast<double>(e->stats.accepted_approximations +
                                                    e->stats.exact_interactions);
    tm.kernel_flops = 20.0 * interactions;
    // walk bytes: every query reads its position and writes 3 doubles; node
    // records are counted once per launch (they stay in L2 across waves).
    tm.kernel_bytes = static_cast<double>(e->bh.n) * ((fp64_exact ? 24.0 : 16.0) + 24.0 + 4.0) +
                      static_cast<double>(e->bh.nodes) * (fp64_exact ? sizeof(NodeF64) : sizeof(NodeF32));
    e->timings = tm;
    e->stats.build_time_seconds = t[4] * 1e-3;
    e->stats.evaluation_time_seconds = tm.walk_ms * 1e-3;
    return N_BODY_SIM_PRO_STATUS_OK;
}

}  // namespace nbx
