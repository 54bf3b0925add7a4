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
// position t starts the cells of depths (delta[t-1], delta[t]] and ends those
// of depths (delta[t], delta[t-1]]. A scan of the start counts gives every
// node's pre-order index; cells are also numbered depth-major ("level
// index"), and the k-th start and k-th end of one depth belong to the same
// cell, which gives each cell's last position (its skip pointer) and parent
// without searching. That pre-order with children 7..0 is exactly the order in
// which the reference's LIFO walk pops nodes (children pushed 0..7,
// barnes_hut.c:617-634), so a walk that follows skip pointers visits, for
// every query, the same nodes in the same order.
//
// Centre of mass. Computed in FP64 from the children in octant order 0..7
// (barnes_hut.c:230-244), one thread per cell, deepest level first: each
// depth is one contiguous range of level indices; each large level is one
// launch and the small levels near the root run in a single one-block launch.
// A cell reads its children through an 8-slot child table written by the
// fill, so results are deterministic and bit-identical to the reference.
//
// Walks. The FP64 walk (and the FP32 walk with reference lists): one wave =
// 64 consecutive queries in tree order; the node index is wave-uniform (node
// records come in through scalar loads). Each lane keeps "skip_until": it is
// active at node i iff i >= skip_until. Active lanes apply the reference's
// own test (d^2 > 0 and 4h^2 < theta^2 d^2, unsoftened) and either take the
// monopole/leaf term (then skip the subtree) or descend. The wave moves to i+1
// if any active lane descends, else to skip[i]. Every lane thus sees exactly
// its reference interaction set, in reference order, with no stack. The FP64
// variant uses the reference arithmetic and is bit-identical. The default FP32
// walk is a group walk over the level-indexed cells (group_walk_kernel).
#include "nbx_internal.h"

#include <cstring>
#include <type_traits>
#include <vector>

namespace nbx {
namespace {

constexpr int kThreads = 256;
constexpr int kExtWords = 6;
constexpr int kMaxLevels = 129;  // a shared depth-129 cell means depth > 128 (barnes_hut.c:149-153)
enum { FLAG_MISMATCH = 0, FLAG_TIE = 1, FLAG_OVERFLOW = 2, FLAG_LONG_RUN = 3 };

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

// Octant digit of tree position t at level `level` (the octant of its depth
// level+1 cell inside its depth `level` cell).
__device__ inline int digit_at(const uint64_t* __restrict__ w, const uint64_t* __restrict__ ext,
                               int t, int level) {
    const uint64_t word = level < 21 ? w[t]
                                     : ext[static_cast<size_t>(t) * kExtWords + (level / 21 - 1)];
    return static_cast<int>((word >> (60 - 3 * (level % 21))) & 7u);
}

// ---------------------------------------------------------------- inputs and bounding cube
// Build inputs as one (x, y, z, m) record per particle (type In): the
// FP32-rounded values as float4 (hip_barnes_hut; every FP64 operation of the
// build sees them widened back, exactly) or the FP64 values as double4 (the
// FP64 walk). Half the bytes for the FP32 build in every pass that moves them.
__device__ inline double4 widen(const double4& v) {
    return v;
}

__device__ inline double4 widen(const float4& v) {
    return make_double4(static_cast<double>(v.x), static_cast<double>(v.y),
                        static_cast<double>(v.z), static_cast<double>(v.w));
}

template <typename In>
__device__ inline In build_input(const double* __restrict__ px, const double* __restrict__ py,
                                 const double* __restrict__ pz, const double* __restrict__ m,
                                 int i) {
    if constexpr (std::is_same_v<In, float4>) {
        return make_float4(static_cast<float>(px[i]), static_cast<float>(py[i]),
                           static_cast<float>(pz[i]), static_cast<float>(m[i]));
    } else {
        return make_double4(px[i], py[i], pz[i], m[i]);
    }
}

// barnes_hut.c:453-483. min/max are exact, so the reduction order is free.
// The same grid-stride pass writes the build inputs.
constexpr int kBBoxBlocks = 512;

template <typename In>
__global__ __launch_bounds__(kThreads) void bbox_partial_kernel(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ m, In* __restrict__ b4, int n, double* __restrict__ partial) {
    __shared__ double s[6][kThreads];
    const double4 p0 = widen(build_input<In>(px, py, pz, m, 0));
    double lo[3] = {p0.x, p0.y, p0.z};
    double hi[3] = {lo[0], lo[1], lo[2]};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const In v = build_input<In>(px, py, pz, m, i);
        b4[i] = v;
        const double4 q = widen(v);
        const double p[3] = {q.x, q.y, q.z};
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
    if (threadIdx.x < 6) {
        partial[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
    }
}

// Also clears the build's flag words and the walk's counters (interactions,
// deferred queries; no fill launches of their own inside the evaluation).
__global__ __launch_bounds__(kThreads) void bbox_final_kernel(const double* __restrict__ partial,
                                                             int blocks,
                                                             double* __restrict__ root,
                                                             unsigned* __restrict__ flags,
                                                             int flag_words,
                                                             unsigned long long* __restrict__ counters) {
    __shared__ double s[6][kThreads];
    for (int k = threadIdx.x; k < flag_words; k += kThreads) {
        flags[k] = 0u;
    }
    if (threadIdx.x < 3) {
        counters[threadIdx.x] = 0ull;
    }
    double lo[3] = {partial[0], partial[1], partial[2]};
    double hi[3] = {partial[3], partial[4], partial[5]};
    for (int b = threadIdx.x; b < blocks; b += kThreads) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double l = partial[b * 6 + a], h = partial[b * 6 + 3 + a];
            if (l < lo[a]) lo[a] = l;
            if (h > hi[a]) hi[a] = h;
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
    if (threadIdx.x == 0) {
        const double l[3] = {s[0][0], s[1][0], s[2][0]};
        const double h[3] = {s[3][0], s[4][0], s[5][0]};
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
    }
}

// ---------------------------------------------------------------- keys
template <typename In>
__global__ void keys_kernel(const In* __restrict__ b4, int n, const double* __restrict__ root,
                            uint64_t* __restrict__ mkey, uint64_t* __restrict__ tkey,
                            uint64_t* __restrict__ sortkey, uint32_t* __restrict__ idx,
                            unsigned* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const double cx = root[0], cy = root[1], cz = root[2], h = root[3];
    const double4 p = widen(b4[i]);
    const double x = p.x, y = p.y, z = p.z;
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

// The Morton sort runs its radix passes on the top kSortTopBits key bits
// only (levels 0..11: four of the seven 9-bit passes); sort_finish_kernel
// then orders each run of equal top bits by the low kSortLowBits, stably.
// The runs are short (two-galaxy input: at most 10 particles at 1M, 53 at
// 10M). A run longer than kFinishRun raises FLAG_LONG_RUN, and the host sorts
// again on all 63 bits after reading the flags.
constexpr int kSortLowBits = 27;
constexpr int kSortTopBits = 63 - kSortLowBits;
constexpr int kFinishRun = 64;
static_assert(kSortTopBits % 9 == 0, "whole 9-bit radix passes");

// Input: sorted (stably, from index order) by key >> kSortLowBits. Element p
// goes to its run's first position plus the number of run members before it
// in (low bits, position) order. Elements of a run longer than kFinishRun
// stay where they are (the output is still a permutation) and raise the flag.
__global__ __launch_bounds__(kThreads) void sort_finish_kernel(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, int n,
    unsigned* __restrict__ flags) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= n) {
        return;
    }
    const uint64_t key = kin[p];
    const uint64_t top = key >> kSortLowBits;
    const uint64_t low = key & ((1ull << kSortLowBits) - 1);
    // the run's bounds, at most kFinishRun steps to each side
    int a = p;
    while (a > 0 && p - a < kFinishRun && (kin[a - 1] >> kSortLowBits) == top) {
        --a;
    }
    int b = p;
    while (b < n - 1 && b - p < kFinishRun && (kin[b + 1] >> kSortLowBits) == top) {
        ++b;
    }
    int dst = p;
    if (b - a + 1 > kFinishRun) {
        // every member of a run this long sees a window of more than
        // kFinishRun positions, so the whole run stays in place
        atomicOr(&flags[FLAG_LONG_RUN], 1u);
    } else {
        int rank = 0;
        for (int q = a; q <= b; ++q) {
            const uint64_t lq = kin[q] & ((1ull << kSortLowBits) - 1);
            rank += (lq < low || (lq == low && q < p)) ? 1 : 0;
        }
        dst = a + rank;
    }
    kout[dst] = key;
    vout[dst] = vin[p];
}

// ext[pos*6+k] = octant digits of levels 21(k+1) .. 21(k+1)+20 of particle orig[pos].
template <typename In>
__global__ void ext_words_kernel(const In* __restrict__ b4, int n,
                                 const double* __restrict__ root,
                                 const uint32_t* __restrict__ orig, uint64_t* __restrict__ ext) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) {
        return;
    }
    const double4 q = widen(b4[orig[p]]);
    const double x = q.x, y = q.y, z = q.z;
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
// Without asc_rank the order is the Morton order and, unless the keys_kernel
// flagged a true-path key that differs from its Morton key (then this pass is
// redone with asc_rank), the true keys are the sorted Morton keys: read in
// order instead of gathered. The tree-order inputs: FP32 s4 always, FP64 sp4
// for FP64 inputs only (an FP32 build widens s4 where it needs FP64).
template <typename In>
__global__ void tree_order_kernel(const uint32_t* __restrict__ asc_rank,
                                  const uint32_t* __restrict__ mperm,
                                  const uint64_t* __restrict__ mkey_sorted,
                                  const uint64_t* __restrict__ tkey_orig,
                                  const uint64_t* __restrict__ ext_asc,
                                  const In* __restrict__ b4, int n, uint32_t* __restrict__ tperm, uint32_t* __restrict__ trank,
                                  uint64_t* __restrict__ tword, uint64_t* __restrict__ ext_tree,
                                  double4* __restrict__ sp4, float4* __restrict__ s4) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) {
        return;
    }
    const int k = n - 1 - t;
    const uint32_t r = asc_rank != nullptr ? asc_rank[k] : static_cast<uint32_t>(k);
    const uint32_t o = mperm[r];
    tperm[t] = o;
    trank[t] = r;
    tword[t] = asc_rank == nullptr ? mkey_sorted[r] : tkey_orig[o];
    if (ext_asc != nullptr) {
        for (int w = 0; w < kExtWords; ++w) {
            ext_tree[static_cast<size_t>(t) * kExtWords + w] =
                ext_asc[static_cast<size_t>(k) * kExtWords + w];
        }
    }
    const In p = b4[o];
    if constexpr (std::is_same_v<In, float4>) {
        s4[t] = p;
    } else {
        sp4[t] = p;
        s4[t] = make_float4(static_cast<float>(p.x), static_cast<float>(p.y),
                            static_cast<float>(p.z), static_cast<float>(p.w));
    }
}

// ---------------------------------------------------------------- topology
// Internal cells are numbered by level index L = doff[d] + k: the k-th cell of
// depth d in tree order (doff = exclusive prefix of the per-depth counts), so
// each depth is one contiguous range for the centre-of-mass pass. Tree
// position t starts the cells of depths (delta[t-1], delta[t]] and ends those
// of depths (delta[t], delta[t-1]]; cells of one depth are disjoint, so the
// k-th start of depth d and the k-th end of depth d belong to the same cell.
// The topology kernels give each workgroup one tile of kTopoTile consecutive
// positions: per-tile counts of starts and ends per depth, scanned over the
// tiles, plus wave ballots inside a tile, give every start and end its rank
// in position order, with no search and no shared global counters.
constexpr int kTopoItems = 4;
constexpr int kTopoTile = kThreads * kTopoItems;
constexpr int kTopoWaves = kThreads / 64;

// Internal cells by level index.
struct InternalSoA {
    double4* com;     // FP64 centre of mass (x, y, z) and mass
    int4* rec;        // node index, skip (node index past the subtree), first and last tree position
    int32_t* depth;
};

__device__ inline int wave_min_int(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    return v;
}

__device__ inline int wave_max_int(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    return v;
}

// delta[t] = common levels of tree positions t and t+1 (-1 for the last);
// cnt[t] = cells started at t; tile_hist[(k * kMaxLevels + d) * tiles + b] =
// cells of depth d started (k = 0) or ended (k = 1) in tile b.
__global__ __launch_bounds__(kThreads) void delta_kernel(const uint64_t* __restrict__ w,
                                                        const uint64_t* __restrict__ ext, int n,
                                                        int tiles, int32_t* __restrict__ delta,
                                                        uint32_t* __restrict__ cnt,
                                                        unsigned* __restrict__ flags,
                                                        unsigned* __restrict__ tile_hist) {
    __shared__ unsigned hist[2 * kMaxLevels];
    __shared__ unsigned tile_flags;
    for (int d = threadIdx.x; d < 2 * kMaxLevels; d += kThreads) {
        hist[d] = 0;
    }
    if (threadIdx.x == 0) {
        tile_flags = 0;
    }
    __syncthreads();
    unsigned mine = 0;
    for (int i = 0; i < kTopoItems; ++i) {
        const int t = blockIdx.x * kTopoTile + i * kThreads + threadIdx.x;
        if (t >= n) {
            break;
        }
        const int dprev = t > 0 ? common_levels(w, ext, t - 1, t) : -1;
        const int dcur = t < n - 1 ? common_levels(w, ext, t, t + 1) : -1;
        delta[t] = dcur;
        const int c = dcur > dprev ? dcur - dprev : 0;
        cnt[t] = static_cast<uint32_t>(c);
        const int top = dcur < kMaxLevels ? dcur : kMaxLevels - 1;
        for (int d = dprev + 1; d <= top; ++d) {
            atomicAdd(&hist[d], 1u);
        }
        const int etop = dprev < kMaxLevels ? dprev : kMaxLevels - 1;
        for (int d = dcur + 1; d <= etop; ++d) {
            atomicAdd(&hist[kMaxLevels + d], 1u);
        }
        if (ext == nullptr && dcur >= 21) {
            mine |= 1u;
        }
        if (dcur >= kMaxLevels) {
            mine |= 2u;
        }
    }
    if (mine != 0) {
        atomicOr(&tile_flags, mine);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < 2 * kMaxLevels; d += kThreads) {
        tile_hist[static_cast<size_t>(d) * tiles + blockIdx.x] = hist[d];
    }
    if (threadIdx.x == 0) {
        if (tile_flags & 1u) {
            atomicOr(&flags[FLAG_TIE], 1u);
        }
        if (tile_flags & 2u) {
            atomicOr(&flags[FLAG_OVERFLOW], 1u);
        }
    }
}

// One workgroup per (kind, depth) row: tile_base[row * tiles + b] = sum of
// tile_hist[row * tiles + b'] over b' < b; dcount[d] = the depth's cell count.
__global__ __launch_bounds__(kThreads) void depth_scan_kernel(const unsigned* __restrict__ tile_hist,
                                                             int tiles,
                                                             unsigned* __restrict__ tile_base,
                                                             unsigned* __restrict__ dcount) {
    __shared__ unsigned wave_sum[kTopoWaves];
    const int row_id = blockIdx.x;
    const unsigned* row = tile_hist + static_cast<size_t>(row_id) * tiles;
    unsigned* out = tile_base + static_cast<size_t>(row_id) * tiles;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned running = 0;
    for (int base = 0; base < tiles; base += kThreads) {
        const int i = base + threadIdx.x;
        const unsigned v = i < tiles ? row[i] : 0u;
        unsigned x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned y = __shfl_up(x, off, 64);
            if (lane >= off) {
                x += y;
            }
        }
        if (lane == 63) {
            wave_sum[wv] = x;
        }
        __syncthreads();
        unsigned before = running, total = 0;
#pragma unroll
        for (int k = 0; k < kTopoWaves; ++k) {
            const unsigned ws = wave_sum[k];
            before += k < wv ? ws : 0u;
            total += ws;
        }
        if (i < tiles) {
            out[i] = before + x - v;
        }
        running += total;
        __syncthreads();
    }
    if (threadIdx.x == 0 && row_id < kMaxLevels) {
        dcount[row_id] = running;
    }
}

// Exclusive offsets of the per-depth level-index ranges, into LDS, by the
// block's first wave (depth pairs per lane, one wave scan). Callers
// synchronise before reading them.
static_assert(kMaxLevels <= 2 * 64 + 1, "depth offsets: two depths per lane plus one");
__device__ inline void depth_offsets(const unsigned* __restrict__ dcount, unsigned* doff) {
    if (threadIdx.x >= 64) {
        return;
    }
    const int l = threadIdx.x;
    const unsigned v0 = 2 * l < kMaxLevels ? dcount[2 * l] : 0u;
    const unsigned v1 = 2 * l + 1 < kMaxLevels ? dcount[2 * l + 1] : 0u;
    const unsigned pair = v0 + v1;
    unsigned incl = pair;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned y = __shfl_up(incl, off, 64);
        if (l >= off) {
            incl += y;
        }
    }
    const unsigned excl = incl - pair;
    if (2 * l < kMaxLevels) {
        doff[2 * l] = excl;
    }
    if (2 * l + 1 < kMaxLevels) {
        doff[2 * l + 1] = excl + v0;
    }
    if (l == 63 && 2 * 64 < kMaxLevels) {
        doff[2 * 64] = incl;
    }
}

// Per-tile rank bookkeeping shared by ends_kernel and fill_kernel, which run
// one tile per workgroup of kTopoTile threads (one position per thread):
// each wave records how many of its lanes hold each depth (wcount), which
// then becomes the exclusive prefix over the waves; `base` holds this tile's
// row of the per-depth counts of the earlier tiles (read once per
// workgroup). Three barriers per workgroup in all. (Measured on MI355X, fill
// and ends: four 256-position chunks per workgroup with four barriers each,
// 40 and 23 us at 1M, 544 and 184 us at 10M; one tile per workgroup, 32 and
// 20 us at 1M, the 10M tree phase 1.18 -> 1.13 ms.)
constexpr int kTopoBlockWaves = kTopoTile / 64;

struct TileRanks {
    unsigned wcount[kTopoBlockWaves][kMaxLevels];
    unsigned doff[kMaxLevels];
    unsigned base[kMaxLevels];
};

__device__ inline void tile_ranks_init(TileRanks& r, const unsigned* __restrict__ dcount,
                                       const unsigned* __restrict__ base_row, int tiles) {
    for (int k = threadIdx.x; k < kTopoBlockWaves * kMaxLevels; k += kTopoTile) {
        (&r.wcount[0][0])[k] = 0;
    }
    for (int d = threadIdx.x; d < kMaxLevels; d += kTopoTile) {
        r.base[d] = base_row[static_cast<size_t>(d) * tiles + blockIdx.x];
    }
    depth_offsets(dcount, r.doff);
}

// Lanes holding depth d (in [a, b]) counted per wave, for d in [wlo, whi].
__device__ inline void tile_ranks_count(TileRanks& r, int wlo, int whi, int a, int b) {
    const int wv = threadIdx.x >> 6;
    for (int d = wlo; d <= whi; ++d) {
        const unsigned long long m = __ballot(a <= d && d <= b);
        if ((threadIdx.x & 63) == 0) {
            r.wcount[wv][d] = static_cast<unsigned>(__popcll(m));
        }
    }
}

// wcount[w][d] := the holders of depth d in the waves before w.
__device__ inline void tile_ranks_prefix(TileRanks& r) {
    for (int d = threadIdx.x; d < kMaxLevels; d += kTopoTile) {
        unsigned run = 0;
#pragma unroll
        for (int k = 0; k < kTopoBlockWaves; ++k) {
            const unsigned v = r.wcount[k][d];
            r.wcount[k][d] = run;
            run += v;
        }
    }
}

// Rank, among the holders of depth d, of the lanes before this one: earlier
// tiles, earlier waves of this tile, earlier lanes.
__device__ inline unsigned tile_rank_before(const TileRanks& r, int d, unsigned long long mask) {
    return r.base[d] + r.wcount[threadIdx.x >> 6][d] + lanes_below(mask);
}

// The end-side fields of every cell L's record, from its last tree position
// e: rec[L].y = skip (the node index past its subtree) and rec[L].w = e. The
// fill writes the start-side fields (.x, .z) from the cell's first position.
// Each cell's child table starts empty here (-1 in all eight slots); the
// fill writes the occupied ones.
__global__ __launch_bounds__(kTopoTile) void ends_kernel(const int32_t* __restrict__ delta,
                                                        const uint32_t* __restrict__ cnt,
                                                        const uint32_t* __restrict__ scan, int n,
                                                        int tiles,
                                                        const unsigned* __restrict__ dcount,
                                                        const unsigned* __restrict__ end_base,
                                                        int4* __restrict__ rec,
                                                        GroupNode* __restrict__ gnode) {
    __shared__ TileRanks r;
    tile_ranks_init(r, dcount, end_base, tiles);
    const int e = blockIdx.x * kTopoTile + threadIdx.x;
    const bool valid = e < n && e > 0;
    const int a = valid ? delta[e] + 1 : 0;
    const int b = valid ? delta[e - 1] : -1;
    // node index past the subtree of every cell ending at e
    const int skip = valid ? static_cast<int>(scan[e] + cnt[e]) + e + 1 : 0;
    const bool has = a <= b;
    const int wlo = wave_min_int(has ? a : 0x7FFFFFFF);
    const int whi = wave_max_int(has ? b : -1);
    __syncthreads();   // wcount cleared
    tile_ranks_count(r, wlo, whi, has ? a : 1, has ? b : 0);
    __syncthreads();
    tile_ranks_prefix(r);
    __syncthreads();
    for (int d = wlo; d <= whi; ++d) {
        const bool in = has && a <= d && d <= b;
        const unsigned long long m = __ballot(in);
        if (in) {
            const unsigned L = r.doff[d] + tile_rank_before(r, d, m);
            int* rl = reinterpret_cast<int*>(rec + L);
            rl[1] = skip;
            rl[3] = e;
            int4* kids = reinterpret_cast<int4*>(gnode + L) + 2;
            kids[0] = make_int4(-1, -1, -1, -1);
            kids[1] = make_int4(-1, -1, -1, -1);
        }
    }
}

// Static fields of the internal cells (by level index), the FP64 walk records
// of the leaves (FP64 builds), and every node's slot in its parent's child table
// gnode[parent L].kids[octant] (internal child: its level index; leaf child:
// -(tree position + 2)). A position's cells are visited shallow to deep, so each one's parent
// is the cell visited just before it: the depth-delta[t-1] cell that holds
// t-1 and t (already started) for the shallowest, the previous one otherwise.
__global__ __launch_bounds__(kTopoTile) void fill_kernel(
    const uint64_t* __restrict__ w, const uint64_t* __restrict__ ext,
    const int32_t* __restrict__ delta, const uint32_t* __restrict__ cnt,
    const uint32_t* __restrict__ scan, int n, int tiles, const double4* __restrict__ sp4,
    const double* __restrict__ root, InternalSoA in, NodeF64* __restrict__ n64,
    const unsigned* __restrict__ dcount, const unsigned* __restrict__ start_base,
    GroupNode* __restrict__ gnode) {
    __shared__ TileRanks r;
    tile_ranks_init(r, dcount, start_base, tiles);
    const double root_h = root[3];
    const int t = blockIdx.x * kTopoTile + threadIdx.x;
    const bool valid = t < n;
    const int dprev = valid && t > 0 ? delta[t - 1] : -1;
    const int hi = valid ? delta[t] : -1;
    const int lo = dprev + 1;
    const bool has = valid && lo <= hi;
    const int wlo = wave_min_int(valid ? (dprev > 0 ? dprev : 0) : 0x7FFFFFFF);
    const int whi = wave_max_int(valid ? (dprev > hi ? dprev : hi) : -1);
    __syncthreads();   // wcount cleared
    tile_ranks_count(r, wlo, whi, has ? lo : 1, has ? hi : 0);
    __syncthreads();
    tile_ranks_prefix(r);
    __syncthreads();
    int parent = -1;
    const uint32_t sc = valid ? scan[t] : 0u;
    for (int d = wlo; d <= whi; ++d) {
        const bool starts = has && lo <= d && d <= hi;
        const unsigned long long m = __ballot(starts);
        if (valid && (starts || d == dprev)) {
            // a start of its own, or the latest depth-dprev start before t
            const int L = static_cast<int>(r.doff[d] + tile_rank_before(r, d, m)) -
                          (starts ? 0 : 1);
            if (starts) {
                // the start-side fields; ends_kernel wrote .y and .w
                int* rl = reinterpret_cast<int*>(in.rec + L);
                rl[0] = static_cast<int>(sc) + t + (d - lo);
                rl[2] = t;
                in.depth[L] = d;
                if (parent >= 0) {
                    gnode[parent].kids[digit_at(w, ext, t, d - 1)] = L;
                }
            }
            parent = L;
        }
    }
    if (valid) {
        const int depth = 1 + (dprev > hi ? dprev : hi);
        if (depth > 0 && parent >= 0) {
            gnode[parent].kids[digit_at(w, ext, t, depth - 1)] = -(t + 2);
        }
        if (n64 != nullptr) {   // FP64 records only for the exact FP64 walk
            const int li = static_cast<int>(sc + cnt[t]) + t;
            const double4 q = sp4[t];
            NodeF64 a;
            a.cx = q.x;
            a.cy = q.y;
            a.cz = q.z;
            a.m = q.w;
            a.h = ldexp(root_h, -depth);
            a.skip = li + 1;
            a.leaf = t;
            n64[li] = a;
        }
    }
}

// The FP32 pre-order record of a leaf (tree position t, node index li) and of
// an internal cell (level index L, its FP64 centre of mass and mass, half
// size h, pre-order skip). Only the per-query walks read these records, so
// node_records_kernel writes them on demand after the build.
__device__ inline NodeF32 leaf_record(const float4& q, int li, int t) {
    NodeF32 b;
    b.cx = q.x;
    b.cy = q.y;
    b.cz = q.z;
    b.m = q.w;
    b.size2 = 0.f;
    b.skip = li + 1;
    b.leaf = t;
    b.inv_size2 = 0.f;
    return b;
}

__device__ inline NodeF32 cell_record(const double4& c, double h, int skip, int L) {
    NodeF32 b;
    b.cx = static_cast<float>(c.x);
    b.cy = static_cast<float>(c.y);
    b.cz = static_cast<float>(c.z);
    b.m = static_cast<float>(c.w);
    b.size2 = static_cast<float>(4.0 * h * h);
    b.skip = skip;
    b.leaf = -(L + 2);   // internal: -(level index + 2)
    b.inv_size2 = 1.f / b.size2;
    return b;
}

// One node of the centre-of-mass pass (barnes_hut.c:214-253): children
// summed in octant order 0..7; a zero-mass cell keeps the position of the
// particle that created it (lowest Morton rank in the cell; for the root the
// caller's particle 0, barnes_hut.c:504-512). Children are one level deeper
// and were finished by the previous launch (or barrier). A child is one
// record: the tree-order particle (In: FP32 inputs as float4, widened) or the
// child's FP64 centre of mass.
template <bool QUAD, typename In>
__device__ inline void com_node(int L, int depth, GroupNode* __restrict__ gnode,
                                InternalSoA in, const uint32_t* __restrict__ trank,
                                const In* __restrict__ sp4, const In* __restrict__ b4,
                                double root_h, NodeF64* __restrict__ n64,
                                double* __restrict__ im2, float4* __restrict__ gquad) {
    const int4 r = in.rec[L];
    const int p = r.x;
    const int4 k0 = reinterpret_cast<const int4*>(gnode + L)[2];
    const int4 k1 = reinterpret_cast<const int4*>(gnode + L)[3];
    const int kid[8] = {k0.x, k0.y, k0.z, k0.w, k1.x, k1.y, k1.z, k1.w};
    double mass = 0.0, x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        const int c = kid[o];
        if (c == -1) {
            continue;
        }
        const double4 q = c < -1 ? widen(sp4[-c - 2]) : in.com[c];
        mass += q.w;
        x += q.w * q.x;
        y += q.w * q.y;
        z += q.w * q.z;
    }
    double ox, oy, oz;
    if (mass != 0.0) {
        ox = x / mass;
        oy = y / mass;
        oz = z / mass;
    } else if (p == 0) {
        const double4 p0 = widen(b4[0]);
        ox = p0.x;
        oy = p0.y;
        oz = p0.z;
    } else {
        // the cell's particles are tree positions r.z .. r.w
        int minpos = r.z;
        uint32_t minrank = trank[r.z];
        for (int t = r.z + 1; t <= r.w; ++t) {
            const uint32_t k = trank[t];
            if (k < minrank) {
                minrank = k;
                minpos = t;
            }
        }
        const double4 q = widen(sp4[minpos]);
        ox = q.x;
        oy = q.y;
        oz = q.z;
    }
    in.com[L] = make_double4(ox, oy, oz, mass);
    const double h = ldexp(root_h, -depth);
    if (n64 != nullptr) {
        NodeF64 a;
        a.cx = ox;
        a.cy = oy;
        a.cz = oz;
        a.m = mass;
        a.h = h;
        a.skip = r.y;
        a.leaf = -1;
        n64[p] = a;
    }
    const NodeF32 b = cell_record(make_double4(ox, oy, oz, mass), h, r.y, L);
    if (QUAD) {
        // second moment about the centre of mass from the children's
        // (parallel axes), FP64; the walk's traceless FP32 quadrupole
        double xx = 0.0, yy = 0.0, zz = 0.0, xy = 0.0, xz = 0.0, yz = 0.0;
        if (mass != 0.0) {
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                const int c = kid[o];
                if (c == -1) {
                    continue;
                }
                const double4 q = c < -1 ? widen(sp4[-c - 2]) : in.com[c];
                const double dx = q.x - ox, dy = q.y - oy, dz = q.z - oz;
                xx += q.w * dx * dx;
                yy += q.w * dy * dy;
                zz += q.w * dz * dz;
                xy += q.w * dx * dy;
                xz += q.w * dx * dz;
                yz += q.w * dy * dz;
                if (c >= 0) {
                    const double* cm = im2 + 6 * static_cast<size_t>(c);
                    xx += cm[0];
                    yy += cm[1];
                    zz += cm[2];
                    xy += cm[3];
                    xz += cm[4];
                    yz += cm[5];
                }
            }
        }
        double* om = im2 + 6 * static_cast<size_t>(L);
        om[0] = xx;
        om[1] = yy;
        om[2] = zz;
        om[3] = xy;
        om[4] = xz;
        om[5] = yz;
        const double tr = xx + yy + zz;
        gquad[2 * L] = make_float4(static_cast<float>(3.0 * xx - tr), static_cast<float>(3.0 * yy - tr),
                                   static_cast<float>(3.0 * zz - tr), static_cast<float>(3.0 * xy));
        gquad[2 * L + 1] = make_float4(static_cast<float>(3.0 * xz), static_cast<float>(3.0 * yz),
                                       b.size2, b.inv_size2);
    }
    float4* g = reinterpret_cast<float4*>(gnode + L);
    g[0] = make_float4(b.cx, b.cy, b.cz, b.m);
    g[1] = make_float4(b.size2, b.inv_size2, 0.f, 0.f);
}

// One depth level per launch (large levels): cells first .. first + count - 1.
template <bool QUAD, typename In>
__global__ __launch_bounds__(kThreads) void com_level_kernel(
    int first, int count, int depth, GroupNode* __restrict__ gnode,
    InternalSoA in, const uint32_t* __restrict__ trank, const In* __restrict__ sp4,
    const In* __restrict__ b4, const double* __restrict__ root, NodeF64* __restrict__ n64,
    double* __restrict__ im2, float4* __restrict__ gquad) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k < count) {
        com_node<QUAD, In>(first + k, depth, gnode, in, trank, sp4, b4, root[3], n64, im2, gquad);
    }
}

// Consecutive small depth levels (deepest first) in one workgroup; the
// barrier orders each level's results before the next level reads them.
template <bool QUAD, typename In>
__global__ __launch_bounds__(1024) void com_small_levels_kernel(
    const unsigned* __restrict__ dcount, int d_hi, int d_lo, GroupNode* __restrict__ gnode,
    InternalSoA in, const uint32_t* __restrict__ trank, const In* __restrict__ sp4,
    const In* __restrict__ b4, const double* __restrict__ root, NodeF64* __restrict__ n64,
    double* __restrict__ im2, float4* __restrict__ gquad) {
    __shared__ unsigned doff[kMaxLevels];
    depth_offsets(dcount, doff);
    __syncthreads();
    const double root_h = root[3];
    for (int d = d_hi; d >= d_lo; --d) {
        const int count = static_cast<int>(dcount[d]);
        const int first = static_cast<int>(doff[d]);
        for (int k = threadIdx.x; k < count; k += blockDim.x) {
            com_node<QUAD, In>(first + k, d, gnode, in, trank, sp4, b4, root_h, n64, im2, gquad);
        }
        __syncthreads();
    }
}

// The FP32 pre-order records of every node, for the per-query walks, from a
// finished build: thread k < n writes the leaf at tree position k (node index
// scan[k] + cnt[k] + k), thread n + L the internal cell of level index L (node
// index rec[L].x). The values are the build's own FP64 values rounded to FP32
// (the leaves' are the tree-order s4).
__global__ __launch_bounds__(kThreads) void node_records_kernel(
    int n, int internal, const uint32_t* __restrict__ scan, const uint32_t* __restrict__ cnt,
    const float4* __restrict__ s4, InternalSoA in, const double* __restrict__ root,
    NodeF32* __restrict__ n32) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k < n) {
        const int li = static_cast<int>(scan[k] + cnt[k]) + k;
        n32[li] = leaf_record(s4[k], li, k);
    } else if (k - n < internal) {
        const int L = k - n;
        const int4 r = in.rec[L];
        n32[r.x] = cell_record(in.com[L], ldexp(root[3], -in.depth[L]), r.y, L);
    }
}


// ---------------------------------------------------------------- walk
// FP64 variant: the reference arithmetic (barnes_hut.c:574-613), bit-identical.
__global__ __launch_bounds__(kThreads) void walk_f64_kernel(
    const NodeF64* __restrict__ n64, int nodes, const double4* __restrict__ sp4,
    const uint32_t* __restrict__ tperm, int n, double G, double eps2, double theta2,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    unsigned long long* __restrict__ counters) {
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
    const double4 qp = sp4[tc];
    const double qx = qp.x, qy = qp.y, qz = qp.z;
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
        atomicAdd(&block_counts[0], static_cast<unsigned long long>(approx));
        atomicAdd(&block_counts[1], static_cast<unsigned long long>(exact));
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        atomicAdd(&counters[threadIdx.x], block_counts[threadIdx.x]);
    }
}

// MI355X dispatches workgroup b to XCD (b mod 8), and each XCD has its own L2.
// Remap so that XCD x walks one contiguous eighth of the queries (Morton
// order): its L2 then holds that eighth's part of the tree. `grid` is a
// multiple of kXcds; surplus blocks get indices >= n and only walk the root.
constexpr int kXcds = 8;
#ifndef NBX_WALK_QUERIES
#define NBX_WALK_QUERIES 2
#endif
// Queries per lane of the FP32 walk.
constexpr int kWalkQueries = NBX_WALK_QUERIES;
__device__ inline int xcd_block(int b, int grid) {
    return (b % kXcds) * (grid / kXcds) + b / kXcds;
}

// Per-query accumulation of the FP32 walks: the softened interaction terms
// and, with FEATURES, the walk statistics of the learned correction
// (quadrupole force of the accepted cells, sum and minimum of (distance /
// cell edge)^2). G is applied once, when the result is stored.
template <bool FEATURES>
struct WalkAcc {
    float ax = 0.f, ay = 0.f, az = 0.f;
    unsigned approx = 0, exact = 0;
    float qax = 0.f, qay = 0.f, qaz = 0.f;
    float q2_sum = 0.f, q2_min = 3.0e38f;

    __device__ void leaf(float dx, float dy, float dz, float m, float eps2) {
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float s = (m * inv) * (inv * inv);
        ax = fmaf(s, dx, ax);
        ay = fmaf(s, dy, ay);
        az = fmaf(s, dz, az);
        ++exact;
    }

    // Group-walk terms: the same arithmetic as leaf(), counted per flush by
    // the caller. A staged leaf can be the query itself (dx = dy = dz = 0
    // exactly, both positions being the same float4): with eps2 > 0 its term
    // is s * 0 = +0 and needs no test; with eps2 = 0 leaf_or_self_term gives
    // it no contribution (and no division by a zero distance), branch-free.
    __device__ void leaf_term(float dx, float dy, float dz, float m, float eps2) {
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float s = (m * inv) * (inv * inv);
        ax = fmaf(s, dx, ax);
        ay = fmaf(s, dy, ay);
        az = fmaf(s, dz, az);
    }

    __device__ void leaf_or_self_term(float dx, float dy, float dz, float m, bool self,
                                      float eps2) {
        const float r2 = self ? 1.f : fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float s = ((self ? 0.f : m) * inv) * (inv * inv);
        ax = fmaf(s, dx, ax);
        ay = fmaf(s, dy, ay);
        az = fmaf(s, dz, az);
    }

    // accepted cell, monopole term only
    __device__ void cell_softened(float dx, float dy, float dz, float m, float eps2) {
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float s = (m * inv) * (inv * inv);
        ax = fmaf(s, dx, ax);
        ay = fmaf(s, dy, ay);
        az = fmaf(s, dz, az);
        ++approx;
    }

    // accepted cell at softened squared distance r2 (the force) and
    // unsoftened d2 (the features); with FEATURES its quadrupole record
    // q0 = (Qxx, Qyy, Qzz, Qxy), q1 = (Qxz, Qyz, size^2, 1 / size^2),
    // Q = sum m (3 x x^T - |x|^2 I) about the centre of mass
    __device__ void cell(float dx, float dy, float dz, float r2, float d2, float m, float4 q0,
                         float4 q1) {
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float inv2 = inv * inv;
        const float s = (m * inv) * inv2;
        ax = fmaf(s, dx, ax);
        ay = fmaf(s, dy, ay);
        az = fmaf(s, dz, az);
        ++approx;
        if (FEATURES) {
            // quadrupole term with d = (cell - query):
            // G (5/2 (d.Q.d) d / r^7 - Q.d / r^5), softened r
            const float qdx = fmaf(q0.x, dx, fmaf(q0.w, dy, q1.x * dz));
            const float qdy = fmaf(q0.w, dx, fmaf(q0.y, dy, q1.y * dz));
            const float qdz = fmaf(q1.x, dx, fmaf(q1.y, dy, q0.z * dz));
            const float dqd = fmaf(dx, qdx, fmaf(dy, qdy, dz * qdz));
            const float inv5 = (inv2 * inv2) * inv;
            const float c = (2.5f * dqd) * inv2;
            qax = fmaf(fmaf(c, dx, -qdx), inv5, qax);
            qay = fmaf(fmaf(c, dy, -qdy), inv5, qay);
            qaz = fmaf(fmaf(c, dz, -qdz), inv5, qaz);
            const float q2 = d2 * q1.w;   // (distance / cell edge)^2
            q2_sum += q2;
            // (d / edge)^2 >= 0, where unsigned order is float order: one
            // v_min_u32 and no NaN canonicalisation
            q2_min = __uint_as_float(min(__float_as_uint(q2_min), __float_as_uint(q2)));
        }
    }

    // FEATURES: the kWalkRecord-float record of tree position t (G not
    // applied; its consumer scatters G * sum). Otherwise the accelerations of
    // particle tperm[t], in original order.
    __device__ void store(int t, const uint32_t* __restrict__ tperm, float G,
                          double* __restrict__ oax, double* __restrict__ oay,
                          double* __restrict__ oaz, float* __restrict__ features) const {
        if (!FEATURES) {
            const uint32_t o = tperm[t];
            oax[o] = static_cast<double>(G * ax);
            oay[o] = static_cast<double>(G * ay);
            oaz[o] = static_cast<double>(G * az);
        } else {
            // total sum xyz, accepted cells; quadrupole sum xyz, exact leaf
            // terms; sum and minimum of (d / edge)^2; zeros
            float4* f = reinterpret_cast<float4*>(features + static_cast<size_t>(t) * kWalkRecord);
            f[0] = make_float4(ax, ay, az, static_cast<float>(approx));
            f[1] = make_float4(qax, qay, qaz, static_cast<float>(exact));
            f[2] = make_float4(q2_sum, approx > 0 ? q2_min : 0.f, 0.f, 0.f);
            f[3] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
};

// One wave-step of the per-query walk at node `cur` (index i) for the query
// at tree position t, position (qx, qy, qz): returns true when this query
// descends into the node.
template <bool FEATURES>
__device__ inline bool walk_visit(const NodeF32& cur, int i, int t, float qx, float qy, float qz,
                                  float eps2, float theta2, int& skip_until,
                                  WalkAcc<FEATURES>& acc, float4 q0, float4 q1) {
    if (i < skip_until) {
        return false;
    }
    const float dx = cur.cx - qx;
    const float dy = cur.cy - qy;
    const float dz = cur.cz - qz;
    if (cur.leaf >= 0) {
        if (cur.leaf != t) {
            acc.leaf(dx, dy, dz, cur.m, eps2);
        }
        return false;
    }
    const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    if (d2 > 0.f && cur.size2 < theta2 * d2) {
        acc.cell(dx, dy, dz, d2 + eps2, d2, cur.m, q0, q1);
        skip_until = cur.skip;
        return false;
    }
    return true;
}

// FP32 walk with per-query reference interaction lists: one node per wave
// step, node fields in SGPRs (scalar loads of a wave-uniform index); lanes with
// nothing to do at a node are masked off. Each lane owns Q queries (tree
// positions base + 64k + lane, k < Q), so a wave walks the union of 64 Q
// consecutive queries' lists and every node record fetched (the walk's
// latency chain) feeds Q times the arithmetic; each query keeps its own
// skip_until and its reference visit order. (Measured alternatives, slower on
// MI355X: branch-free steps, vector-loaded nodes, prefetching both
// successors, four queries per lane.)
// The quadrupole record of pre-order node `cur` for the per-query walks: an
// internal node carries its level index as leaf = -(L + 2); a leaf has none.
__device__ inline void node_quad(const NodeF32& cur, const float4* __restrict__ gquad,
                                 float4& q0, float4& q1) {
    const int L = -cur.leaf - 2;
    if (L >= 0) {
        q0 = gquad[2 * L];
        q1 = gquad[2 * L + 1];
    } else {
        q0 = make_float4(0.f, 0.f, 0.f, 0.f);
        q1 = q0;
    }
}

// Queries: tree positions 0 .. nq - 1, or with `list` the tree positions
// list[0 .. nq - 1] (the queries a group walk deferred).
template <bool FEATURES, int Q>
__global__ __launch_bounds__(kThreads) void walk_f32_kernel(
    const NodeF32* __restrict__ n32, int nodes, const float4* __restrict__ s4,
    const uint32_t* __restrict__ tperm, const int* __restrict__ list, int nq, float G,
    float eps2, float theta2, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, unsigned long long* __restrict__ counters,
    float* __restrict__ features, const float4* __restrict__ gquad) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = xcd_block(blockIdx.x, gridDim.x) * (kThreads / kWave) + threadIdx.x / kWave;
    const int base = wave * (kWave * Q);
    int t[Q], skip_until[Q];
    bool valid[Q];
    float qx[Q], qy[Q], qz[Q];
    WalkAcc<FEATURES> acc[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int qi = base + k * kWave + lane;
        valid[k] = qi < nq;
        t[k] = valid[k] ? (list != nullptr ? list[qi] : qi) : -1;
        const float4 q = s4[valid[k] ? t[k] : 0];
        qx[k] = q.x;
        qy[k] = q.y;
        qz[k] = q.z;
        skip_until[k] = valid[k] ? 0 : 0x7FFFFFFF;
    }
    int i = 0;
    while (i < nodes) {
        const int iu = __builtin_amdgcn_readfirstlane(i);
        const NodeF32 cur = n32[iu];
        // with FEATURES the node's quadrupole record
        float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0;
        if (FEATURES) {
            node_quad(cur, gquad, q0, q1);
        }
        bool descend = false;
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            descend |= walk_visit<FEATURES>(cur, i, t[k], qx[k], qy[k], qz[k], eps2, theta2,
                                            skip_until[k], acc[k], q0, q1);
        }
        i = cur.leaf >= 0 || __builtin_amdgcn_ballot_w64(descend) != 0 ? i + 1 : cur.skip;
    }
    unsigned long long wa = 0, we = 0;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        if (valid[k]) {
            wa += acc[k].approx;
            we += acc[k].exact;
            acc[k].store(t[k], tperm, G, ax, ay, az, features);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        wa += __shfl_xor(wa, off);
        we += __shfl_xor(we, off);
    }
    if (lane == 0) {
        atomicAdd(&counters[0], wa);
        atomicAdd(&counters[1], we);
    }
}

// Group walk (FP32). A wave takes 64 Q consecutive tree positions (Q per
// lane) as one group and walks the tree once for the group's bounding box: a
// cell is accepted when the reference's test (d^2 > 0 and 4h^2 < theta^2 d^2)
// holds for the box distance d, which implies it holds for every member, so
// each member's interaction set refines its reference set. The walk runs on
// the level-indexed cells: a per-wave stack of cells in LDS is popped up to 64
// at a time (one cell per lane, independent loads), accepted cells are
// staged, and an opened cell pushes its internal children and stages its
// leaf children. The dependent chain is one record load per step, about the
// tree depth in steps, instead of one per node visited in pre-order. Staged
// cells and leaves are evaluated by all 64 lanes without divergence (leaf
// positions are gathered into LDS just before). A group that is not compact
// (64 consecutive positions of the Morton curve can straddle one of its
// jumps) would get a box with most of the tree in its near field: it is split
// into compact parts. The members of a part the split cannot reach within its
// budget, or whose stack would overflow, are deferred to the per-query walk
// with reference lists, which runs after this kernel.
constexpr int kGroupList = 128;      // staged cells per wave
// 96 staged leaves: less LDS per wave, so more waves per CU (measured on
// MI355X against 128: 3% faster walk at 1M particles, 6% at 10M)
constexpr int kGroupLeaves = 96;     // staged leaves per wave
constexpr int kGroupStack = 512;     // pending cells per wave
constexpr int kRangeStack = 32;      // pending sub-ranges per wave
constexpr int kMaxSubWalks = 16;     // shared-list walks per wave
constexpr int kGroupCell = 2048;     // compact: the group's smallest enclosing cell holds <= kGroupCell

// Whether the run of tree positions that share the first c levels with a and
// b (a < b, which share exactly c; tree order is key order, so the run is an
// interval) holds at most kGroupCell positions. Each lane probes one position
// past each end at a stride of kGroupCell / 64 (the in-run probes are a prefix
// of the lanes), then half the lanes probe each end's last stride position by
// position: two rounds of independent loads instead of two galloping binary
// searches of up to ~40 dependent loads.
constexpr int kRunStride = kGroupCell / kWave;
static_assert(kRunStride == kWave / 2, "the fine round probes one stride per half wave");

__device__ inline bool run_compact(const uint64_t* __restrict__ w, const uint64_t* __restrict__ ext,
                                   int n, int a, int b, int c, int lane) {
    const int r = b + kRunStride * (lane + 1);
    const int l = a - kRunStride * (lane + 1);
    const bool in_r = r < n && common_levels(w, ext, b, r) >= c;
    const bool in_l = l >= 0 && common_levels(w, ext, l, a) >= c;
    const int kr = __popcll(__builtin_amdgcn_ballot_w64(in_r));
    const int kl = __popcll(__builtin_amdgcn_ballot_w64(in_l));
    if (kr == kWave || kl == kWave) {
        return false;   // the run reaches kGroupCell positions past one end
    }
    // lanes 0..30 probe the 31 positions after b + stride * kr, lanes 32..62
    // the 31 before a - stride * kl
    const int f = lane & (kRunStride - 1);
    const bool right = lane < kRunStride;
    const int p = right ? b + kRunStride * kr + f + 1 : a - kRunStride * kl - f - 1;
    bool in = false;
    if (f < kRunStride - 1) {
        in = right ? p < n && common_levels(w, ext, b, p) >= c
                   : p >= 0 && common_levels(w, ext, p, a) >= c;
    }
    const uint64_t fine = __builtin_amdgcn_ballot_w64(in);
    const int last = b + kRunStride * kr + __popcll(fine & 0xFFFFFFFFull);
    const int first = a - kRunStride * kl - __popcll(fine >> 32);
    return last - first + 1 <= kGroupCell;
}
#ifndef NBX_GROUP_THREADS
#define NBX_GROUP_THREADS 64
#endif
// Groups are dealt to the XCDs in chunks of kGroupChunk consecutive groups
// (chunk c to XCD c mod 8, each XCD's chunks in tree order): every XCD gets a
// share of the dense and the sparse regions, and a chunk's waves still share
// cell records in their XCD's L2. (Measured on MI355X, two-galaxy: with each
// XCD on one contiguous eighth of the groups, the XCDs finished between 630
// and 846 us at 1M and between 12.8 and 16.2 ms at 10M, theta = 0.5; walk
// 0.78 / 16.4 ms. Chunks of 8, 32, 128 groups: 0.80 / 14.4, 0.80 / 14.2,
// 0.78 / 14.2 ms.)
constexpr int kGroupChunk = 128;

__device__ inline int group_of_block(int b) {
    const int slot = b / kXcds;   // workgroup b runs on XCD b mod 8
    return ((slot / kGroupChunk) * kXcds + b % kXcds) * kGroupChunk + slot % kGroupChunk;
}

// One wave per workgroup: a wave's LDS is released as soon as it finishes
// (waves of one region take very different times; measured on MI355X, 1M
// particles: 1.45 ms with one wave per workgroup, 1.66 ms with four).
constexpr int kGroupThreads = NBX_GROUP_THREADS;
// One query per lane in the group walk (measured on MI355X: two per lane, a
// group of 128, is 7% slower at 10M particles and 16% slower at 1M).

// Exclusive prefix max over the wave of non-negative ints, with DPP row
// shifts and row broadcasts (no LDS round trips); `total` = max over the wave.
__device__ inline int dpp_max(int v, int shifted) {
    return v > shifted ? v : shifted;
}

__device__ inline int wave_prefix_max_exclusive(int v, int lane, int& total) {
    // inclusive scan within each 16-lane row (out-of-row sources read 0)
    v = dpp_max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));   // row_shr:1
    v = dpp_max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));   // row_shr:2
    v = dpp_max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));   // row_shr:4
    v = dpp_max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));   // row_shr:8
    // across rows: lane 15 of each row into the next row, lane 31 into rows 2, 3
    v = dpp_max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));   // row_bcast:15
    v = dpp_max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));   // row_bcast:31
    total = __builtin_amdgcn_readlane(v, kWave - 1);
    const int e = __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false);      // wave_shr:1
    return lane == 0 ? 0 : e;
}

__device__ inline float wave_min(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v = fminf(v, __shfl_xor(v, off));
    }
    return v;
}

__device__ inline float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v = fmaxf(v, __shfl_xor(v, off));
    }
    return v;
}

// Orders one wave's LDS writes before its later LDS reads.
__device__ inline void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool FEATURES>
__global__ __launch_bounds__(kGroupThreads) void group_walk_kernel(
    int* __restrict__ deferred, const GroupNode* __restrict__ gnode,
    int internal, const float4* __restrict__ s4,
    const uint32_t* __restrict__ tperm, const uint64_t* __restrict__ tword,
    const uint64_t* __restrict__ ext, int n, float G, float eps2, float theta2,
    double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az, unsigned long long* __restrict__ counters,
    float* __restrict__ features, const float4* __restrict__ gquad,
    int max_walks, bool member_test) {
    constexpr int kWaves = kGroupThreads / kWave;
    // per wave: pending cells; staged cells (COM, mass; with FEATURES the
    // quadrupole record); staged leaves (tree position, then position and
    // mass); whether each lane's own leaf was staged; pending sub-ranges
    __shared__ int stack_all[kWaves][kGroupStack];
    // FEATURES stages fewer cells per flush: their quadrupole records take LDS.
    // (Measured on MI355X, learned walk at 10M: 96 or 128 staged cells, 8.9 or
    // 10.4 KB of LDS, are 4% and 17% slower; six waves per SIMD through 80
    // VGPRs and 64 staged leaves is 5% slower. A list holds up to one step's
    // accepted cells, one per lane, so it cannot be shorter than a wave.)
    // A staged cell is kCellWords float4 in a row (COM and mass, then with
    // FEATURES the two quadrupole float4), so one LDS pointer walks the list.
    constexpr int kList = FEATURES ? kGroupList / 2 : kGroupList;
    static_assert(kList >= kWave, "a step stages up to one accepted cell per lane");
    static_assert(kGroupLeaves >= kWave, "the member test stages a step's candidates in lleaf");
    // The plain walk evaluates its staged lists four entries per loop trip
    // (LDS loads run ahead, one address update per four; same summation
    // order). Measured on MI355X against one per trip: walk 7% faster at 1M
    // particles and 5% at 10M, at 78 VGPRs; eight per trip takes 98 VGPRs.
    // The feature walk stays at one (two per trip: 99 VGPRs, fewer waves).
    constexpr int kEvalUnroll = FEATURES ? 1 : 4;
    constexpr int kCellWords = FEATURES ? 3 : 1;
    __shared__ float4 lcell_all[kWaves][kList * kCellWords];
    __shared__ int ltag_all[kWaves][kGroupLeaves];
    __shared__ float4 lleaf_all[kWaves][kGroupLeaves];
    __shared__ unsigned char selfhit_all[kWaves][kWave];
    __shared__ int2 ranges_all[kWaves][kRangeStack];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    int* stack = stack_all[wid];
    float4* lcell = lcell_all[wid];
    int* ltag = ltag_all[wid];
    float4* lleaf = lleaf_all[wid];
    unsigned char* selfhit = selfhit_all[wid];
    int2* ranges = ranges_all[wid];
    static_assert(kWaves == 1, "one group per workgroup");
    const int wave = group_of_block(static_cast<int>(blockIdx.x));
    const int t0 = wave * kWave;
    if (t0 >= n) {
        return;   // whole wave beyond the particles
    }
    // the group: tree positions t0 + lane
    const int t = t0 + lane;
    const bool valid = t < n;
    const float4 q = s4[valid ? t : n - 1];
    const float qx = q.x, qy = q.y, qz = q.z;
    WalkAcc<FEATURES> acc;
    bool per_query = false;   // this lane's sub-group could not share a list

    // One shared-list walk for the members [a, b] (tree positions) of the
    // group: box acceptance against the members' bounding box. Returns false,
    // with the members' sums cleared, when the pending cells would overflow
    // the stack.
    auto shared_walk = [&](int a, int b) -> bool {
        const bool member = valid && t >= a && t <= b;
        const float bx0 = wave_min(member ? qx : 3.0e38f);
        const float by0 = wave_min(member ? qy : 3.0e38f);
        const float bz0 = wave_min(member ? qz : 3.0e38f);
        const float bx1 = wave_max(member ? qx : -3.0e38f);
        const float by1 = wave_max(member ? qy : -3.0e38f);
        const float bz1 = wave_max(member ? qz : -3.0e38f);
        int ccnt = 0, lcnt = 0;            // staged cells, leaves (wave-uniform)
        unsigned ncell = 0, nleaf = 0;     // evaluated per member (wave-uniform)
        selfhit[lane] = 0;
        // evaluate the staged lists for every member
        auto flush = [&]() {
            wave_lds_fence();
            for (int e = lane; e < lcnt; e += kWave) {
                lleaf[e] = s4[ltag[e]];
            }
            wave_lds_fence();
            if (member) {
#pragma unroll kEvalUnroll
                for (int e = 0; e < ccnt; ++e) {
                    const float4 p = lcell[kCellWords * e];
                    const float dx = p.x - qx;
                    const float dy = p.y - qy;
                    const float dz = p.z - qz;
                    if (FEATURES) {
                        const float4 q0 = lcell[kCellWords * e + 1];
                        const float4 q1 = lcell[kCellWords * e + 2];
                        // softened distance from the unsoftened one (within an
                        // ulp of the plain walk's FMA chain)
                        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                        acc.cell(dx, dy, dz, d2 + eps2, d2, p.w, q0, q1);
                    } else {
                        // softened distance straight from the FMA chain (the
                        // unsoftened one only feeds the features)
                        acc.cell_softened(dx, dy, dz, p.w, eps2);
                    }
                }
                if (eps2 > 0.f) {
#pragma unroll kEvalUnroll
                    for (int e = 0; e < lcnt; ++e) {
                        const float4 p = lleaf[e];
                        acc.leaf_term(p.x - qx, p.y - qy, p.z - qz, p.w, eps2);
                    }
                } else {
                    for (int e = 0; e < lcnt; ++e) {
                        const float4 p = lleaf[e];
                        acc.leaf_or_self_term(p.x - qx, p.y - qy, p.z - qz, p.w, ltag[e] == t,
                                              eps2);
                    }
                }
            }
            ncell += static_cast<unsigned>(ccnt);
            nleaf += static_cast<unsigned>(lcnt);
            ccnt = 0;
            lcnt = 0;
            wave_lds_fence();
        };
        int sp = 0;   // pending cells (wave-uniform); the root is level index 0
        if (internal > 0) {
            if (lane == 0) {
                stack[0] = 0;
            }
            sp = 1;
        }
        bool overflow = false;
        wave_lds_fence();
        while (sp > 0) {
            // (sp goes through readfirstlane so the compiler keeps it, and
            // every test on it, in scalar registers; with the unmasked record
            // loads below, measured on MI355X: walk 0.817 -> 0.794 ms at 1M,
            // 17.07 -> 16.3 ms at 10M and theta = 0.5)
            sp = __builtin_amdgcn_readfirstlane(sp);
            // pop at most 64 cells, and few enough that their children (8
            // each) fit: a wide level then spreads over more steps instead
            // of overflowing
            const int room = (kGroupStack - sp) / 8;
            const int cap = room > 1 ? (room < kWave ? room : kWave) : 1;
            const int take = sp < cap ? sp : cap;
            const bool act = lane < take;
            const int ref = act ? stack[sp - take + lane] : 0;
            sp = __builtin_amdgcn_readfirstlane(sp - take);
            wave_lds_fence();   // pops read before pushes reuse their slots
            // Every lane loads a record (the lanes past `take` the root's) and
            // tests it; `act` masks the outcomes, so idle lanes need no defaults.
            const float4* g = reinterpret_cast<const float4*>(gnode + ref);
            const float4 c0 = g[0];
            const float4 c1 = g[1];
            const int4 k0 = reinterpret_cast<const int4*>(g)[2];
            const int4 k1 = reinterpret_cast<const int4*>(g)[3];
            float4 gq0 = make_float4(0.f, 0.f, 0.f, 0.f), gq1 = gq0;
            if (FEATURES) {
                // the quadrupole record comes with the cell (loads in flight together)
                gq0 = gquad[2 * ref];
                gq1 = gquad[2 * ref + 1];
            }
            const float ddx = fmaxf(0.f, fmaxf(bx0 - c0.x, c0.x - bx1));
            const float ddy = fmaxf(0.f, fmaxf(by0 - c0.y, c0.y - by1));
            const float ddz = fmaxf(0.f, fmaxf(bz0 - c0.z, c0.z - bz1));
            const float d2 = fmaf(ddx, ddx, fmaf(ddy, ddy, ddz * ddz));
            bool accepted = act & (d2 > 0.f) & (c1.x < theta2 * d2);
            if (member_test) {
                // A cell the box test opens is still accepted when every
                // member's own reference test (walk_visit's arithmetic)
                // accepts it: the lists keep refining the members' reference
                // lists. Only cells that pass at the box's far corner from
                // their centre of mass can pass for some member; each such
                // candidate is broadcast and tested by all members at once.
                const float fx = fmaxf(fabsf(c0.x - bx0), fabsf(c0.x - bx1));
                const float fy = fmaxf(fabsf(c0.y - by0), fabsf(c0.y - by1));
                const float fz = fmaxf(fabsf(c0.z - bz0), fabsf(c0.z - bz1));
                const bool cand = act & !accepted &
                                  (c1.x < theta2 * fmaf(fx, fx, fmaf(fy, fy, fz * fz)));
                // The candidates go to LDS in lane order (the staged-leaf
                // buffer: it only holds data inside a flush) and every member
                // reads each one by broadcast; the tests take no branches.
                const uint64_t pending = __builtin_amdgcn_ballot_w64(cand);
                if (pending != 0) {
                    const uint64_t members = __builtin_amdgcn_ballot_w64(member);
                    const int slot = static_cast<int>(lanes_below(pending));
                    if (cand) {
                        lleaf[slot] = make_float4(c0.x, c0.y, c0.z, c1.x);
                    }
                    wave_lds_fence();
                    const int nc = static_cast<int>(__popcll(pending));
                    uint64_t pass = 0;   // bit j: every member accepts candidate j
                    for (int j = 0; j < nc; ++j) {
                        const float4 cj = lleaf[j];
                        const float dx = cj.x - qx;
                        const float dy = cj.y - qy;
                        const float dz = cj.z - qz;
                        const float dj2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                        // (each comparison balloted on its own: its lane
                        // mask is the ballot, with no VALU round trip)
                        const uint64_t ok = __builtin_amdgcn_ballot_w64(dj2 > 0.f) &
                                            __builtin_amdgcn_ballot_w64(cj.w < theta2 * dj2);
                        if ((members & ~ok) == 0) {
                            pass |= 1ull << j;
                        }
                    }
                    wave_lds_fence();   // candidates read before the buffer is reused
                    accepted = accepted | (cand & (((pass >> slot) & 1ull) != 0));
                }
            }
            if (ccnt > kList - kWave) {
                flush();
            }
            const uint64_t cbal = __ballot(accepted);
            if (accepted) {
                const int pos = ccnt + static_cast<int>(lanes_below(cbal));
                lcell[kCellWords * pos] = c0;
                if (FEATURES) {
                    lcell[kCellWords * pos + 1] = gq0;
                    lcell[kCellWords * pos + 2] = gq1;
                }
            }
            ccnt += static_cast<int>(__popcll(cbal));
            // opened cells: children in octant order, one octant per round
            // (at most 64 new leaves and 64 new cells per round)
            const bool open = act && !accepted;
            const int kid[8] = {k0.x, k0.y, k0.z, k0.w, k1.x, k1.y, k1.z, k1.w};
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                const int ch = open ? kid[o] : -1;
                const uint64_t lb = __ballot(ch < -1);
                const uint64_t ib = __ballot(ch >= 0);
                if (lb != 0) {
                    if (lcnt > kGroupLeaves - kWave) {
                        flush();
                    }
                    if (ch < -1) {
                        const int tag = -ch - 2;
                        ltag[lcnt + static_cast<int>(lanes_below(lb))] = tag;
                        const unsigned rel = static_cast<unsigned>(tag - t0);
                        if (rel < static_cast<unsigned>(kWave)) {
                            selfhit[rel] = 1;
                        }
                    }
                    lcnt += static_cast<int>(__popcll(lb));
                }
                const int np = static_cast<int>(__popcll(ib));
                if (sp + np > kGroupStack) {
                    overflow = true;
                } else if (np != 0) {
                    if (ch >= 0) {
                        stack[sp + static_cast<int>(lanes_below(ib))] = ch;
                    }
                    sp = __builtin_amdgcn_readfirstlane(sp + np);
                }
            }
            if (overflow) {
                break;
            }
            wave_lds_fence();   // pushes visible before the next pops
        }
        if (!overflow && ccnt + lcnt > 0) {
            flush();
        }
        wave_lds_fence();
        if (member) {
            if (overflow) {
                acc = WalkAcc<FEATURES>();
            } else {
                // each member evaluated every staged entry but its own leaf
                acc.approx = ncell;
                acc.exact = nleaf - selfhit[lane];
            }
        }
        wave_lds_fence();   // selfhit read before the next walk clears it
        return !overflow;
    };

    // Sub-ranges of the group, split at the Morton curve's jumps until each
    // is compact: its smallest enclosing cell holds <= kGroupCell particles
    // (a range straddling a jump, an outlier, a dense core and its outskirts
    // would otherwise get a box with most of the tree in its near field). A
    // range that is not compact is split at the octant changes of its common
    // cell's children (tree order is key order, so each octant is a run); the
    // members of a range that cannot be split within the budget of
    // `max_walks` shared walks per wave are deferred to the per-query walk.
    const int last = t0 + kWave - 1 < n ? t0 + kWave - 1 : n - 1;
    const int max_level = ext != nullptr ? kMaxLevels : 21;
    int rsp = 1, walks = 0;
    if (lane == 0) {
        ranges[0] = make_int2(t0, last);
    }
    wave_lds_fence();
    while (rsp > 0) {
        --rsp;
        const int2 r = ranges[rsp];
        const int a = __builtin_amdgcn_readfirstlane(r.x);
        const int b = __builtin_amdgcn_readfirstlane(r.y);
        wave_lds_fence();   // the pop is read before a push reuses the slot
        const int c =
            __builtin_amdgcn_readfirstlane(a < b ? common_levels(tword, ext, a, b) : max_level);
        const bool shared = c >= max_level || run_compact(tword, ext, n, a, b, c, lane);
        uint64_t cut = 0;
        if (!shared && rsp + 8 <= kRangeStack && walks + rsp + 2 <= max_walks) {
            // octant of each member's depth c + 1 cell; runs of equal octants
            // (a and b differ in that octant, so there are at least two)
            const int dig = digit_at(tword, ext, valid ? t : a, c);
            const int prev = __shfl_up(dig, 1);
            cut = __ballot(t > a && t <= b && dig != prev);
        }
        if (cut != 0) {
            int s = a;
            while (cut != 0) {
                const int j = __builtin_ctzll(cut);
                cut &= cut - 1;
                if (lane == 0) {
                    ranges[rsp] = make_int2(s, t0 + j - 1);
                }
                ++rsp;
                s = t0 + j;
            }
            if (lane == 0) {
                ranges[rsp] = make_int2(s, b);
            }
            ++rsp;
            wave_lds_fence();
            continue;
        }
        if (shared) {
            ++walks;
            if (!shared_walk(a, b)) {
                per_query = per_query || (valid && t >= a && t <= b);
            }
        } else {
            per_query = per_query || (valid && t >= a && t <= b);
        }
    }
    // The lanes whose range could not share a list go to the deferred list
    // (counters[2] entries): the per-query reference-list walk takes them after
    // this kernel, and writes their results. Their sums here are empty (never
    // walked, or cleared on overflow).
    const uint64_t deferring = __builtin_amdgcn_ballot_w64(per_query);
    if (deferring != 0) {
        unsigned long long slot = 0;
        if (lane == 0) {
            slot = atomicAdd(&counters[2], static_cast<unsigned long long>(__popcll(deferring)));
        }
        slot = __shfl(slot, 0);
        const unsigned long long k = slot + lanes_below(deferring);
        if (per_query && k < static_cast<unsigned long long>(n)) {
            deferred[k] = t;
        }
    }
    unsigned long long wa = 0, we = 0;
    if (valid && !per_query) {
        wa = acc.approx;
        we = acc.exact;
        acc.store(t, tperm, G, ax, ay, az, features);
    }
    for (int off = 32; off > 0; off >>= 1) {
        wa += __shfl_xor(wa, off);
        we += __shfl_xor(we, off);
    }
    if (lane == 0) {
        atomicAdd(&counters[0], wa);
        atomicAdd(&counters[1], we);
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

InternalSoA internal_soa(BHState& s) {
    InternalSoA in;
    in.com = B<double4>(s.icom);
    in.rec = B<int4>(s.irec);
    in.depth = B<int32_t>(s.idepth);
    return in;
}

// Device word layout of s.flags: [0..7] flags, [8..8+129) depth counts.
constexpr int kDCount = 8;
constexpr int kFlagRead = 256;
constexpr int kFlagWords = 512;

}  // namespace

NBodySimProStatus bh_build(Engine* e, bool fp32_inputs, NBodySimProError* error,
                           bool quadrupoles) {
    BHState& s = e->bh;
    const size_t n = e->n;
    // the FP32 pre-order records are written on demand (bh_node_records)
    s.has_n32 = false;
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
    // build inputs: float4 for FP32-rounded inputs, double4 for FP64 ones
    NBX_TRY(s.b4.need<char>(n * (fp32_inputs ? sizeof(float4) : sizeof(double4)), error));
    float4* b4f = B<float4>(s.b4);
    double4* b4d = B<double4>(s.b4);
    s.inputs_fp64 = !fp32_inputs;
    const unsigned bb_blocks = g < static_cast<unsigned>(kBBoxBlocks) ? g : kBBoxBlocks;
    NBX_TRY(s.partial.need<double>(6 * kBBoxBlocks + 8, error));
    NBX_TRY(s.flags.need<unsigned>(kFlagWords, error));
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
    if (!fp32_inputs) {
        NBX_TRY(s.sp4.need<double4>(n, error));
    }
    NBX_TRY(s.s4.need<float4>(n, error));
    NBX_TRY(s.counters.need<unsigned long long>(3, error));
    NBX_TRY(s.deferred.need<int32_t>(n, error));
    const size_t tiles = (n + kTopoTile - 1) / kTopoTile;
    NBX_TRY(s.tile_hist.need<unsigned>(2 * kMaxLevels * tiles, error));
    NBX_TRY(s.tile_base.need<unsigned>(2 * kMaxLevels * tiles, error));
    const int nt = static_cast<int>(tiles);

    double* root = B<double>(s.partial) + 6 * kBBoxBlocks;
    unsigned* flags = B<unsigned>(s.flags);
    if (fp32_inputs) {
        bbox_partial_kernel<float4><<<bb_blocks, kThreads, 0, st>>>(e->px, e->py, e->pz, e->m,
                                                                    b4f, ni, B<double>(s.partial));
    } else {
        bbox_partial_kernel<double4><<<bb_blocks, kThreads, 0, st>>>(e->px, e->py, e->pz, e->m,
                                                                     b4d, ni, B<double>(s.partial));
    }
    bbox_final_kernel<<<1, kThreads, 0, st>>>(B<double>(s.partial), static_cast<int>(bb_blocks),
                                              root, flags, kFlagWords,
                                              B<unsigned long long>(s.counters));
    if (fp32_inputs) {
        keys_kernel<float4><<<g, kThreads, 0, st>>>(b4f, ni, root, B<uint64_t>(s.mkey),
                                                    B<uint64_t>(s.tkey), B<uint64_t>(s.sk0),
                                                    B<uint32_t>(s.sv0), flags);
    } else {
        keys_kernel<double4><<<g, kThreads, 0, st>>>(b4d, ni, root, B<uint64_t>(s.mkey),
                                                     B<uint64_t>(s.tkey), B<uint64_t>(s.sk0),
                                                     B<uint32_t>(s.sv0), flags);
    }
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[1], st));

    // Morton order: stable sort of (key, original index) on the top key bits
    // (the result stays in sk0/sv0), then each run of equal top bits ordered
    // by the low bits. (Measured on MI355X, two-galaxy, against seven passes
    // on all 63 bits: sort 1M 0.171 -> 0.108 ms, 10M 0.894 -> 0.641 ms.)
    NBX_TRY(radix_sort_pairs(st, B<uint64_t>(s.sk0), B<uint32_t>(s.sv0), B<uint64_t>(s.sk1),
                             B<uint32_t>(s.sv1), n, kSortTopBits, s.hist, error, nullptr,
                             nullptr, kSortLowBits));
    sort_finish_kernel<<<g, kThreads, 0, st>>>(B<uint64_t>(s.sk0), B<uint32_t>(s.sv0),
                                               B<uint64_t>(s.mkey_sorted), B<uint32_t>(s.mperm),
                                               ni, flags);
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[9], st));

    // Tree order, neighbour depths and per-depth node counts for a given
    // ascending true-path order (null: Morton order), then one host read of
    // the flags, counts and root cube.
    if (s.host_flags == nullptr) {
        NBX_HIP(error, hipHostMalloc(&s.host_flags, kFlagRead * sizeof(unsigned) + 4 * sizeof(double),
                                     hipHostMallocDefault));
    }
    unsigned* hflags = static_cast<unsigned*>(s.host_flags);
    double* hroot = reinterpret_cast<double*>(hflags + kFlagRead);
    auto topology = [&](const uint32_t* asc_rank, const uint64_t* ext_asc, uint64_t* ext_t)
        -> NBodySimProStatus {
        if (fp32_inputs) {
            tree_order_kernel<float4><<<g, kThreads, 0, st>>>(
                asc_rank, B<uint32_t>(s.mperm), B<uint64_t>(s.mkey_sorted), B<uint64_t>(s.tkey),
                ext_asc, b4f, ni, B<uint32_t>(s.tperm), B<uint32_t>(s.trank),
                B<uint64_t>(s.tword), ext_t, nullptr, B<float4>(s.s4));
        } else {
            tree_order_kernel<double4><<<g, kThreads, 0, st>>>(
                asc_rank, B<uint32_t>(s.mperm), B<uint64_t>(s.mkey_sorted), B<uint64_t>(s.tkey),
                ext_asc, b4d, ni, B<uint32_t>(s.tperm), B<uint32_t>(s.trank),
                B<uint64_t>(s.tword), ext_t, B<double4>(s.sp4), B<float4>(s.s4));
        }
        delta_kernel<<<nt, kThreads, 0, st>>>(B<uint64_t>(s.tword), ext_t, ni, nt,
                                              B<int32_t>(s.delta), B<uint32_t>(s.cnt), flags,
                                              B<unsigned>(s.tile_hist));
        depth_scan_kernel<<<2 * kMaxLevels, kThreads, 0, st>>>(B<unsigned>(s.tile_hist), nt,
                                                           B<unsigned>(s.tile_base),
                                                           flags + kDCount);
        NBX_HIP(error, hipGetLastError());
        NBX_HIP(error, hipMemcpyAsync(hflags, flags, kFlagRead * sizeof(unsigned),
                                      hipMemcpyDeviceToHost, st));
        NBX_HIP(error, hipMemcpyAsync(hroot, root, 4 * sizeof(double), hipMemcpyDeviceToHost,
                                      st));
        NBX_HIP(error, hipEventRecord(e->ev[10], st));
        // node numbering (internal nodes started per tree position, scanned)
        // runs on the device while the host waits for the flags; the copies
        // go to pinned memory the engine owns, so they land whether or not
        // this returns early
        const NBodySimProStatus scanned =
            exclusive_scan_u32(st, B<uint32_t>(s.cnt), B<uint32_t>(s.scan), n, s.scan_part,
                               error);
        // Poll rather than block: the device idles until the host has read
        // the per-depth counts and enqueued the rest of the build.
        hipError_t done;
        while ((done = hipEventQuery(e->ev[10])) == hipErrorNotReady) {
        }
        NBX_HIP(error, done);
        std::memcpy(s.root, hroot, sizeof(s.root));
        return scanned;
    };
    // The Morton order is the tree order unless some true-path key differs
    // from its Morton key (keys_kernel raises the flag): build on it and redo
    // the topology in that rare case.
    NBX_TRY(topology(nullptr, nullptr, nullptr));
    s.mismatch_path = hflags[FLAG_MISMATCH] != 0;
    s.tie_path = false;
    if (hflags[FLAG_LONG_RUN] != 0) {
        // a run of equal top key bits too long for sort_finish_kernel (the
        // Morton order above was wrong there): sort on all 63 bits, then
        // redo the topology
        NBX_HIP(error, hipMemcpyAsync(s.sk0.p, s.mkey.p, n * sizeof(uint64_t),
                                      hipMemcpyDeviceToDevice, st));
        iota_kernel<<<g, kThreads, 0, st>>>(B<uint32_t>(s.sv0), ni);
        NBX_TRY(radix_sort_pairs(st, B<uint64_t>(s.sk0), B<uint32_t>(s.sv0), B<uint64_t>(s.sk1),
                                 B<uint32_t>(s.sv1), n, 63, s.hist, error,
                                 B<uint64_t>(s.mkey_sorted), B<uint32_t>(s.mperm)));
        NBX_HIP(error, hipMemsetAsync(flags, 0, kFlagWords * sizeof(unsigned), st));
        NBX_TRY(topology(nullptr, nullptr, nullptr));
    }
    if (s.mismatch_path) {
        // asc_rank[k]: Morton rank of the k-th particle in ascending true-path
        // order (ties keep Morton order): the true keys in Morton order,
        // sorted like the Morton keys (top bits, then the short runs).
        gather_u64_kernel<<<g, kThreads, 0, st>>>(B<uint64_t>(s.tkey), B<uint32_t>(s.mperm),
                                                  B<uint64_t>(s.sk0), B<uint32_t>(s.sv0), ni, 1, 0);
        NBX_TRY(radix_sort_pairs(st, B<uint64_t>(s.sk0), B<uint32_t>(s.sv0), B<uint64_t>(s.sk1),
                                 B<uint32_t>(s.sv1), n, kSortTopBits, s.hist, error, nullptr,
                                 nullptr, kSortLowBits));
        NBX_HIP(error, hipMemsetAsync(flags, 0, kFlagWords * sizeof(unsigned), st));
        sort_finish_kernel<<<g, kThreads, 0, st>>>(B<uint64_t>(s.sk0), B<uint32_t>(s.sv0),
                                                   B<uint64_t>(s.sk1), B<uint32_t>(s.tmp32), ni,
                                                   flags);
        NBX_HIP(error, hipGetLastError());
        NBX_TRY(topology(B<uint32_t>(s.tmp32), nullptr, nullptr));
        if (hflags[FLAG_LONG_RUN] != 0) {
            // a run of equal top bits too long for the finish kernel: all 63 bits
            gather_u64_kernel<<<g, kThreads, 0, st>>>(B<uint64_t>(s.tkey), B<uint32_t>(s.mperm),
                                                      B<uint64_t>(s.sk0), B<uint32_t>(s.sv0), ni,
                                                      1, 0);
            NBX_TRY(radix_sort_pairs(st, B<uint64_t>(s.sk0), B<uint32_t>(s.sv0),
                                     B<uint64_t>(s.sk1), B<uint32_t>(s.sv1), n, 63, s.hist,
                                     error));
            NBX_HIP(error, hipMemcpyAsync(s.tmp32.p, s.sv0.p, n * sizeof(uint32_t),
                                          hipMemcpyDeviceToDevice, st));
            NBX_HIP(error, hipMemsetAsync(flags, 0, kFlagWords * sizeof(unsigned), st));
            NBX_TRY(topology(B<uint32_t>(s.tmp32), nullptr, nullptr));
        }
    }

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
        if (fp32_inputs) {
            ext_words_kernel<float4><<<g, kThreads, 0, st>>>(b4f, ni, root, B<uint32_t>(s.mperm),
                                                             ext_m);
        } else {
            ext_words_kernel<double4><<<g, kThreads, 0, st>>>(b4d, ni, root,
                                                              B<uint32_t>(s.mperm), ext_m);
        }
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
            gather_u64_kernel<<<g, kThreads, 0, st>>>(ext_m, P, B<uint64_t>(s.sk0), nullptr, ni,
                                                      kExtWords, k);
            NBX_HIP(error, hipMemcpy2DAsync(ext_a + k, kExtWords * sizeof(uint64_t), s.sk0.p,
                                            sizeof(uint64_t), sizeof(uint64_t), n,
                                            hipMemcpyDeviceToDevice, st));
        }
        NBX_HIP(error, hipMemsetAsync(flags, 0, kFlagWords * sizeof(unsigned), st));
        NBX_TRY(topology(P, ext_a, ext_t));
        ext_tree = ext_t;
    }
    if (hflags[FLAG_OVERFLOW] != 0) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_OVERFLOW,
                 "Barnes-Hut tree exceeded maximum depth (duplicate positions?)");
    }
    size_t internal = 0;
    int deepest = -1;
    for (int d = 0; d < kMaxLevels; ++d) {
        internal += hflags[kDCount + d];
        if (hflags[kDCount + d] != 0) deepest = d;
    }

    const size_t nodes = n + internal;
    NBX_TRY(s.icom.need<double4>(internal + 1, error));
    NBX_TRY(s.irec.need<int4>(internal + 1, error));
    NBX_TRY(s.idepth.need<int32_t>(internal + 1, error));
    NBX_TRY(s.gnode.need<GroupNode>(internal + 1, error));
    // FP64 node records feed only the exact FP64 walk
    NodeF64* n64_out = nullptr;
    s.has_n64 = false;
    // quadrupole records only for the learned correction's feature walk
    double* im2 = nullptr;
    float4* gquad = nullptr;
    s.has_quad = false;
    if (quadrupoles) {
        NBX_TRY(s.im2.need<double>(6 * (internal + 1), error));
        NBX_TRY(s.gquad.need<float4>(2 * (internal + 1), error));
        im2 = B<double>(s.im2);
        gquad = B<float4>(s.gquad);
        s.has_quad = true;
    }
    if (!fp32_inputs) {
        NBX_TRY(s.n64.need<NodeF64>(nodes, error));
        n64_out = B<NodeF64>(s.n64);
        s.has_n64 = true;
    }
    const InternalSoA in = internal_soa(s);
    // tree-order FP64 inputs (FP64 builds; FP32 builds widen s4)
    const double4* sp4 = fp32_inputs ? nullptr : B<double4>(s.sp4);
    const float4* s4 = B<float4>(s.s4);
    s.ext_tree = ext_tree;
    const unsigned* start_base = B<unsigned>(s.tile_base);
    const unsigned* end_base = start_base + static_cast<size_t>(kMaxLevels) * tiles;
    ends_kernel<<<nt, kTopoTile, 0, st>>>(B<int32_t>(s.delta), B<uint32_t>(s.cnt),
                                         B<uint32_t>(s.scan), ni, nt, flags + kDCount, end_base,
                                         in.rec, B<GroupNode>(s.gnode));
    fill_kernel<<<nt, kTopoTile, 0, st>>>(B<uint64_t>(s.tword), ext_tree, B<int32_t>(s.delta),
                                         B<uint32_t>(s.cnt), B<uint32_t>(s.scan), ni, nt, sp4,
                                         root, in, n64_out, flags + kDCount,
                                         start_base, B<GroupNode>(s.gnode));
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[3], st));
    std::vector<unsigned> offs(kMaxLevels, 0);
    for (int d = 0, run = 0; d < kMaxLevels; ++d) {
        offs[d] = static_cast<unsigned>(run);
        run += static_cast<int>(hflags[kDCount + d]);
    }
    // Deepest level first. Runs of small levels share one single-workgroup
    // launch; a large level gets a full-grid launch of its own. (Measured on
    // MI355X, 1M two-galaxy: small levels up to 512 cells 0.113 ms for the
    // whole pass, up to 1,536 cells 0.119 ms, up to 2,560 cells 0.136 ms.
    // Two adjacent large levels per launch, each workgroup computing its
    // cells' children first: 0.100 -> 0.114 ms at 1M, 0.40 -> 0.45 ms at 10M.)
    constexpr unsigned kSmallLevel = 512;
    // the four centre-of-mass variants: quadrupole moments or not, FP32 or
    // FP64 particle inputs
    auto com_level = [&](int d, unsigned count) {
        const unsigned grid = grid_for(count, kThreads);
        const int first = static_cast<int>(offs[d]);
        const int c = static_cast<int>(count);
        GroupNode* gn = B<GroupNode>(s.gnode);
        const uint32_t* tr = B<uint32_t>(s.trank);
        if (fp32_inputs) {
            if (quadrupoles) {
                com_level_kernel<true, float4><<<grid, kThreads, 0, st>>>(
                    first, c, d, gn, in, tr, s4, b4f, root, n64_out, im2, gquad);
            } else {
                com_level_kernel<false, float4><<<grid, kThreads, 0, st>>>(
                    first, c, d, gn, in, tr, s4, b4f, root, n64_out, nullptr, nullptr);
            }
        } else {
            if (quadrupoles) {
                com_level_kernel<true, double4><<<grid, kThreads, 0, st>>>(
                    first, c, d, gn, in, tr, sp4, b4d, root, n64_out, im2, gquad);
            } else {
                com_level_kernel<false, double4><<<grid, kThreads, 0, st>>>(
                    first, c, d, gn, in, tr, sp4, b4d, root, n64_out, nullptr, nullptr);
            }
        }
    };
    auto com_small_levels = [&](int d_hi, int d_lo) {
        GroupNode* gn = B<GroupNode>(s.gnode);
        const uint32_t* tr = B<uint32_t>(s.trank);
        const unsigned* dc = flags + kDCount;
        if (fp32_inputs) {
            if (quadrupoles) {
                com_small_levels_kernel<true, float4><<<1, 1024, 0, st>>>(
                    dc, d_hi, d_lo, gn, in, tr, s4, b4f, root, n64_out, im2, gquad);
            } else {
                com_small_levels_kernel<false, float4><<<1, 1024, 0, st>>>(
                    dc, d_hi, d_lo, gn, in, tr, s4, b4f, root, n64_out, nullptr, nullptr);
            }
        } else {
            if (quadrupoles) {
                com_small_levels_kernel<true, double4><<<1, 1024, 0, st>>>(
                    dc, d_hi, d_lo, gn, in, tr, sp4, b4d, root, n64_out, im2, gquad);
            } else {
                com_small_levels_kernel<false, double4><<<1, 1024, 0, st>>>(
                    dc, d_hi, d_lo, gn, in, tr, sp4, b4d, root, n64_out, nullptr, nullptr);
            }
        }
    };
    for (int d = deepest; d >= 0;) {
        const unsigned count = hflags[kDCount + d];
        if (count > kSmallLevel) {
            com_level(d, count);
            --d;
            continue;
        }
        int lo = d;
        while (lo - 1 >= 0 && hflags[kDCount + lo - 1] <= kSmallLevel) {
            --lo;
        }
        com_small_levels(d, lo);
        d = lo - 1;
    }
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[4], st));
    s.n = n;
    s.internal = internal;
    s.nodes = nodes;
    // levels of the reference's compute_maximum_depth (barnes_hut.c:269-283):
    // deepest leaf depth + 1, where the deepest leaf sits one below the deepest cell
    s.max_depth = n == 1 ? 0 : deepest + 1;
    e->stats.node_count = nodes;
    e->stats.leaf_count = n;
    e->stats.internal_node_count = internal;
    e->stats.maximum_depth = static_cast<size_t>(s.max_depth) + 1;
    e->stats.accepted_approximations = 0;
    e->stats.exact_interactions = 0;
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus bh_node_records(Engine* e, NBodySimProError* error) {
    BHState& s = e->bh;
    if (s.n == 0) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_STATE, "no Barnes-Hut build");
    }
    if (s.has_n32) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    NBX_TRY(s.n32.need<NodeF32>(s.nodes, error));
    const int ni = static_cast<int>(s.n);
    const int ic = static_cast<int>(s.internal);
    node_records_kernel<<<grid_for(s.nodes, kThreads), kThreads, 0, e->stream>>>(
        ni, ic, B<uint32_t>(s.scan), B<uint32_t>(s.cnt), B<float4>(s.s4), internal_soa(s),
        B<double>(s.partial) + 6 * kBBoxBlocks, B<NodeF32>(s.n32));
    NBX_HIP(error, hipGetLastError());
    s.has_n32 = true;
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus bh_walk(Engine* e, const NBodySimProGravity* g, bool fp64_exact, float* features,
                          NBodySimProError* error) {
    BHState& s = e->bh;
    hipStream_t st = e->stream;
    const int ni = static_cast<int>(s.n);
    // the counters were cleared by this build's bbox_final_kernel
    NBX_HIP(error, hipEventRecord(e->ev[5], st));
    const double theta2 = e->theta * e->theta;
    const unsigned blocks = grid_for(s.n, kThreads);
    if (fp64_exact && !s.has_n64) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_STATE,
                 "the exact walk needs a build on FP64 inputs");
    }
    if (features != nullptr && !s.has_quad) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_STATE,
                 "the feature walk needs a build with quadrupole records");
    }
    const float4* gq = features != nullptr ? B<float4>(s.gquad) : nullptr;
    const float Gf = static_cast<float>(g->gravitational_constant);
    const float e2 = static_cast<float>(g->softening_squared);
    const float th2 = static_cast<float>(theta2);
    unsigned long long* cnts = B<unsigned long long>(s.counters);
    // per-query reference-list walk of nq queries (all tree positions, or the
    // listed ones) on the FP32 pre-order records
    auto per_query_walk = [&](const int* list, int nq) -> NBodySimProStatus {
        NBX_TRY(bh_node_records(e, error));
        const unsigned wblocks = grid_for(static_cast<size_t>(nq), kThreads * kWalkQueries);
        const unsigned wgrid = (wblocks + kXcds - 1) / kXcds * kXcds;
        const NodeF32* n32 = B<NodeF32>(s.n32);
        const int nn = static_cast<int>(s.nodes);
        if (features != nullptr) {
            walk_f32_kernel<true, kWalkQueries><<<wgrid, kThreads, 0, st>>>(
                n32, nn, B<float4>(s.s4), B<uint32_t>(s.tperm), list, nq, Gf, e2, th2, e->ax,
                e->ay, e->az, cnts, features, gq);
        } else {
            walk_f32_kernel<false, kWalkQueries><<<wgrid, kThreads, 0, st>>>(
                n32, nn, B<float4>(s.s4), B<uint32_t>(s.tperm), list, nq, Gf, e2, th2, e->ax,
                e->ay, e->az, cnts, nullptr, nullptr);
        }
        NBX_HIP(error, hipGetLastError());
        return N_BODY_SIM_PRO_STATUS_OK;
    };
    bool group = false;
    if (fp64_exact) {
        walk_f64_kernel<<<blocks, kThreads, 0, st>>>(
            B<NodeF64>(s.n64), static_cast<int>(s.nodes), B<double4>(s.sp4), B<uint32_t>(s.tperm), ni, g->gravitational_constant,
            g->softening_squared, theta2, e->ax, e->ay, e->az, cnts);
    } else if (!(e->flags & NBX_FLAG_BH_REFERENCE_LISTS)) {
        group = true;
        const GroupNode* gn = B<GroupNode>(s.gnode);
        const int ni_cells = static_cast<int>(s.internal);
        int* deferred = B<int32_t>(s.deferred);
        // one shared walk per group (no split) or the split's budget
        const int max_walks = (e->flags & NBX_FLAG_BH_NO_SPLIT) ? 1 : kMaxSubWalks;
        // The member test is on unless the box test alone is asked for; the
        // feature walk of the learned path always takes it. (Measured on
        // MI355X, two-galaxy, box test alone -> with the member test: plain
        // walk at theta = 0.7, 1M: 0.844 -> 0.819 ms, 1.53G -> 1.18G
        // interactions; theta = 0.5, 10M: 18.3 -> 17.2 ms; feature walk at
        // theta = 1.0, 10M: 9.61 -> 7.96 ms.)
        const bool member_test =
            features != nullptr || (e->flags & NBX_FLAG_BH_BOX_TEST) == 0;
        // whole rounds of one chunk per XCD (blocks past the last group exit)
        const size_t chunks = (s.n + static_cast<size_t>(kWave) * kGroupChunk - 1) /
                              (static_cast<size_t>(kWave) * kGroupChunk);
        const unsigned grid_g =
            static_cast<unsigned>((chunks + kXcds - 1) / kXcds * kXcds * kGroupChunk);
        if (features != nullptr) {
            group_walk_kernel<true><<<grid_g, kGroupThreads, 0, st>>>(
                deferred, gn, ni_cells, B<float4>(s.s4), B<uint32_t>(s.tperm),
                B<uint64_t>(s.tword), s.ext_tree, ni, Gf, e2, th2, e->ax, e->ay, e->az, cnts,
                features, gq, max_walks, member_test);
        } else {
            group_walk_kernel<false><<<grid_g, kGroupThreads, 0, st>>>(
                deferred, gn, ni_cells, B<float4>(s.s4), B<uint32_t>(s.tperm),
                B<uint64_t>(s.tword), s.ext_tree, ni, Gf, e2, th2, e->ax, e->ay, e->az, cnts,
                nullptr, nullptr, max_walks, member_test);
        }
        NBX_HIP(error, hipGetLastError());
    } else {
        NBX_TRY(per_query_walk(nullptr, ni));
    }
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[6], st));
    unsigned long long counts[3] = {0, 0, 0};
    NBX_HIP(error, hipMemcpyAsync(counts, s.counters.p, sizeof(counts), hipMemcpyDeviceToHost, st));
    NBX_HIP(error, hipStreamSynchronize(st));
    if (group && counts[2] > 0) {
        // Queries the group walk could not give a shared list (a range the
        // split could not make compact within its budget, or a stack
        // overflow; none in most evaluations): their reference lists, after
        // the FP32 pre-order records are written.
        if (counts[2] > s.n) {
            NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_STATE,
                     "deferred query count out of range");
        }
        NBX_TRY(per_query_walk(B<int32_t>(s.deferred), static_cast<int>(counts[2])));
        NBX_HIP(error, hipEventRecord(e->ev[6], st));
        NBX_HIP(error, hipMemcpyAsync(counts, s.counters.p, 2 * sizeof(unsigned long long),
                                      hipMemcpyDeviceToHost, st));
        NBX_HIP(error, hipStreamSynchronize(st));
    }
    e->stats.accepted_approximations = counts[0];
    e->stats.exact_interactions = counts[1];
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus barnes_hut(Engine* e, const NBodySimProGravity* g, bool fp64_exact,
                             NBodySimProError* error) {
    NBX_TRY(bh_build(e, !fp64_exact, error));
    NBX_TRY(bh_walk(e, g, fp64_exact, nullptr, error));
    float t[10] = {0};
    for (int k : {1, 3, 4, 5, 6, 9}) {
        NBX_HIP(error, hipEventElapsedTime(&t[k], e->ev[0], e->ev[k]));
    }
    NBXTimings tm{};
    tm.bbox_keys_ms = t[1];
    tm.sort_ms = t[9] - t[1];
    tm.tree_ms = t[3] - t[9];
    tm.com_ms = t[4] - t[3];
    tm.walk_ms = t[6] - t[5];
    tm.force_ms = t[6];
    tm.kernel_ms = tm.walk_ms;
    const double interactions = static_cast<double>(e->stats.accepted_approximations +
                                                    e->stats.exact_interactions);
    tm.kernel_flops = 20.0 * interactions;
    // walk bytes: every query reads its position (and permutation entry) and
    // writes 3 doubles; the cell records the walk reads are counted once per
    // launch (they stay in L2 across waves): GroupNode records for the group
    // walk, pre-order node records for the per-query walks.
    const bool group = !fp64_exact && !(e->flags & NBX_FLAG_BH_REFERENCE_LISTS);
    const double node_bytes =
        fp64_exact ? static_cast<double>(e->bh.nodes) * sizeof(NodeF64)
        : group    ? static_cast<double>(e->bh.internal) * sizeof(GroupNode)
                   : static_cast<double>(e->bh.nodes) * sizeof(NodeF32);
    tm.kernel_bytes = static_cast<double>(e->bh.n) * ((fp64_exact ? 24.0 : 16.0) + 24.0 + 4.0) +
                      node_bytes;
    e->timings = tm;
    e->stats.build_time_seconds = t[4] * 1e-3;
    e->stats.evaluation_time_seconds = tm.walk_ms * 1e-3;
    return N_BODY_SIM_PRO_STATUS_OK;
}

}  // namespace nbx
