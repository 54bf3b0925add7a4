// nbx_internal.h - engine state and helpers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "nbx_engine.h"

namespace nbx {

constexpr int kWave = 64;
// Floats per query of the FP32 walk's statistics record (learned correction).
constexpr int kWalkRecord = 16;

// Fills the optional error record (reference contract: core/status.h:37-49).
void set_error(NBodySimProError* error, NBodySimProStatus status, const char* file, int line,
               const char* message);
NBodySimProStatus map_hip_error(hipError_t e);

#define NBX_FAIL(err, st, msg)                                 \
    do {                                                       \
        ::nbx::set_error((err), (st), __FILE__, __LINE__, (msg)); \
        return (st);                                           \
    } while (0)

#define NBX_HIP(err, expr)                                                        \
    do {                                                                          \
        hipError_t nbx_e_ = (expr);                                               \
        if (nbx_e_ != hipSuccess) {                                               \
            NBodySimProStatus nbx_s_ = ::nbx::map_hip_error(nbx_e_);              \
            char nbx_m_[200];                                                     \
            std::snprintf(nbx_m_, sizeof(nbx_m_), "%s: %s", #expr,                \
                          hipGetErrorString(nbx_e_));                             \
            ::nbx::set_error((err), nbx_s_, __FILE__, __LINE__, nbx_m_);          \
            return nbx_s_;                                                        \
        }                                                                         \
    } while (0)

#define NBX_TRY(expr)                               \
    do {                                            \
        NBodySimProStatus nbx_t_ = (expr);          \
        if (nbx_t_ != N_BODY_SIM_PRO_STATUS_OK) {   \
            return nbx_t_;                          \
        }                                           \
    } while (0)

// Grows a device allocation to at least `bytes` (contents not preserved).
NBodySimProStatus ensure(void** ptr, size_t* cap, size_t bytes, NBodySimProError* error);
template <typename T>
inline NBodySimProStatus ensure_n(T** ptr, size_t* cap, size_t count, NBodySimProError* error) {
    return ensure(reinterpret_cast<void**>(ptr), cap, count * sizeof(T), error);
}

struct Buf {
    void* p = nullptr;
    size_t cap = 0;
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
    template <typename T>
    NBodySimProStatus need(size_t count, NBodySimProError* error) {
        return ensure(&p, &cap, count * sizeof(T), error);
    }
};

// FP32 walk node: centre of mass, mass, opening size^2 (4h^2), pre-order skip,
// leaf, and 1 / size^2 (walk statistics of the learned correction; 0 for leaves).
struct alignas(32) NodeF32 {
    float cx, cy, cz, m;
    float size2;
    int32_t skip;
    int32_t leaf;   // tree position of the leaf particle; -(level index + 2) for an internal cell
    float inv_size2;
};

// Internal cell by level index for the group walk: FP32 centre of mass and
// mass, opening size^2 and its reciprocal, and the child table (internal
// child: its level index; leaf child: -(tree position + 2); empty: -1).
struct alignas(64) GroupNode {
    float cx, cy, cz, m;
    float size2, inv_size2;
    int32_t pad[2];
    int32_t kids[8];
};

// FP64 walk node (exact reference arithmetic).
struct alignas(16) NodeF64 {
    double cx, cy, cz, m;
    double h;       // cell half size at this depth
    int32_t skip;
    int32_t leaf;
};

// Barnes-Hut device workspace (all sized by particle or node capacity).
struct BHState {
    size_t n = 0;             // particles in the last build
    size_t internal = 0;      // internal nodes in the last build
    size_t nodes = 0;         // n + internal
    int max_depth = 0;
    bool tie_path = false;
    bool mismatch_path = false;
    bool has_n64 = false;     // the last build wrote FP64 node records
    bool has_quad = false;    // the last build wrote quadrupole records
    bool has_n32 = false;     // the FP32 pre-order records of the last build are written
    bool inputs_fp64 = false; // the last build ran on FP64 inputs (b4 and sp4 are double4)
    const uint64_t* ext_tree = nullptr;   // tree-order extension words (tie path) or null
    double root[4] = {0, 0, 0, 1};
    // pinned host block the build reads its flags, per-depth counts and root
    // cube into (asynchronous copies; owned here so they always land)
    void* host_flags = nullptr;
    // build inputs (x, y, z, m): FP32-rounded copies of the master arrays
    // (float4) or the FP64 values (double4)
    Buf b4;                   // (x, y, z, m) per particle in original order
    Buf partial;              // bbox partials + root
    Buf flags;                // flags, per-depth counts, offsets and cursors
    Buf mkey, tkey;           // Morton / true-path keys in original order
    Buf sk0, sk1, sv0, sv1;   // sort ping-pong
    Buf hist;                 // sort histograms
    Buf mperm;                // Morton order: sorted position -> original index
    Buf mkey_sorted;          // Morton keys ascending
    Buf tperm;                // tree position -> original index
    Buf trank;                // tree position -> Morton rank
    Buf tword;                // tree-order true keys (descending)
    Buf ext;                  // extension key words (6 per particle) on the tie path
    Buf delta;                // int32 common levels between tree positions t, t+1
    Buf cnt, scan;            // internal nodes started per tree position, exclusive scan
    Buf scan_part;            // scan block partials
    Buf sp4;                  // tree-order FP64 (x, y, z, m), FP64 builds only
    Buf s4;                   // tree-order FP32 x, y, z, m
    Buf tile_hist, tile_base; // per-tile, per-depth internal-node counts and their scan
    // nodes (pre-order, children in descending octant order)
    Buf n64;                  // NodeF64 (FP64 builds)
    Buf n32;                  // NodeF32, written on demand (bh_node_records)
    // internal cells by level index (depth-major, tree order within a depth):
    // FP64 centre of mass and mass (double4), (node index, skip, first and
    // last tree position) (int4), depth, GroupNode (FP32 record and child
    // table)
    Buf icom, irec, idepth, gnode;
    // second moments about the centre of mass (FP64, 6 per cell) and the
    // walk's quadrupole records (2 float4 per cell, by level index; the
    // per-query walks reach them through the level index in their node
    // records), learned path only
    Buf im2, gquad;
    Buf tmp32;                // scratch permutation
    Buf counters;             // unsigned long long [approx, exact, deferred queries]
    Buf deferred;             // tree positions the group walk leaves to the per-query walk
};

struct Engine {
    int device = 0;
    int cu_count = 256;       // compute units of the device (grid shape choices)
    hipStream_t stream = nullptr;
    size_t capacity = 0;
    size_t n = 0;
    double theta = 0.7;
    uint32_t flags = 0;       // NBXConfig.flags
    // FP64 master state (SoA)
    double *px = nullptr, *py = nullptr, *pz = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *ax = nullptr, *ay = nullptr, *az = nullptr;
    double* m = nullptr;
    float4* pos4 = nullptr;   // FP32 (x, y, z, m) mirror of the master
    bool pos4_valid = false;
    Buf scratch;              // diagnostics / misc
    Buf scratch2;
    BHState bh;
    NodeF32* learned_dummy = nullptr;
    NBodySimProBarnesHutStats stats{};
    NBXTimings timings{};
    hipEvent_t ev[12]{};
    // learned BH correction
    void* mlp = nullptr;
};

// ---- entry points implemented per translation unit ----
NBodySimProStatus direct_fp64(Engine* e, const NBodySimProGravity* g, NBodySimProError* error);
NBodySimProStatus direct_fp32(Engine* e, const NBodySimProGravity* g, NBodySimProError* error);
NBodySimProStatus refresh_pos4(Engine* e, NBodySimProError* error);
NBodySimProStatus barnes_hut(Engine* e, const NBodySimProGravity* g, bool fp64_exact,
                             NBodySimProError* error);
NBodySimProStatus bh_build(Engine* e, bool fp32_inputs, NBodySimProError* error,
                           bool quadrupoles = false);
// The FP32 pre-order node records of the last build (NodeF32, for the
// per-query walks and the node export), written once per build on first use.
NBodySimProStatus bh_node_records(Engine* e, NBodySimProError* error);
// One walk per bh_build: the build clears the walk's interaction counters.
NBodySimProStatus bh_walk(Engine* e, const NBodySimProGravity* g, bool fp64_exact,
                          float* features, NBodySimProError* error);
NBodySimProStatus bh_learned(Engine* e, const NBodySimProGravity* g, NBodySimProError* error);

// Stable LSD radix sort of (64-bit key, 32-bit value) pairs, ascending key;
// equal keys keep their input order. It sorts by key bits first_bit ..
// first_bit + bits - 1 (the bits below first_bit are ignored: equal there
// counts as equal). The result goes to keys_out/vals_out when given (they
// must not alias the alternate buffers), else back into keys/vals.
NBodySimProStatus radix_sort_pairs(hipStream_t s, uint64_t* keys, uint32_t* vals,
                                   uint64_t* keys_alt, uint32_t* vals_alt, size_t n, int bits,
                                   Buf& hist, NBodySimProError* error,
                                   uint64_t* keys_out = nullptr, uint32_t* vals_out = nullptr,
                                   int first_bit = 0);
// Exclusive scan of uint32 (total written to out[n]).
NBodySimProStatus exclusive_scan_u32(hipStream_t s, const uint32_t* in, uint32_t* out, size_t n,
                                     Buf& partial, NBodySimProError* error);

// Bits of `mask` set in the lanes below this one (v_mbcnt_lo/hi: two VALU
// instructions, where and-ing a per-lane mask and counting bits takes four).
__device__ inline unsigned lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u));
}

inline unsigned grid_for(size_t n, unsigned block) {
    return static_cast<unsigned>((n + block - 1) / block);
}

}  // namespace nbx

struct NBXEngine {
    nbx::Engine e;
};
