// nbx_sort.hip - stable LSD radix sort of (u64 key, u32 value) pairs and an
// exclusive u32 scan, both for gfx950 (wave64).
//
// The reference sorts a size_t permutation by 63-bit Morton keys with three
// stable 21-bit LSD passes and a 2^21-entry counting array
// (src/c/barnes_hut/barnes_hut.c:312-339). Any stable sort on the full key
// yields the same permutation (ties keep ascending original index), so here
// the digit is 8 bits: 256 buckets fit one LDS histogram per tile and the
// per-tile offsets stay small, which suits a 256-CU device better than a
// 16 MiB counting array.
#include "nbx_internal.h"

namespace nbx {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortIpt = 8;
constexpr int kSortTile = kSortThreads * kSortIpt;
constexpr int kRadix = 256;

__global__ __launch_bounds__(kSortThreads) void sort_histogram(const uint64_t* __restrict__ keys,
                                                              size_t n, int shift,
                                                              uint32_t* __restrict__ hist,
                                                              unsigned tiles) {
    __shared__ uint32_t h[kRadix];
    const unsigned tile = blockIdx.x;
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = static_cast<size_t>(tile) * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortIpt; ++r) {
        const size_t i = base + static_cast<size_t>(r) * kSortThreads + threadIdx.x;
        if (i < n) {
            atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
        }
    }
    __syncthreads();
    hist[static_cast<size_t>(threadIdx.x) * tiles + tile] = h[threadIdx.x];
}

// One block per digit: exclusive scan of that digit's per-tile counts.
__global__ __launch_bounds__(kSortThreads) void sort_scan_digit(uint32_t* __restrict__ hist,
                                                               unsigned tiles,
                                                               uint32_t* __restrict__ totals) {
    __shared__ uint32_t s[kSortThreads];
    uint32_t* row = hist + static_cast<size_t>(blockIdx.x) * tiles;
    uint32_t carry = 0;
    for (unsigned start = 0; start < tiles; start += kSortThreads) {
        const unsigned i = start + threadIdx.x;
        const uint32_t v = i < tiles ? row[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < kSortThreads; off <<= 1) {
            const uint32_t t = threadIdx.x >= static_cast<unsigned>(off) ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < tiles) {
            row[i] = carry + s[threadIdx.x] - v;
        }
        carry += s[kSortThreads - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        totals[blockIdx.x] = carry;
    }
}

__global__ __launch_bounds__(kSortThreads) void sort_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, size_t n, int shift,
    const uint32_t* __restrict__ hist, unsigned tiles, const uint32_t* __restrict__ totals) {
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t running[kRadix];
    __shared__ uint32_t wcount[kSortThreads / kWave][kRadix];
    const unsigned tid = threadIdx.x;
    const unsigned lane = tid & (kWave - 1);
    const unsigned wave = tid / kWave;
    const unsigned tile = blockIdx.x;

    // Digit bases: exclusive scan of the digit totals.
    base[tid] = totals[tid];
    __syncthreads();
    for (int off = 1; off < kRadix; off <<= 1) {
        const uint32_t t = tid >= static_cast<unsigned>(off) ? base[tid - off] : 0u;
        __syncthreads();
        base[tid] += t;
        __syncthreads();
    }
    const uint32_t my_total = totals[tid];
    __syncthreads();
    base[tid] = base[tid] - my_total + hist[static_cast<size_t>(tid) * tiles + tile];
    running[tid] = 0;
    __syncthreads();

    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (kWave - lane));
    const size_t tbase = static_cast<size_t>(tile) * kSortTile;
    for (int r = 0; r < kSortIpt; ++r) {
        const size_t i = tbase + static_cast<size_t>(r) * kSortThreads + tid;
        const bool valid = i < n;
        uint64_t key = 0;
        uint32_t val = 0;
        unsigned d = 0;
        if (valid) {
            key = kin[i];
            val = vin[i];
            d = static_cast<unsigned>((key >> shift) & 0xFFu);
        }
        // Lanes of this wave holding the same digit.
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const unsigned rank_in_wave = __popcll(peers & lt);
        for (unsigned k = tid; k < (kSortThreads / kWave) * kRadix; k += kSortThreads) {
            (&wcount[0][0])[k] = 0;
        }
        __syncthreads();
        if (valid && rank_in_wave == 0) {
            wcount[wave][d] = static_cast<uint32_t>(__popcll(peers));
        }
        __syncthreads();
        if (valid) {
            uint32_t off = running[d] + rank_in_wave;
            for (unsigned w = 0; w < wave; ++w) {
                off += wcount[w][d];
            }
            const uint32_t dst = base[d] + off;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (unsigned w = 0; w < kSortThreads / kWave; ++w) {
            add += wcount[w][tid];
        }
        running[tid] += add;
        __syncthreads();
    }
}

constexpr int kScanThreads = 256;
constexpr int kScanIpt = 4;
constexpr int kScanTile = kScanThreads * kScanIpt;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* s, uint32_t* total) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const uint32_t t = threadIdx.x >= static_cast<unsigned>(off) ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    const uint32_t incl = s[threadIdx.x];
    *total = s[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_tiles(const uint32_t* __restrict__ in,
                                                          size_t n,
                                                          uint32_t* __restrict__ partial) {
    __shared__ uint32_t s[kScanThreads];
    const size_t base = static_cast<size_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanIpt;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanIpt; ++k) {
        sum += (base + k < n) ? in[base + k] : 0u;
    }
    uint32_t total;
    block_exclusive_scan(sum, s, &total);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = total;
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_partials(uint32_t* __restrict__ partial,
                                                             unsigned count) {
    __shared__ uint32_t s[kScanThreads];
    uint32_t carry = 0;
    for (unsigned start = 0; start < count; start += kScanThreads) {
        const unsigned i = start + threadIdx.x;
        const uint32_t v = i < count ? partial[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, s, &total);
        if (i < count) {
            partial[i] = carry + ex;
        }
        carry += total;
    }
    if (threadIdx.x == 0) {
        partial[count] = carry;
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_apply(const uint32_t* __restrict__ in,
                                                          size_t n,
                                                          const uint32_t* __restrict__ partial,
                                                          unsigned tiles,
                                                          uint32_t* __restrict__ out) {
    __shared__ uint32_t s[kScanThreads];
    const size_t base = static_cast<size_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanIpt;
    uint32_t v[kScanIpt];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanIpt; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, s, &total) + partial[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanIpt; ++k) {
        if (base + k < n) {
            out[base + k] = run;
        }
        run += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out[n] = partial[tiles];
    }
}

}  // namespace

NBodySimProStatus radix_sort_pairs(hipStream_t s, uint64_t* keys, uint32_t* vals,
                                   uint64_t* keys_alt, uint32_t* vals_alt, size_t n, int bits,
                                   Buf& hist, NBodySimProError* error) {
    if (n <= 1) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    if (n > 0xFFFFFFFFull) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "sort size exceeds 2^32");
    }
    const unsigned tiles = grid_for(n, kSortTile);
    NBX_TRY(hist.need<uint32_t>(static_cast<size_t>(kRadix) * tiles + kRadix, error));
    uint32_t* h = hist.as<uint32_t>();
    uint32_t* totals = h + static_cast<size_t>(kRadix) * tiles;
    const int passes = (bits + 7) / 8;
    uint64_t* kin = keys;
    uint32_t* vin = vals;
    uint64_t* kout = keys_alt;
    uint32_t* vout = vals_alt;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        sort_histogram<<<tiles, kSortThreads, 0, s>>>(kin, n, shift, h, tiles);
        sort_scan_digit<<<kRadix, kSortThreads, 0, s>>>(h, tiles, totals);
        sort_scatter<<<tiles, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, shift, h, tiles,
                                                    totals);
        uint64_t* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
    }
    NBX_HIP(error, hipGetLastError());
    if (kin != keys) {
        NBX_HIP(error, hipMemcpyAsync(keys, kin, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        NBX_HIP(error, hipMemcpyAsync(vals, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus exclusive_scan_u32(hipStream_t s, const uint32_t* in, uint32_t* out, size_t n,
                                     Buf& partial, NBodySimProError* error) {
    if (n == 0) {
        NBX_HIP(error, hipMemsetAsync(out, 0, sizeof(uint32_t), s));
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    const unsigned tiles = grid_for(n, kScanTile);
    NBX_TRY(partial.need<uint32_t>(tiles + 1, error));
    uint32_t* part = partial.as<uint32_t>();
    scan_tiles<<<tiles, kScanThreads, 0, s>>>(in, n, part);
    scan_partials<<<1, kScanThreads, 0, s>>>(part, tiles);
    scan_apply<<<tiles, kScanThreads, 0, s>>>(in, n, part, tiles, out);
    NBX_HIP(error, hipGetLastError());
    return N_BODY_SIM_PRO_STATUS_OK;
}

}  // namespace nbx
