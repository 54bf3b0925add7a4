// nbx_sort.hip - stable LSD radix sort of (u64 key, u32 value) pairs and an
// exclusive u32 scan, both for gfx950 (wave64).
//
// The reference sorts a size_t permutation by 63-bit Morton keys with three
// stable 21-bit LSD passes and a 2^21-entry counting array
// (src/c/barnes_hut/barnes_hut.c:312-339). Any stable sort on the full key
// yields the same permutation (ties keep ascending original index). Here the
// digit is 9 bits (7 passes cover 63 bits; a caller may sort a window of the
// key bits, as the Morton sort does with the top 36) and a pass is three
// launches:
//   sort_histogram   per-tile digit counts (2048-item tiles, 4096 for large
//                    inputs; LDS atomics);
//   sort_scan_digit  per digit, exclusive scan over tiles (+ digit totals);
//   sort_scatter     each tile ranks its items stably in LDS (wave ballots
//                    match equal digits; waves and rounds keep input order),
//                    stages them digit-sorted in LDS, then writes each
//                    digit's run to its global slot with consecutive lanes
//                    on consecutive addresses.
#include "nbx_internal.h"

namespace nbx {
namespace {

constexpr int kSortThreads = 512;
constexpr int kSortWaves = kSortThreads / kWave;
// Items per thread: 4 (2048-item tiles), or 8 (4096) from kLargeSort items on.
// (Measured on MI355X, 63-bit keys: 1M items 0.173 ms with 4 and 0.178 ms
// with 8 per thread; 10M 1.01 ms and 0.90 ms; 2 per thread is slower at both.)
constexpr size_t kLargeSort = size_t{1} << 22;
constexpr int kDigitBits = 9;
constexpr int kRadix = 1 << kDigitBits;             // 512
constexpr int kDigitsPerThread = kRadix / kSortThreads;
static_assert(kRadix % kSortThreads == 0, "each thread owns whole digits");

__device__ inline unsigned digit_of(uint64_t key, int shift) {
    return static_cast<unsigned>((key >> shift) & (kRadix - 1));
}

template <int IPT>
__global__ __launch_bounds__(kSortThreads) void sort_histogram(const uint64_t* __restrict__ keys,
                                                              size_t n, int shift,
                                                              uint32_t* __restrict__ hist,
                                                              unsigned tiles) {
    constexpr int kSortTile = kSortThreads * IPT;
    __shared__ uint32_t h[kRadix];
    const unsigned tile = blockIdx.x;
    for (int k = threadIdx.x; k < kRadix; k += kSortThreads) {
        h[k] = 0;
    }
    __syncthreads();
    const size_t base = static_cast<size_t>(tile) * kSortTile;
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const size_t i = base + static_cast<size_t>(r) * kSortThreads + threadIdx.x;
        if (i < n) {
            atomicAdd(&h[digit_of(keys[i], shift)], 1u);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kRadix; k += kSortThreads) {
        hist[static_cast<size_t>(k) * tiles + tile] = h[k];
    }
}

// Inclusive scan over one wave.
__device__ inline uint32_t wave_inclusive_scan(uint32_t v, unsigned lane) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t t = __shfl_up(v, off);
        if (lane >= static_cast<unsigned>(off)) {
            v += t;
        }
    }
    return v;
}

// Exclusive scan over the block's kSortThreads threads; `total` gets the block sum.
// `wsum` holds kSortWaves entries; the caller syncs before reusing it.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum, uint32_t& total) {
    const unsigned lane = threadIdx.x & (kWave - 1);
    const unsigned wave = threadIdx.x / kWave;
    const uint32_t incl = wave_inclusive_scan(v, lane);
    if (lane == kWave - 1) {
        wsum[wave] = incl;
    }
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
        before += static_cast<unsigned>(w) < wave ? wsum[w] : 0u;
        total += wsum[w];
    }
    return before + incl - v;
}

// One block per digit: exclusive scan of that digit's per-tile counts.
__global__ __launch_bounds__(kSortThreads) void sort_scan_digit(uint32_t* __restrict__ hist,
                                                               unsigned tiles,
                                                               uint32_t* __restrict__ totals) {
    __shared__ uint32_t wsum[kSortWaves];
    uint32_t* row = hist + static_cast<size_t>(blockIdx.x) * tiles;
    uint32_t carry = 0;
    for (unsigned start = 0; start < tiles; start += kSortThreads) {
        const unsigned i = start + threadIdx.x;
        const uint32_t v = i < tiles ? row[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, wsum, total);
        if (i < tiles) {
            row[i] = carry + ex;
        }
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        totals[blockIdx.x] = carry;
    }
}

template <int IPT>
__global__ __launch_bounds__(kSortThreads) void sort_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, size_t n, int shift,
    const uint32_t* __restrict__ hist, unsigned tiles, const uint32_t* __restrict__ totals) {
    constexpr int kSortTile = kSortThreads * IPT;
    __shared__ uint64_t skey[kSortTile];
    __shared__ uint32_t sval[kSortTile];
    __shared__ uint32_t gbase[kRadix];     // global slot of this tile's first item of digit d
    __shared__ uint32_t lbase[kRadix];     // start of digit d in the tile's staged order
    __shared__ uint32_t running[kRadix];   // items of digit d staged so far
    __shared__ uint32_t wcount[kSortWaves][kRadix];
    __shared__ uint32_t wsum[kSortWaves];
    const unsigned tid = threadIdx.x;
    const unsigned lane = tid & (kWave - 1);
    const unsigned wave = tid / kWave;
    const unsigned tile = blockIdx.x;

    // Digit starts over the whole array and within this tile; each thread
    // owns kDigitsPerThread consecutive digits.
    uint32_t tot[kDigitsPerThread], cnt[kDigitsPerThread], off[kDigitsPerThread];
    uint32_t tsum = 0, csum = 0;
#pragma unroll
    for (int j = 0; j < kDigitsPerThread; ++j) {
        const unsigned d = tid * kDigitsPerThread + j;
        tot[j] = totals[d];
        off[j] = hist[static_cast<size_t>(d) * tiles + tile];
        const uint32_t next = tile + 1 < tiles ? hist[static_cast<size_t>(d) * tiles + tile + 1]
                                               : tot[j];
        cnt[j] = next - off[j];
        tsum += tot[j];
        csum += cnt[j];
    }
    uint32_t total;
    uint32_t gstart = block_exclusive_scan(tsum, wsum, total);
    __syncthreads();
    uint32_t lstart = block_exclusive_scan(csum, wsum, total);
#pragma unroll
    for (int j = 0; j < kDigitsPerThread; ++j) {
        const unsigned d = tid * kDigitsPerThread + j;
        gbase[d] = gstart + off[j];
        lbase[d] = lstart;
        running[d] = 0;
        gstart += tot[j];
        lstart += cnt[j];
    }
    for (int k = tid; k < kSortWaves * kRadix; k += kSortThreads) {
        (&wcount[0][0])[k] = 0;
    }
    __syncthreads();

    // Stable ranks: rounds in input order, waves in order within a round,
    // lanes in order within a wave.
    const size_t tbase = static_cast<size_t>(tile) * kSortTile;
    for (int r = 0; r < IPT; ++r) {
        const size_t i = tbase + static_cast<size_t>(r) * kSortThreads + tid;
        const bool valid = i < n;
        uint64_t key = 0;
        uint32_t val = 0;
        unsigned d = 0;
        if (valid) {
            key = kin[i];
            val = vin[i];
            d = digit_of(key, shift);
        }
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const unsigned rank = lanes_below(peers);
        const bool leader = valid && rank == 0;
        const uint32_t group = static_cast<uint32_t>(__popcll(peers));
        if (leader) {
            wcount[wave][d] = group;
        }
        __syncthreads();
        if (valid) {
            uint32_t o = running[d] + rank;
            for (unsigned w = 0; w < wave; ++w) {
                o += wcount[w][d];
            }
            const uint32_t pos = lbase[d] + o;
            skey[pos] = key;
            sval[pos] = val;
        }
        __syncthreads();
        if (leader) {
            atomicAdd(&running[d], group);
            wcount[wave][d] = 0;
        }
        __syncthreads();
    }

    // Staged order is digit-sorted: write each digit's run contiguously.
    const unsigned m = static_cast<unsigned>(n - tbase < static_cast<size_t>(kSortTile)
                                                 ? n - tbase
                                                 : static_cast<size_t>(kSortTile));
    for (unsigned idx = tid; idx < m; idx += kSortThreads) {
        const uint64_t key = skey[idx];
        const unsigned d = digit_of(key, shift);
        const uint32_t dst = gbase[d] + (idx - lbase[d]);
        kout[dst] = key;
        vout[dst] = sval[idx];
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

// One workgroup: each thread sums a contiguous chunk of the tile totals, one
// block scan of the chunk sums, then each thread writes its chunk's running
// offsets. (Measured on MI355X: 9,766 tiles at 10M particles took 42 us as
// one block scan per 256 totals.)
__global__ __launch_bounds__(kScanThreads) void scan_partials(uint32_t* __restrict__ partial,
                                                             unsigned count) {
    __shared__ uint32_t s[kScanThreads];
    const unsigned per = (count + kScanThreads - 1) / kScanThreads;
    const unsigned lo = min(count, threadIdx.x * per);
    const unsigned hi = min(count, lo + per);
    uint32_t sum = 0;
    for (unsigned i = lo; i < hi; ++i) {
        sum += partial[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, s, &total);
    for (unsigned i = lo; i < hi; ++i) {
        const uint32_t v = partial[i];
        partial[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) {
        partial[count] = total;
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
                                   Buf& hist, NBodySimProError* error, uint64_t* keys_out,
                                   uint32_t* vals_out, int first_bit) {
    if (keys_out == nullptr || vals_out == nullptr) {
        keys_out = keys;
        vals_out = vals;
    }
    if (n == 0) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    if (n > 0xFFFFFFFFull) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "sort size exceeds 2^32");
    }
    if (n == 1) {
        if (keys_out != keys) {
            NBX_HIP(error, hipMemcpyAsync(keys_out, keys, sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
            NBX_HIP(error, hipMemcpyAsync(vals_out, vals, sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        }
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    const bool large = n >= kLargeSort;
    const unsigned tiles = grid_for(n, kSortThreads * (large ? 8 : 4));
    NBX_TRY(hist.need<uint32_t>(static_cast<size_t>(kRadix) * tiles + kRadix, error));
    uint32_t* h = hist.as<uint32_t>();
    uint32_t* totals = h + static_cast<size_t>(kRadix) * tiles;
    const int passes = (bits + kDigitBits - 1) / kDigitBits;
    const uint64_t* kin = keys;
    const uint32_t* vin = vals;
    for (int p = 0; p < passes; ++p) {
        const int shift = first_bit + kDigitBits * p;
        // ping-pong between the two buffers; the last pass writes the output
        const bool last = p == passes - 1;
        uint64_t* kout = last ? keys_out : ((p & 1) ? keys : keys_alt);
        uint32_t* vout = last ? vals_out : ((p & 1) ? vals : vals_alt);
        if (last && (kout == kin || static_cast<const uint32_t*>(vout) == vin)) {
            kout = (kin == keys) ? keys_alt : keys;     // output aliases the input: one more hop
            vout = (vin == vals) ? vals_alt : vals;
        }
        if (large) {
            sort_histogram<8><<<tiles, kSortThreads, 0, s>>>(kin, n, shift, h, tiles);
        } else {
            sort_histogram<4><<<tiles, kSortThreads, 0, s>>>(kin, n, shift, h, tiles);
        }
        sort_scan_digit<<<kRadix, kSortThreads, 0, s>>>(h, tiles, totals);
        if (large) {
            sort_scatter<8><<<tiles, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, shift, h,
                                                           tiles, totals);
        } else {
            sort_scatter<4><<<tiles, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, shift, h,
                                                           tiles, totals);
        }
        kin = kout;
        vin = vout;
    }
    NBX_HIP(error, hipGetLastError());
    if (kin != keys_out) {
        NBX_HIP(error, hipMemcpyAsync(keys_out, kin, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        NBX_HIP(error, hipMemcpyAsync(vals_out, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
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
