// Stable LSD radix sort, exclusive scan, CSR offsets and run flags for gfx950.
//
// Sort: 8-bit digits. Each pass is three launches -- a per-tile digit
// histogram in LDS, one exclusive scan over the digit-major table of tile
// counts, and a stable scatter in which every wave ranks its keys with eight
// ballots (the lanes holding the same digit) and waves are ordered through LDS.
// A tile is 256 threads x 8 rounds = 2048 keys.
#include <hip/hip_runtime.h>

#include "primitives.hpp"

namespace nricp_hip {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kTile = kSortThreads * kSortRounds;
constexpr int kRadix = 256;

__global__ void __launch_bounds__(kSortThreads)
digit_histogram(const uint64_t* __restrict__ keys, int64_t n, int shift, uint32_t* __restrict__ counts,
                int tiles) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t begin = static_cast<int64_t>(blockIdx.x) * kTile;
    const int64_t end = begin + kTile < n ? begin + kTile : n;
    for (int64_t i = begin + threadIdx.x; i < end; i += kSortThreads) {
        atomicAdd(&h[(keys[i] >> shift) & 0xff], 1u);
    }
    __syncthreads();
    counts[static_cast<int64_t>(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of n uint32 in place, one workgroup of 1024 threads: each
// thread scans a contiguous chunk, the chunk totals are scanned in LDS.
__global__ void __launch_bounds__(1024) scan_u32_inplace(uint32_t* data, int64_t n) {
    __shared__ uint32_t totals[1024];
    const int64_t chunk = (n + 1023) / 1024;
    const int64_t begin = threadIdx.x * chunk;
    const int64_t end = begin + chunk < n ? begin + chunk : n;
    uint32_t sum = 0;
    for (int64_t i = begin; i < end; ++i) sum += data[i];
    totals[threadIdx.x] = sum;
    __syncthreads();
    for (int offset = 1; offset < 1024; offset <<= 1) {
        const uint32_t v = threadIdx.x >= offset ? totals[threadIdx.x - offset] : 0u;
        __syncthreads();
        totals[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x > 0 ? totals[threadIdx.x - 1] : 0u;
    for (int64_t i = begin; i < end; ++i) {
        const uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

__device__ inline uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__global__ void __launch_bounds__(kSortThreads)
digit_scatter(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n, int shift,
              const uint32_t* __restrict__ offsets, int tiles, uint64_t* __restrict__ out_keys,
              uint32_t* __restrict__ out_vals) {
    constexpr int waves = kSortThreads / 64;
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t wave_count[waves][kRadix];
    base[threadIdx.x] = offsets[static_cast<int64_t>(threadIdx.x) * tiles + blockIdx.x];
    const int wave = threadIdx.x / 64;
    const uint64_t below = lanemask_lt();
    const int64_t tile_begin = static_cast<int64_t>(blockIdx.x) * kTile;
    for (int round = 0; round < kSortRounds; ++round) {
        for (int w = 0; w < waves; ++w) wave_count[w][threadIdx.x] = 0;
        __syncthreads();
        const int64_t i = tile_begin + static_cast<int64_t>(round) * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        const uint64_t key = valid ? keys[i] : 0ull;
        const uint32_t val = (valid && vals) ? vals[i] : 0u;
        const uint32_t digit = static_cast<uint32_t>((key >> shift) & 0xff);
        uint64_t same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (digit >> b) & 1u;
            const uint64_t set = __ballot(bit);
            same &= bit ? set : ~set;
        }
        const uint32_t rank = static_cast<uint32_t>(__popcll(same & below));
        if (valid && rank == 0) wave_count[wave][digit] = static_cast<uint32_t>(__popcll(same));
        __syncthreads();
        if (valid) {
            uint32_t before = 0;
            for (int w = 0; w < wave; ++w) before += wave_count[w][digit];
            const uint32_t pos = base[digit] + before + rank;
            out_keys[pos] = key;
            if (vals) out_vals[pos] = val;
        }
        __syncthreads();
        uint32_t total = 0;
        for (int w = 0; w < waves; ++w) total += wave_count[w][threadIdx.x];
        base[threadIdx.x] += total;
        __syncthreads();
    }
}

__global__ void copy_pairs(const uint64_t* __restrict__ k, const uint32_t* __restrict__ v, int64_t n,
                           uint64_t* __restrict__ ko, uint32_t* __restrict__ vo) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ko[i] = k[i];
    if (v) vo[i] = v[i];
}

template <typename T>
__global__ void __launch_bounds__(1024) scan_exclusive(const T* __restrict__ in, T* __restrict__ out, int64_t n) {
    __shared__ T totals[1024];
    const int64_t chunk = (n + 1023) / 1024;
    const int64_t begin = threadIdx.x * chunk;
    const int64_t end = begin + chunk < n ? begin + chunk : n;
    T sum = 0;
    for (int64_t i = begin; i < end; ++i) sum += in[i];
    totals[threadIdx.x] = sum;
    __syncthreads();
    for (int offset = 1; offset < 1024; offset <<= 1) {
        const T v = threadIdx.x >= offset ? totals[threadIdx.x - offset] : T(0);
        __syncthreads();
        totals[threadIdx.x] += v;
        __syncthreads();
    }
    T run = threadIdx.x > 0 ? totals[threadIdx.x - 1] : T(0);
    for (int64_t i = begin; i < end; ++i) {
        const T v = in[i];
        out[i] = run;
        run += v;
    }
    if (threadIdx.x == 1023) out[n] = totals[1023];
}

__global__ void sorted_offsets(const uint64_t* __restrict__ keys, int64_t n, int shift, uint64_t mask,
                               int64_t rows, int32_t* __restrict__ offsets) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r > rows) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (static_cast<int64_t>((keys[mid] >> shift) & mask) < r)
            lo = mid + 1;
        else
            hi = mid;
    }
    offsets[r] = static_cast<int32_t>(lo);
}

__global__ void run_start_flags(const uint64_t* __restrict__ keys, int64_t n, int32_t* __restrict__ flags) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

}  // namespace

int bits_for(uint64_t v) {
    int b = 0;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

void radix_sort_pairs(uint64_t* keys, uint32_t* vals, int64_t n, int key_bits, SortScratch& s,
                      hipStream_t stream) {
    if (n <= 1 || key_bits <= 0) return;
    if (n > 0xffffffffLL) throw InvalidArgument("radix sort: more than 2^32 keys");
    const int tiles = blocks_for(n, kTile);
    s.keys_alt.resize(n);
    if (vals) s.vals_alt.resize(n);
    s.counts.resize(static_cast<int64_t>(kRadix) * tiles);
    uint64_t* kin = keys;
    uint32_t* vin = vals;
    uint64_t* kout = s.keys_alt.get();
    uint32_t* vout = vals ? s.vals_alt.get() : nullptr;
    const int passes = (key_bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        digit_histogram<<<tiles, kSortThreads, 0, stream>>>(kin, n, shift, s.counts.get(), tiles);
        NRICP_LAUNCH_CHECK("digit_histogram");
        scan_u32_inplace<<<1, 1024, 0, stream>>>(s.counts.get(), static_cast<int64_t>(kRadix) * tiles);
        NRICP_LAUNCH_CHECK("scan_u32_inplace");
        digit_scatter<<<tiles, kSortThreads, 0, stream>>>(kin, vin, n, shift, s.counts.get(), tiles, kout, vout);
        NRICP_LAUNCH_CHECK("digit_scatter");
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys) {
        copy_pairs<<<blocks_for(n, 256), 256, 0, stream>>>(kin, vin, n, keys, vals);
        NRICP_LAUNCH_CHECK("copy_pairs");
    }
}

void exclusive_scan(const int64_t* in, int64_t* out, int64_t n, hipStream_t stream) {
    scan_exclusive<int64_t><<<1, 1024, 0, stream>>>(in, out, n);
    NRICP_LAUNCH_CHECK("scan_exclusive64");
}

void exclusive_scan32(const int32_t* in, int32_t* out, int64_t n, hipStream_t stream) {
    scan_exclusive<int32_t><<<1, 1024, 0, stream>>>(in, out, n);
    NRICP_LAUNCH_CHECK("scan_exclusive32");
}

void offsets_from_sorted(const uint64_t* keys, int64_t n, int shift, uint64_t mask, int64_t rows,
                         int32_t* offsets, hipStream_t stream) {
    sorted_offsets<<<blocks_for(rows + 1, 256), 256, 0, stream>>>(keys, n, shift, mask, rows, offsets);
    NRICP_LAUNCH_CHECK("sorted_offsets");
}

void run_starts(const uint64_t* keys, int64_t n, int32_t* flags, hipStream_t stream) {
    if (n <= 0) return;
    run_start_flags<<<blocks_for(n, 256), 256, 0, stream>>>(keys, n, flags);
    NRICP_LAUNCH_CHECK("run_start_flags");
}

}  // namespace nricp_hip
