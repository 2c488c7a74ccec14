// Stable LSD radix sort, exclusive scan, CSR offsets and run flags for gfx950.
//
// Sort: 8-bit digits. Each pass is a per-tile digit histogram in LDS, one
// exclusive scan over the digit-major table of tile counts, and a stable
// scatter in which every wave ranks its keys with eight
// ballots (the lanes holding the same digit) and waves are ordered through LDS.
// A tile is 256 threads x 8 rounds = 2048 keys.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

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

// ------------------------------------------------------------------ scan --
//
// Exclusive scan in three launches: per-tile sums, a scan of the tile sums
// (one workgroup, chunked with a carry), and a per-tile scan that adds the
// tile's offset. A tile is 256 threads x 16 items, loaded striped (coalesced)
// into LDS and scanned in contiguous runs of 16 per thread. Every sum has a
// fixed order, and each tile is read whole before it is written, so the scan
// may run in place.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

// Exclusive scan of one value per thread over the block; returns the block
// total through `total`.
template <typename T>
__device__ inline T block_exclusive(T v, T* wave_totals, T& total) {
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    T incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wave_totals[wave] = incl;
    __syncthreads();
    T before = 0;
    total = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) {
        if (w < wave) before += wave_totals[w];
        total += wave_totals[w];
    }
    __syncthreads();
    return before + incl - v;
}

template <typename T>
__global__ void __launch_bounds__(kScanThreads) scan_tile_sums(const T* __restrict__ in, int64_t n,
                                                               T* __restrict__ partials) {
    __shared__ T wave_totals[kScanThreads / 64];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    T sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k * kScanThreads + threadIdx.x;
        if (i < n) sum += in[i];
    }
    T total;
    block_exclusive(sum, wave_totals, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// In-place exclusive scan of the tile sums by one workgroup.
template <typename T>
__global__ void __launch_bounds__(kScanThreads) scan_partials(T* __restrict__ partials, int64_t tiles) {
    __shared__ T wave_totals[kScanThreads / 64];
    T carry = 0;
    for (int64_t base = 0; base < tiles; base += kScanThreads) {
        const int64_t i = base + threadIdx.x;
        const T v = i < tiles ? partials[i] : T(0);
        T total;
        const T ex = block_exclusive(v, wave_totals, total);
        if (i < tiles) partials[i] = carry + ex;
        carry += total;
    }
}

template <typename T>
__global__ void __launch_bounds__(kScanThreads)
scan_tiles(const T* in, T* out, int64_t n, const T* __restrict__ partials, T* total_out) {
    __shared__ T tile[kScanTile];
    __shared__ T wave_totals[kScanThreads / 64];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k * kScanThreads + threadIdx.x;
        tile[k * kScanThreads + threadIdx.x] = i < n ? in[i] : T(0);
    }
    __syncthreads();
    T run[kScanItems];
    T sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        run[k] = sum;
        sum += tile[threadIdx.x * kScanItems + k];
    }
    T total;
    const T offset = partials[blockIdx.x] + block_exclusive(sum, wave_totals, total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) tile[threadIdx.x * kScanItems + k] = offset + run[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k * kScanThreads + threadIdx.x;
        if (i < n) out[i] = tile[k * kScanThreads + threadIdx.x];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total_out = partials[blockIdx.x] + total;
}

template <typename T>
void scan_exclusive(const T* in, T* out, int64_t n, T* total_out, DeviceBuffer<int64_t>& scratch,
                    hipStream_t stream) {
    if (n <= 0) {
        if (total_out) NRICP_CHECK(hipMemsetAsync(total_out, 0, sizeof(T), stream));
        return;
    }
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    scratch.resize((tiles * static_cast<int64_t>(sizeof(T)) + 7) / 8);
    T* partials = reinterpret_cast<T*>(scratch.get());
    scan_tile_sums<T><<<static_cast<unsigned>(tiles), kScanThreads, 0, stream>>>(in, n, partials);
    NRICP_LAUNCH_CHECK("scan_tile_sums");
    scan_partials<T><<<1, kScanThreads, 0, stream>>>(partials, tiles);
    NRICP_LAUNCH_CHECK("scan_partials");
    scan_tiles<T><<<static_cast<unsigned>(tiles), kScanThreads, 0, stream>>>(in, out, n, partials, total_out);
    NRICP_LAUNCH_CHECK("scan_tiles");
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

namespace {

// Grid-barrier launches in flight per device, counted in compute units (a
// persistent group holds one). Several jobs can run on one device at once
// (each on its own streams); admitting a persistent launch only while the
// launches in flight leave room for all of its groups keeps every admitted
// launch fully resident once the ordinary kernels ahead of it retire, so no
// two grid barriers can wait on each other's compute units. The dispatcher
// deals a launch's groups to the XCDs in turn, so a launch of g groups needs
// up to ceil(g / xcds) compute units on every XCD: a launch is charged
// ceil(g / xcds) * xcds units, which keeps each XCD's share within its own
// compute units whatever XCD the launch starts on.
//
// The ledger: launches admitted and not yet known to be finished, each
// with the event recorded after it. Charges return when the events complete,
// found by querying them at the next admission (an event costs the stream
// nothing, where a host callback would hold its later work back until the
// host had run it).
struct Admitted {
    int device, charge;
    hipEvent_t done;
    hipStream_t stream;  // the launch's stream, where `done` is recorded
};

struct Spare {
    int device;
    hipEvent_t event;
    hipStream_t stream;  // where the event was last recorded
};

// An event is queried only while the stream it was last recorded on is
// alive: before a stream is destroyed (synchronised first), its launches
// leave the ledger and their events are destroyed (launches_retired).
struct Residency {
    std::mutex mutex;
    std::map<int, int> used;         // device -> compute units charged
    std::deque<Admitted> in_flight;  // admission order
    std::vector<Spare> spare;        // events of finished launches, for reuse
};

Residency& residency() {
    static Residency* r = new Residency;  // never destroyed: used until exit
    return *r;
}

bool finished(hipEvent_t e) {
    const hipError_t q = hipEventQuery(e);
    if (q == hipSuccess) return true;
    // Not ready is no error; any other status surfaces on the job's own
    // stream. Either way leave no trace for the thread's next launch check.
    (void)hipGetLastError();
    return q != hipErrorNotReady;
}

// Returns the charges of every finished launch. Caller holds the mutex.
void reclaim(Residency& r) {
    for (auto it = r.in_flight.begin(); it != r.in_flight.end();) {
        if (finished(it->done)) {
            r.used[it->device] -= it->charge;
            r.spare.push_back(Spare{it->device, it->done, it->stream});
            it = r.in_flight.erase(it);
        } else {
            ++it;
        }
    }
}

hipEvent_t take_event(Residency& r, int device) {
    for (auto it = r.spare.begin(); it != r.spare.end(); ++it) {
        if (it->device == device) {
            const hipEvent_t e = it->event;
            r.spare.erase(it);
            return e;
        }
    }
    hipEvent_t e = nullptr;
    NRICP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
}

struct DeviceShape {
    int cus = 0, xcds = 1;
};

// Compute units and XCDs of the current device, queried once per device
// (the queries cost tens of microseconds, paid before every solve's launch).
DeviceShape device_shape(int* device_out) {
    int device = 0;
    NRICP_CHECK(hipGetDevice(&device));
    *device_out = device;
    static std::mutex mutex;
    static std::map<int, DeviceShape>* known = new std::map<int, DeviceShape>;
    std::lock_guard<std::mutex> lock(mutex);
    auto it = known->find(device);
    if (it != known->end()) return it->second;
    DeviceShape d;
    NRICP_CHECK(hipDeviceGetAttribute(&d.cus, hipDeviceAttributeMultiprocessorCount, device));
    if (hipDeviceGetAttribute(&d.xcds, hipDeviceAttributeNumberOfXccs, device) != hipSuccess || d.xcds < 1) {
        (void)hipGetLastError();
        d.xcds = 1;
    }
    (*known)[device] = d;
    return d;
}

}  // namespace

int persistent_group_budget(const void* kernel, int threads) {
    int device = 0;
    const int cus = device_shape(&device).cus;
    // Whether one group fits a compute unit, checked once per (device,
    // kernel, block size).
    static std::mutex mutex;
    static std::set<std::tuple<int, const void*, int>>* fits = new std::set<std::tuple<int, const void*, int>>;
    {
        std::lock_guard<std::mutex> lock(mutex);
        if (!fits->count({device, kernel, threads})) {
            int per_cu = 0;
            NRICP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
            if (per_cu < 1) throw DeviceError("a persistent kernel does not fit on a compute unit");
            fits->insert({device, kernel, threads});
        }
    }
    return std::max(1, cus / 2);
}

void launch_persistent(const void* kernel, int groups, int threads, void** args, hipStream_t stream) {
    if (groups < 1 || groups > persistent_group_budget(kernel, threads))
        throw InvalidArgument("persistent launch of " + std::to_string(groups) + " groups exceeds the residency budget");
    int device = 0;
    const DeviceShape shape = device_shape(&device);
    const int cus = shape.cus, xcds = shape.xcds;
    const int charge = std::min(cus, (groups + xcds - 1) / xcds * xcds);
    Residency& r = residency();
    std::unique_lock<std::mutex> lock(r.mutex);
    // Until the charge fits: poll every admitted launch's event (launches
    // of different jobs finish in no particular order), yielding the core
    // between polls at first and then sleeping.
    for (int polls = 0;; ++polls) {
        reclaim(r);
        if (r.used[device] + charge <= cus) break;
        bool any = false;
        for (const Admitted& a : r.in_flight) any = any || a.device == device;
        if (!any) throw DeviceError("persistent launch ledger out of step");
        lock.unlock();
        if (polls < 512)
            std::this_thread::yield();
        else
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        lock.lock();
    }
    const hipEvent_t done = take_event(r, device);
    // Launch and record under the lock, so the ledger never holds a charge
    // without its event.
    hipError_t e = hipLaunchKernel(kernel, dim3(groups), dim3(threads), args, 0, stream);
    if (e == hipSuccess) e = hipEventRecord(done, stream);
    if (e != hipSuccess) {
        r.spare.push_back(Spare{device, done, stream});
        lock.unlock();
        (void)hipStreamSynchronize(stream);
        check(e, "persistent launch");
    }
    r.used[device] += charge;
    r.in_flight.push_back(Admitted{device, charge, done, stream});
}

void launches_teardown() {
    Residency& r = residency();
    std::lock_guard<std::mutex> lock(r.mutex);
    for (const Admitted& a : r.in_flight) (void)hipEventDestroy(a.done);
    for (const Spare& sp : r.spare) (void)hipEventDestroy(sp.event);
    r.in_flight.clear();
    r.spare.clear();
    r.used.clear();
}

void launches_completed(hipStream_t stream) {
    Residency& r = residency();
    std::lock_guard<std::mutex> lock(r.mutex);
    // The caller has synchronised the stream: its launches are complete, and
    // their events go back to the spares for the next admissions.
    for (auto it = r.in_flight.begin(); it != r.in_flight.end();) {
        if (it->stream == stream) {
            r.used[it->device] -= it->charge;
            r.spare.push_back(Spare{it->device, it->done, it->stream});
            it = r.in_flight.erase(it);
        } else {
            ++it;
        }
    }
}

void launches_retired(hipStream_t stream) {
    Residency& r = residency();
    std::lock_guard<std::mutex> lock(r.mutex);
    // The stream has drained: its launches are complete.
    for (auto it = r.in_flight.begin(); it != r.in_flight.end();) {
        if (it->stream == stream) {
            r.used[it->device] -= it->charge;
            (void)hipEventDestroy(it->done);
            it = r.in_flight.erase(it);
        } else {
            ++it;
        }
    }
    for (auto it = r.spare.begin(); it != r.spare.end();) {
        if (it->stream == stream) {
            (void)hipEventDestroy(it->event);
            it = r.spare.erase(it);
        } else {
            ++it;
        }
    }
}

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
        scan_exclusive<uint32_t>(s.counts.get(), s.counts.get(), static_cast<int64_t>(kRadix) * tiles, nullptr,
                                 s.scan_partials, stream);
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

void exclusive_scan(const int64_t* in, int64_t* out, int64_t n, SortScratch& s, hipStream_t stream) {
    scan_exclusive<int64_t>(in, out, n, out + n, s.scan_partials, stream);
}

void exclusive_scan32(const int32_t* in, int32_t* out, int64_t n, SortScratch& s, hipStream_t stream) {
    scan_exclusive<int32_t>(in, out, n, out + n, s.scan_partials, stream);
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
