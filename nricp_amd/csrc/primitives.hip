// Stable LSD radix sort, exclusive scan, CSR offsets and run flags for gfx950.
//
// Sort: 8-bit digits, one launch per digit pass after one histogram launch
// per sort (a single-pass "onesweep" scheme: Merrill and Garland's decoupled
// look-back, applied per digit). The histogram launch counts every pass's
// digits in one read of the keys (the multiset of keys is the same in every
// pass). Each pass's launch takes its tiles in the order workgroups start
// (a ticket), counts its tile's digits, publishes them, and walks back over
// the tiles before it for each digit's running offset; then it scatters its
// keys stably (every wave ranks its keys with eight ballots, the lanes
// holding the same digit, and waves are ordered through LDS). A tile is 256
// threads x 8 rounds = 2048 keys.
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
constexpr int kSortHistogramBlocks = 256;

__device__ inline uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Every pass's digit counts in one read of the keys: per-workgroup counts in
// LDS, then one atomic add per nonzero bin into this sort's table (zeroed by
// the previous sort's histogram launch, as this launch's workgroup 0 zeroes
// the other table -- counts and tickets -- for the next sort).
__global__ void __launch_bounds__(kSortThreads)
sort_histogram(const uint64_t* __restrict__ keys, int64_t n, int passes, uint32_t last_mask,
               uint32_t* __restrict__ table, uint32_t* __restrict__ next_table) {
    __shared__ uint32_t h[kSortMaxPasses][kRadix];
    for (int p = 0; p < passes; ++p) h[p][threadIdx.x] = 0;
    if (blockIdx.x == 0)
        for (int k = threadIdx.x; k < kSortTableWords; k += kSortThreads) next_table[k] = 0;
    __syncthreads();
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kSortThreads + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kSortThreads) {
        const uint64_t key = keys[i];
        for (int p = 0; p < passes; ++p)
            atomicAdd(&h[p][(key >> (8 * p)) & (p == passes - 1 ? last_mask : 0xffu)], 1u);
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p) {
        const uint32_t c = h[p][threadIdx.x];
        if (c) atomicAdd(&table[p * kRadix + threadIdx.x], c);
    }
}

// Exclusive scan of one value per thread over a 256-thread block; returns
// the block total through `total`.
template <typename T>
__device__ inline T block_exclusive(T v, T* wave_totals, T& total);

// Look-back words, one per (tile, digit): the pass's epoch in the high 32
// bits (a word from any other pass or sort never matches), then a flag (the
// count is the digit's inclusive prefix up to this tile, else this tile's own
// count) and the count in the low 31 bits.
constexpr uint64_t kPrefixFlag = 1ull << 31;
constexpr uint64_t kCountMask = kPrefixFlag - 1;
// A wait on a tile that started earlier always ends (that tile is resident
// and waits only on tiles before it); the cap only bounds a wait that a
// malformed launch could leave open.
constexpr uint32_t kLookbackSpinCap = 1u << 20;

__device__ inline void publish(uint64_t* word, uint64_t value) {
    __hip_atomic_store(word, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(kSortThreads)
onesweep_scatter(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n, int shift,
                 uint32_t digit_mask, const uint32_t* __restrict__ counts, uint32_t* __restrict__ ticket,
                 uint64_t* __restrict__ state,
                 uint32_t epoch, uint64_t* __restrict__ out_keys, uint32_t* __restrict__ out_vals) {
    constexpr int waves = kSortThreads / 64;
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t tile_count[kRadix];
    __shared__ uint32_t wave_count[waves][kRadix];
    __shared__ uint32_t wave_totals[waves];
    __shared__ uint32_t tile_slot;
    const int d = threadIdx.x;
    if (d == 0) tile_slot = atomicAdd(ticket, 1u);
    tile_count[d] = 0;
    uint32_t total;
    uint32_t offset = block_exclusive(counts[d], wave_totals, total);  // digits below d, whole input
    __syncthreads();
    const int64_t tile = tile_slot;
    const int64_t tile_begin = tile * kTile;
    uint64_t key[kSortRounds];
    uint32_t val[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const int64_t i = tile_begin + static_cast<int64_t>(r) * kSortThreads + d;
        const bool valid = i < n;
        key[r] = valid ? keys[i] : 0ull;
        val[r] = (valid && vals) ? vals[i] : 0u;
        if (valid) atomicAdd(&tile_count[(key[r] >> shift) & digit_mask], 1u);
    }
    __syncthreads();
    // Thread d: digit d's offset among the tiles before this one.
    const uint32_t mine = tile_count[d];
    const uint64_t tag = static_cast<uint64_t>(epoch) << 32;
    uint64_t* word = state + tile * kRadix + d;
    if (tile == 0) {
        publish(word, tag | kPrefixFlag | mine);
    } else {
        publish(word, tag | mine);
        uint32_t before = 0;
        for (int64_t u = tile - 1; u >= 0; --u) {
            const uint64_t* w = state + u * kRadix + d;
            uint64_t s = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (uint32_t spin = 0; static_cast<uint32_t>(s >> 32) != epoch && spin < kLookbackSpinCap; ++spin) {
                __builtin_amdgcn_s_sleep(1);
                s = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            before += static_cast<uint32_t>(s & kCountMask);
            if (s & kPrefixFlag) break;
        }
        publish(word, tag | kPrefixFlag | (before + mine));
        offset += before;
    }
    base[d] = offset;
    const int wave = d / 64;
    const uint64_t below = lanemask_lt();
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        for (int w = 0; w < waves; ++w) wave_count[w][d] = 0;
        __syncthreads();
        const int64_t i = tile_begin + static_cast<int64_t>(r) * kSortThreads + d;
        const bool valid = i < n;
        const uint32_t digit = static_cast<uint32_t>((key[r] >> shift) & digit_mask);
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
            uint32_t ahead = 0;
            for (int w = 0; w < wave; ++w) ahead += wave_count[w][digit];
            const uint32_t pos = base[digit] + ahead + rank;
            out_keys[pos] = key[r];
            if (vals) out_vals[pos] = val[r];
        }
        __syncthreads();
        uint32_t round_total = 0;
        for (int w = 0; w < waves; ++w) round_total += wave_count[w][d];
        base[d] += round_total;
    }
}

__global__ void fill_u64(uint64_t* __restrict__ p, int64_t n, uint64_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
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

namespace {
struct KernelLaunch {
    const void* kernel;
    int groups, threads;
    void** args;
    hipStream_t stream;
    void (*prelude)(void*);
    void* prelude_arg;
    static hipError_t run(void* p) {
        const KernelLaunch& k = *static_cast<const KernelLaunch*>(p);
        if (k.prelude) k.prelude(k.prelude_arg);
        return hipLaunchKernel(k.kernel, dim3(k.groups), dim3(k.threads), k.args, 0, k.stream);
    }
};
}  // namespace

void launch_persistent(const void* kernel, int groups, int threads, void** args, hipStream_t stream,
                       void (*prelude)(void*), void* prelude_arg) {
    if (groups < 1 || groups > persistent_group_budget(kernel, threads))
        throw InvalidArgument("persistent launch of " + std::to_string(groups) + " groups exceeds the residency budget");
    KernelLaunch k{kernel, groups, threads, args, stream, prelude, prelude_arg};
    admit_persistent(groups, stream, &KernelLaunch::run, &k);
}

void admit_persistent(int groups, hipStream_t stream, hipError_t (*enqueue)(void*), void* arg) {
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
    // Enqueue and record under the lock, so the ledger never holds a charge
    // without its event.
    hipError_t e = hipSuccess;
    try {
        e = enqueue(arg);
    } catch (...) {
        r.spare.push_back(Spare{device, done, stream});
        throw;
    }
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
    if (n > static_cast<int64_t>(kCountMask)) throw InvalidArgument("radix sort: 2^31 keys or more");
    key_bits = std::min(key_bits, 64);
    const int passes = (key_bits + 7) / 8;
    // The last pass's digit holds only the key bits left (so bits above
    // key_bits never order anything).
    const uint32_t last_mask = (1u << (key_bits - 8 * (passes - 1))) - 1u;
    const int64_t tiles = (n + kTile - 1) / kTile;
    s.keys_alt.resize(n);
    if (vals) s.vals_alt.resize(n);
    if (s.tables.size() == 0) {
        // Both tables start at zero; from then on each sort's histogram
        // launch clears the table the next sort uses.
        s.tables.resize(2 * kSortTableWords);
        fill_u32<<<blocks_for(2 * kSortTableWords, 256), 256, 0, stream>>>(s.tables.get(), 2 * kSortTableWords, 0u);
        NRICP_LAUNCH_CHECK("fill_u32");
        s.parity = 0;
    }
    const int64_t words = tiles * kRadix;
    if (words > s.state_words) {
        // Fresh look-back words hold no epoch (epochs start at 1).
        s.state.resize(words);
        fill_u64<<<blocks_for(words, 256), 256, 0, stream>>>(s.state.get(), words, 0ull);
        NRICP_LAUNCH_CHECK("fill_u64");
        s.state_words = words;
    }
    uint32_t* table = s.tables.get() + s.parity * kSortTableWords;
    uint32_t* next = s.tables.get() + (1 - s.parity) * kSortTableWords;
    s.parity = 1 - s.parity;
    const int hist_blocks = static_cast<int>(std::min<int64_t>(tiles, kSortHistogramBlocks));
    sort_histogram<<<hist_blocks, kSortThreads, 0, stream>>>(keys, n, passes, last_mask, table, next);
    NRICP_LAUNCH_CHECK("sort_histogram");
    uint64_t* kin = keys;
    uint32_t* vin = vals;
    uint64_t* kout = s.keys_alt.get();
    uint32_t* vout = vals ? s.vals_alt.get() : nullptr;
    for (int p = 0; p < passes; ++p) {
        if (++s.epoch == 0) {
            // 2^32 passes on this scratch: clear the words before reusing epochs.
            fill_u64<<<blocks_for(s.state_words, 256), 256, 0, stream>>>(s.state.get(), s.state_words, 0ull);
            NRICP_LAUNCH_CHECK("fill_u64");
            s.epoch = 1;
        }
        onesweep_scatter<<<static_cast<unsigned>(tiles), kSortThreads, 0, stream>>>(
            kin, vin, n, 8 * p, p == passes - 1 ? last_mask : 0xffu, table + p * kRadix,
            table + kSortMaxPasses * kRadix + p, s.state.get(), s.epoch,
            kout, vout);
        NRICP_LAUNCH_CHECK("onesweep_scatter");
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
