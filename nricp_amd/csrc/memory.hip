// Process-wide, stream-ordered cache of device memory blocks behind
// DeviceBuffer.
//
// hipFree waits for the whole device, so a job that frees its buffers while
// other jobs run on the GPU would stall all of them. Released blocks are kept
// here per (device, stream, size class). A block belongs to the stream that
// was in scope for the allocating thread; once released, only an allocation
// ordered on that same stream takes it again, so every kernel queued on the
// stream before the release (the ones that still read or write the old
// contents) runs before the new owner's first kernel. Handing the block to
// another stream at once would let that stream's kernels overwrite it while
// the first stream's are still queued. A retired stream (synchronised, then
// destroyed) has no work left, so its blocks become idle and any stream may
// take them. Memory returns to the device only when an allocation would
// otherwise fail.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <set>
#include <tuple>
#include <vector>

#include "buffers.hpp"

namespace nricp_hip {
namespace {

// Powers of two up to 1 MiB, then whole MiB.
int64_t size_class(int64_t bytes) {
    if (bytes <= 256) return 256;
    constexpr int64_t kMiB = int64_t{1} << 20;
    if (bytes <= kMiB) {
        int64_t c = 256;
        while (c < bytes) c <<= 1;
        return c;
    }
    return (bytes + kMiB - 1) / kMiB * kMiB;
}

using Key = std::tuple<int, hipStream_t, int64_t>;  // (device, stream or idle, class)

struct Cache {
    std::mutex mutex;
    std::multimap<Key, void*> free;
    std::set<hipStream_t> live;  // streams created and not yet retired
};

// Never destroyed: buffers with static storage may be released after every
// other static object is gone.
Cache& cache() {
    static Cache* c = new Cache;
    return *c;
}

thread_local int64_t t_bytes = 0;
thread_local hipStream_t t_stream = nullptr;

// Takes a block of `cls` bytes for `stream` (its own free list first, then
// the idle one); nullptr if none is cached. Caller holds the mutex.
void* take(int device, hipStream_t stream, int64_t cls) {
    auto& f = cache().free;
    auto it = stream ? f.find(Key{device, stream, cls}) : f.end();
    if (it == f.end()) it = f.find(Key{device, nullptr, cls});
    if (it == f.end()) return nullptr;
    void* p = it->second;
    f.erase(it);
    return p;
}

// Returns every cached block of `device` to the runtime, after the device
// has drained (blocks on a stream's free list may still have work queued).
void trim(int device) {
    (void)hipDeviceSynchronize();
    std::lock_guard<std::mutex> lock(cache().mutex);
    for (auto it = cache().free.begin(); it != cache().free.end();) {
        if (std::get<0>(it->first) == device) {
            (void)hipFree(it->second);
            it = cache().free.erase(it);
        } else {
            ++it;
        }
    }
}

}  // namespace

void* device_allocate(int64_t bytes, int64_t* granted, int* device_out, hipStream_t* stream_out) {
    const int64_t cls = size_class(bytes);
    int device = 0;
    NRICP_CHECK(hipGetDevice(&device));
    const hipStream_t stream = t_stream;
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(cache().mutex);
        p = take(device, stream, cls);
    }
    if (!p) {
        hipError_t e = hipMalloc(&p, static_cast<size_t>(cls));
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            trim(device);
            e = hipMalloc(&p, static_cast<size_t>(cls));
        }
        check(e, "hipMalloc");
    }
    *granted = cls;
    *device_out = device;
    *stream_out = stream;
    t_bytes += cls;
    return p;
}

void device_release(void* p, int64_t bytes, int device, hipStream_t stream) {
    if (!p) return;
    t_bytes -= bytes;
    if (!stream) {
        // Ordered on no known stream: the block is idle once the device has
        // drained.
        int current = 0;
        (void)hipGetDevice(&current);
        if (current != device) (void)hipSetDevice(device);
        (void)hipDeviceSynchronize();
        if (current != device) (void)hipSetDevice(current);
    }
    std::lock_guard<std::mutex> lock(cache().mutex);
    // A stream retired since the allocation has finished all its work.
    const hipStream_t owner = stream && cache().live.count(stream) ? stream : nullptr;
    cache().free.emplace(Key{device, owner, bytes}, p);
}

int64_t thread_device_bytes() { return t_bytes; }

hipStream_t swap_allocation_stream(hipStream_t s) {
    const hipStream_t old = t_stream;
    t_stream = s;
    return old;
}

void stream_created(hipStream_t s) {
    std::lock_guard<std::mutex> lock(cache().mutex);
    cache().live.insert(s);
}

void stream_retired(hipStream_t s) {
    std::lock_guard<std::mutex> lock(cache().mutex);
    cache().live.erase(s);
    auto& f = cache().free;
    std::vector<std::pair<Key, void*>> moved;
    for (auto it = f.begin(); it != f.end();) {
        if (std::get<1>(it->first) == s) {
            moved.emplace_back(Key{std::get<0>(it->first), nullptr, std::get<2>(it->first)}, it->second);
            it = f.erase(it);
        } else {
            ++it;
        }
    }
    for (auto& m : moved) f.emplace(m.first, m.second);
}

}  // namespace nricp_hip
