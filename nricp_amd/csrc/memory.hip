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
// take them (the policy itself is StreamOrderedPool, stream_pool.hpp). Memory
// returns to the device when an allocation would otherwise fail, or when the
// caller asks for it (release_all, before it tears the runtime down).
#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
#include <vector>

#include "buffers.hpp"
#include "stream_pool.hpp"

namespace nricp_hip {
namespace {

using Pool = StreamOrderedPool<hipStream_t, void*>;

struct Cache {
    std::mutex mutex;
    Pool pool;
    std::atomic<bool> closed{false};  // after release_all: later releases are left to the process's end
};

// Never destroyed: buffers with static storage may be released after every
// other static object is gone.
Cache& cache() {
    static Cache* c = new Cache;
    return *c;
}

thread_local int64_t t_bytes = 0;
thread_local hipStream_t t_stream = nullptr;

// Returns every cached block of `device` to the runtime, after the device
// has drained (blocks on a stream's free list may still have work queued).
void trim(int device) {
    (void)hipDeviceSynchronize();
    std::lock_guard<std::mutex> lock(cache().mutex);
    cache().pool.drain(device, [](int, void* p) { (void)hipFree(p); });
}

}  // namespace

void* device_allocate(int64_t bytes, int64_t* granted, int* device_out, hipStream_t* stream_out) {
    const int64_t cls = Pool::size_class(bytes);
    int device = 0;
    NRICP_CHECK(hipGetDevice(&device));
    const hipStream_t stream = t_stream;
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(cache().mutex);
        if (!cache().closed && !cache().pool.take(device, stream, cls, &p)) p = nullptr;
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
    if (cache().closed) return;  // after release_all: the process's end frees it
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
    cache().pool.release(device, stream, bytes, p);
}

int64_t thread_device_bytes() { return t_bytes; }

hipStream_t swap_allocation_stream(hipStream_t s) {
    const hipStream_t old = t_stream;
    t_stream = s;
    return old;
}

void stream_created(hipStream_t s) {
    std::lock_guard<std::mutex> lock(cache().mutex);
    cache().pool.stream_created(s);
}

void stream_retired(hipStream_t s) {
    std::lock_guard<std::mutex> lock(cache().mutex);
    cache().pool.stream_retired(s);
}

namespace {
struct IdleStreams {
    std::mutex mutex;
    std::map<int, std::vector<hipStream_t>> idle;  // device -> synchronised streams
    bool closed = false;
};
IdleStreams& idle_streams() {
    static IdleStreams* s = new IdleStreams;  // never destroyed: used until exit
    return *s;
}
constexpr size_t kIdleStreamsPerDevice = 64;
}  // namespace

hipStream_t stream_acquire() {
    int device = 0;
    NRICP_CHECK(hipGetDevice(&device));
    hipStream_t s = nullptr;
    {
        IdleStreams& pool = idle_streams();
        std::lock_guard<std::mutex> lock(pool.mutex);
        auto& v = pool.idle[device];
        if (!v.empty()) {
            s = v.back();
            v.pop_back();
        }
    }
    if (!s) NRICP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    stream_created(s);
    return s;
}

void stream_release(hipStream_t s) {
    stream_retired(s);
    int device = -1;
    if (hipStreamGetDevice(s, &device) != hipSuccess) {
        (void)hipGetLastError();
        device = -1;
    }
    IdleStreams& pool = idle_streams();
    bool room = false;
    if (device >= 0) {
        std::lock_guard<std::mutex> lock(pool.mutex);
        room = !pool.closed && pool.idle[device].size() < kIdleStreamsPerDevice;
    }
    if (room) {
        // Synchronised: its persistent launches leave the admission ledger
        // (their events kept for reuse) before another owner can take it and
        // admit launches of its own on it.
        launches_completed(s);
        std::lock_guard<std::mutex> lock(pool.mutex);
        if (!pool.closed) {
            pool.idle[device].push_back(s);
            return;
        }
    }
    launches_retired(s);
    (void)hipStreamDestroy(s);
}

void release_all() {
    Cache& c = cache();
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
    int current = 0;
    (void)hipGetDevice(&current);
    for (int d = 0; d < count; ++d) {
        if (hipSetDevice(d) == hipSuccess) (void)hipDeviceSynchronize();
    }
    {
        std::lock_guard<std::mutex> lock(c.mutex);
        c.closed = true;
        c.pool.drain(-1, [](int device, void* p) {
            (void)hipSetDevice(device);
            (void)hipFree(p);
        });
    }
    (void)hipSetDevice(current);
    {
        IdleStreams& pool = idle_streams();
        std::lock_guard<std::mutex> lock(pool.mutex);
        pool.closed = true;
        for (auto& kv : pool.idle)
            for (hipStream_t s : kv.second) {
                launches_retired(s);
                (void)hipStreamDestroy(s);
            }
        pool.idle.clear();
    }
    launches_teardown();
}

}  // namespace nricp_hip
