// Process-wide cache of device memory blocks behind DeviceBuffer.
//
// hipFree waits for the whole device, so a job that frees its buffers while
// other jobs run on the GPU would stall all of them. Released blocks are kept
// here per (device, size class) and handed to the next request of that class;
// memory returns to the device only when an allocation would otherwise fail.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

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

struct Cache {
    std::mutex mutex;
    std::multimap<std::pair<int, int64_t>, void*> free;  // (device, class) -> block
};

// Never destroyed: buffers with static storage may be released after every
// other static object is gone.
Cache& cache() {
    static Cache* c = new Cache;
    return *c;
}

thread_local int64_t t_bytes = 0;

// Returns every cached block of `device` to the runtime (waits for the device).
void trim(int device) {
    std::lock_guard<std::mutex> lock(cache().mutex);
    for (auto it = cache().free.begin(); it != cache().free.end();) {
        if (it->first.first == device) {
            (void)hipFree(it->second);
            it = cache().free.erase(it);
        } else {
            ++it;
        }
    }
}

}  // namespace

void* device_allocate(int64_t bytes, int64_t* granted, int* device_out) {
    const int64_t cls = size_class(bytes);
    int device = 0;
    NRICP_CHECK(hipGetDevice(&device));
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(cache().mutex);
        auto it = cache().free.find({device, cls});
        if (it != cache().free.end()) {
            p = it->second;
            cache().free.erase(it);
        }
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
    t_bytes += cls;
    return p;
}

void device_release(void* p, int64_t bytes, int device) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lock(cache().mutex);
        cache().free.emplace(std::make_pair(device, bytes), p);
    }
    t_bytes -= bytes;
}

int64_t thread_device_bytes() { return t_bytes; }

}  // namespace nricp_hip
