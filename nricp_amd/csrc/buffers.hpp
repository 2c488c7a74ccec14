// Host-side ownership of device memory: one RAII buffer type, sized once and
// reused, plus typed upload/download on a stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_common.hpp"

namespace nricp_hip {

// Device memory goes through a process-wide, stream-ordered cache. hipFree
// waits for the whole device and would stall every other job on the GPU, so
// released blocks are kept for the next request of their size class. Each
// block belongs to the stream in scope when it was allocated (StreamScope):
// released, it goes back to that stream's free list and only allocations
// ordered on the same stream take it again, so work still queued there that
// uses the old contents runs before any new writer. When a Stream is
// destroyed it is synchronised first and its free list becomes idle, for any
// stream to take. A block allocated with no stream in scope waits for the
// device when released. A buffer that work on another stream also uses must
// outlive that work (the owner synchronises that stream first). Blocks are
// returned to the runtime only when an allocation would otherwise fail. The
// bytes held by the calling thread's live buffers are counted per thread (a
// worker drives one job at a time, so that is the job's footprint).
void* device_allocate(int64_t bytes, int64_t* granted, int* device, hipStream_t* stream);
void device_release(void* p, int64_t bytes, int device, hipStream_t stream);
int64_t thread_device_bytes();
// The calling thread's allocation stream; returns the previous one.
hipStream_t swap_allocation_stream(hipStream_t s);
// Stream lifetime, for the cache: created, and retired once synchronised.
void stream_created(hipStream_t s);
void stream_retired(hipStream_t s);
// Streams for sessions and engines, from a per-device pool of idle ones: a
// stream costs about 0.1 ms of runtime time to create and 0.4 ms to destroy,
// and the runtime serialises both across threads, while sessions come and go
// per job. stream_release takes a synchronised stream back (its cached blocks
// and persistent launches retired as at destruction) and destroys it only
// when the device's pool is full or closed.
hipStream_t stream_acquire();
void stream_release(hipStream_t s);
// A synchronised stream about to be destroyed: its grid-barrier launches
// leave the admission ledger (their events go with them).
void launches_retired(hipStream_t s);
// A synchronised stream that lives on: its grid-barrier launches leave the
// ledger and their events are kept for reuse.
void launches_completed(hipStream_t s);
// The device drained: the ledger's events are destroyed.
void launches_teardown();
// Drains every device, returns every cached block to the runtime and destroys
// the launch ledger's events; blocks released afterwards are left to the
// process's end. For callers about to tear the HIP runtime down; no other
// library call may run meanwhile.
void release_all();

// Orders the buffers the calling thread allocates in this scope on `s`.
class StreamScope {
public:
    explicit StreamScope(hipStream_t s) : saved_(swap_allocation_stream(s)) {}
    ~StreamScope() { swap_allocation_stream(saved_); }
    StreamScope(const StreamScope&) = delete;
    StreamScope& operator=(const StreamScope&) = delete;

private:
    hipStream_t saved_;
};

template <typename T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(int64_t n) { resize(n); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept
        : ptr_(o.ptr_), size_(o.size_), cap_(o.cap_), bytes_(o.bytes_), device_(o.device_), stream_(o.stream_) {
        o.ptr_ = nullptr;
        o.size_ = o.cap_ = o.bytes_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            release();
            ptr_ = o.ptr_;
            size_ = o.size_;
            cap_ = o.cap_;
            bytes_ = o.bytes_;
            device_ = o.device_;
            stream_ = o.stream_;
            o.ptr_ = nullptr;
            o.size_ = o.cap_ = o.bytes_ = 0;
        }
        return *this;
    }

    // Grows the allocation only when needed; contents are not preserved.
    void resize(int64_t n) {
        if (n > cap_) {
            release();
            const int64_t bytes = (n > 0 ? n : 1) * static_cast<int64_t>(sizeof(T));
            ptr_ = static_cast<T*>(device_allocate(bytes, &bytes_, &device_, &stream_));
            cap_ = n;
        }
        size_ = n;
    }
    void release() {
        if (ptr_) device_release(ptr_, bytes_, device_, stream_);
        ptr_ = nullptr;
        size_ = cap_ = 0;
        bytes_ = 0;
    }
    void upload(const T* host, int64_t n, hipStream_t s) {
        resize(n);
        if (n > 0) NRICP_CHECK(hipMemcpyAsync(ptr_, host, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& host, hipStream_t s) { upload(host.data(), static_cast<int64_t>(host.size()), s); }
    void download(T* host, int64_t n, hipStream_t s) const {
        if (n > 0) NRICP_CHECK(hipMemcpyAsync(host, ptr_, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    std::vector<T> to_host(hipStream_t s) const {
        std::vector<T> h(size_);
        download(h.data(), size_, s);
        NRICP_CHECK(hipStreamSynchronize(s));
        return h;
    }
    void zero(hipStream_t s) {
        if (size_ > 0) NRICP_CHECK(hipMemsetAsync(ptr_, 0, size_ * sizeof(T), s));
    }

    T* get() const { return ptr_; }
    int64_t size() const { return size_; }

private:
    T* ptr_ = nullptr;
    int64_t size_ = 0;
    int64_t cap_ = 0;
    int64_t bytes_ = 0;  // granted by the cache
    int device_ = 0;
    hipStream_t stream_ = nullptr;  // the stream the block is ordered on
};

// A stream owned by one object (default priority).
class Stream {
public:
    Stream() : s_(stream_acquire()) {}
    ~Stream() {
        if (s_) {
            (void)hipStreamSynchronize(s_);
            stream_release(s_);
        }
    }
    Stream(const Stream&) = delete;
    Stream& operator=(const Stream&) = delete;
    hipStream_t get() const { return s_; }
    void sync() const { NRICP_CHECK(hipStreamSynchronize(s_)); }

private:
    hipStream_t s_ = nullptr;
};

// Pinned host staging for small readbacks (energies, flags).
template <typename T>
class PinnedBuffer {
public:
    explicit PinnedBuffer(int64_t n) : n_(n) {
        NRICP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault));
    }
    ~PinnedBuffer() {
        if (p_) (void)hipHostFree(p_);
    }
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    T* get() const { return p_; }
    T& operator[](int64_t i) const { return p_[i]; }

private:
    T* p_ = nullptr;
    int64_t n_ = 0;
};

}  // namespace nricp_hip
