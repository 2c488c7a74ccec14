// Host-side ownership of device memory: one RAII buffer type, sized once and
// reused, plus typed upload/download on a stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_common.hpp"

namespace nricp_hip {

// Device memory goes through a process-wide cache: a released block is kept
// for the next request of its size class instead of being returned with
// hipFree, which waits for the whole device and would stall every other job
// on the GPU. Blocks are never returned before the process ends; the bytes
// held by the calling thread's live buffers are counted per thread (a
// worker drives one job at a time, so that is the job's footprint).
void* device_allocate(int64_t bytes, int64_t* granted, int* device);
void device_release(void* p, int64_t bytes, int device);
int64_t thread_device_bytes();

template <typename T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(int64_t n) { resize(n); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept
        : ptr_(o.ptr_), size_(o.size_), cap_(o.cap_), bytes_(o.bytes_), device_(o.device_) {
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
            ptr_ = static_cast<T*>(device_allocate(bytes, &bytes_, &device_));
            cap_ = n;
        }
        size_ = n;
    }
    void release() {
        if (ptr_) device_release(ptr_, bytes_, device_);
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
};

// A stream owned by one object (default priority).
class Stream {
public:
    Stream() { NRICP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }
    ~Stream() {
        if (s_) (void)hipStreamDestroy(s_);
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
