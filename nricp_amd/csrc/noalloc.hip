// The damped block solve without device allocation, for callers that cannot
// allocate (code under graph capture, tight serving loops, harnesses that
// forbid hipMalloc): the system is already on the device in block-sparse row
// form, and every workspace vector of the solve is carved from one
// file-scope array, `scratch`, that each GPU holds for the life of the
// process. It is the solve of nricp_hip_solve_system with the host-side
// conversion left to the caller.
//
// Calls on one GPU share that GPU's array: calls that overlap on one GPU race
// (the entry point is synchronous, so only calls from different threads can
// overlap). Nothing in the array needs to start at zero: the grid-barrier
// counter is cleared by the solve's first kernel, and every other word is
// written before it is read.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "../../include/nricp_hip.h"
#include "kernels.hpp"

namespace {

constexpr int64_t kMaxNodes = NRICP_HIP_NOALLOC_MAX_NODES;
constexpr int64_t kMaxBlocks = NRICP_HIP_NOALLOC_MAX_BLOCKS;
constexpr int64_t kMaxGroups = 512;  // groups of one persistent solve

constexpr size_t round256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Workspace of the largest solve in each form (bytes, every piece 256-aligned).
constexpr size_t kTail = round256(6 * kMaxGroups * 8) + round256(2 * 4) + round256(4 * 8);  // slots, barrier, outcome
constexpr size_t kDoubleBytes = round256(6 * kMaxNodes * 8) + round256(36 * kMaxNodes * 8) +
                                round256(12 * kMaxNodes * 8) + round256(6 * kMaxNodes * 8) + kTail;
constexpr size_t kMixedBytes = round256(36 * kMaxBlocks * 4) + round256(6 * kMaxNodes * 4) +
                               round256(6 * kMaxNodes * 8) + round256(36 * kMaxNodes * 4) +
                               round256(12 * kMaxNodes * 4) + round256(6 * kMaxNodes * 8) + kTail;
constexpr size_t kScratchBytes = kDoubleBytes > kMixedBytes ? kDoubleBytes : kMixedBytes;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace nricp_hip {
namespace {

// The array's device address on the current device, looked up once per
// device with that device current (the name itself is a host-side stand-in).
unsigned char* scratch_on_current_device() {
    static std::mutex mutex;
    static std::map<int, unsigned char*>* known = new std::map<int, unsigned char*>;
    int device = 0;
    NRICP_CHECK(hipGetDevice(&device));
    std::lock_guard<std::mutex> lock(mutex);
    auto it = known->find(device);
    if (it != known->end()) return it->second;
    void* p = nullptr;
    NRICP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw DeviceError("scratch array has no device address");
    (*known)[device] = static_cast<unsigned char*>(p);
    return static_cast<unsigned char*>(p);
}

__global__ void round_blocks(const double* __restrict__ in, int64_t n, float* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<float>(in[i]);
}

struct Carver {
    unsigned char* p;
    template <typename T>
    T* take(int64_t count) {
        T* out = reinterpret_cast<T*>(p);
        p += round256(static_cast<size_t>(count) * sizeof(T));
        return out;
    }
};

void solve(const int32_t* row_off, const int32_t* cols, const int32_t* diag_pos, int64_t nn, int64_t nnzb,
           const double* blocks, const double* rhs, double damping, bool scaling, int precision, double tol,
           int max_iter, double* delta, double* info, hipStream_t stream) {
    if (nn < 1 || nn > kMaxNodes)
        throw InvalidArgument("solve_bsr_noalloc: node_count must be in [1, " + std::to_string(kMaxNodes) + "], not " +
                              std::to_string(nn));
    if (nnzb < nn || nnzb > kMaxBlocks)
        throw InvalidArgument("solve_bsr_noalloc: block_count must be in [node_count, " +
                              std::to_string(kMaxBlocks) + "], not " + std::to_string(nnzb));
    if (!row_off || !cols || !diag_pos || !blocks || !rhs || !delta || !info)
        throw InvalidArgument("solve_bsr_noalloc: null array");
    if (precision != NRICP_HIP_PRECISION_DOUBLE && precision != NRICP_HIP_PRECISION_MIXED &&
        precision != NRICP_HIP_PRECISION_FLOAT)
        throw InvalidArgument("solve_bsr_noalloc: unknown precision");
    if (max_iter < 1) throw InvalidArgument("solve_bsr_noalloc: max_iterations must be at least 1");
    if (pcg_workspace_groups(nn, precision != NRICP_HIP_PRECISION_DOUBLE) > kMaxGroups)
        throw InvalidArgument("solve_bsr_noalloc: the system needs more groups than the workspace holds");
    Carver c{scratch_on_current_device()};
    PcgWorkspaceView w{};
    const float* bsr32 = nullptr;
    if (precision == NRICP_HIP_PRECISION_DOUBLE) {
        w.damp = c.take<double>(6 * kMaxNodes);
        w.minv = c.take<double>(36 * kMaxNodes);
        w.vectors = c.take<double>(12 * kMaxNodes);
    } else {
        float* b32 = c.take<float>(36 * kMaxBlocks);
        round_blocks<<<blocks_for(36 * nnzb, 256), 256, 0, stream>>>(blocks, 36 * nnzb, b32);
        NRICP_LAUNCH_CHECK("round_blocks");
        bsr32 = b32;
        w.damp = c.take<float>(6 * kMaxNodes);
        w.damp64 = c.take<double>(6 * kMaxNodes);
        w.minv = c.take<float>(36 * kMaxNodes);
        w.vectors = c.take<float>(12 * kMaxNodes);
    }
    w.xbuf = c.take<double>(6 * kMaxNodes);
    w.slots = c.take<double>(6 * kMaxGroups);
    w.barrier = c.take<unsigned>(2);
    w.outcome = c.take<double>(4);
    // FLOAT: the rounded blocks are the system (no FP64 refinement matrix).
    const double* bsr64 = precision == NRICP_HIP_PRECISION_FLOAT ? nullptr : blocks;
    pcg_solve_on(row_off, cols, diag_pos, nn, bsr64, bsr32, rhs, damping, scaling, tol, max_iter, delta, w, stream);
    NRICP_CHECK(hipMemcpyAsync(info, w.outcome, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
    NRICP_CHECK(hipStreamSynchronize(stream));
    // The solve has drained from the caller's stream: its persistent-launch
    // charge leaves the ledger now, not when some later launch polls an
    // event recorded on a stream this library does not own.
    launches_completed(stream);
}

}  // namespace

void solve_bsr_noalloc(const int32_t* row_offsets, const int32_t* columns, const int32_t* diag_positions,
                       int64_t node_count, int64_t block_count, const double* blocks, const double* rhs, double damping,
                       bool diagonal_scaling, int precision, double tolerance, int max_iterations, double* delta,
                       double* info, hipStream_t stream) {
    solve(row_offsets, columns, diag_positions, node_count, block_count, blocks, rhs, damping, diagonal_scaling,
          precision, tolerance, max_iterations, delta, info, stream);
}

}  // namespace nricp_hip
