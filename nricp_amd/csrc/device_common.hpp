// Device-side helpers shared by the gfx950 kernels: FP64 vector math, wave64
// reductions, a grid barrier for persistent launches, and error plumbing.
//
// Every kernel is compiled with -ffp-contract=off: the closest-point, sampling
// and binding arithmetic is compared bit for bit with the FP64 CPU route, so
// no multiply-add may be fused behind the source's back. Kernels that want an
// FMA ask for it with fma().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace nricp_hip {

// ------------------------------------------------------------ errors -------

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct InvalidArgument : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

inline void check(hipError_t status, const char* what) {
    if (status != hipSuccess) {
        throw DeviceError(std::string(what) + ": " + hipGetErrorString(status));
    }
}
#define NRICP_CHECK(expr) ::nricp_hip::check((expr), #expr)
#define NRICP_LAUNCH_CHECK(name) ::nricp_hip::check(hipGetLastError(), name)

constexpr int kWave = 64;

inline int blocks_for(int64_t n, int threads) {
    return static_cast<int>((n + threads - 1) / threads);
}

// ------------------------------------------------------------ vectors ------

struct d3 {
    double x, y, z;
};
__host__ __device__ inline d3 mk(double x, double y, double z) { return {x, y, z}; }
__host__ __device__ inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline d3 operator/(d3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline d3 cross(d3 a, d3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline double norm2(d3 a) { return dot(a, a); }
__device__ inline d3 unit(d3 a) {
    const double n = sqrt(dot(a, a));
    return n > 0.0 ? a / n : d3{0.0, 0.0, 0.0};
}
__device__ inline d3 ld3(const double* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ inline void st3(double* p, int64_t i, d3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
// Row-major 3x3 times vector, same operation order as the CPU route.
__device__ inline d3 mul3(const double* r, d3 v) {
    return {r[0] * v.x + r[1] * v.y + r[2] * v.z, r[3] * v.x + r[4] * v.y + r[5] * v.z,
            r[6] * v.x + r[7] * v.y + r[8] * v.z};
}

// IRLS weight of a squared residual s (robust_weight,
// src/energy/robust.cpp:30-58): kind 1 Huber (1 inside the threshold t,
// t / sqrt(s) outside), kind 2 Geman-McClure (1 / (1 + s / t^2)^2), anything
// else or t <= 0 gives 1. The query kernel applies it to every accepted
// correspondence; robust_weights_kernel exposes it elementwise.
__device__ inline double robust_weight(int kind, double s, double t) {
    if (s < 0.0) s = 0.0;
    if (kind == 1) {
        if (t <= 0.0) return 1.0;
        if (s <= t * t) return 1.0;
        return t / sqrt(s);
    }
    if (kind == 2) {
        if (t <= 0.0) return 1.0;
        const double d = 1.0 + s / (t * t);
        return 1.0 / (d * d);
    }
    return 1.0;
}

// ----------------------------------------------------------- reductions ----

// Wave-wide sums with DPP lane moves (register-to-register, no LDS round
// trip): pairs within quads, then the two quads of each half-row
// (half-mirror), then the two halves of each row (mirror) -- after which all
// sixteen lanes of a row hold the row's sum, since a + b == b + a bitwise --
// and the four row sums, read as scalars, in row order. Every lane gets the
// same bits, and the order is fixed, so sums are reproducible.
template <int Ctrl>
__device__ inline int dpp_move(int v) {
    return __builtin_amdgcn_update_dpp(0, v, Ctrl, 0xF, 0xF, false);
}
template <int Ctrl>
__device__ inline double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_move<Ctrl>(static_cast<int>(b));
    const int hi = dpp_move<Ctrl>(static_cast<int>(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <int Ctrl>
__device__ inline float dpp_move(float v) {
    return __int_as_float(dpp_move<Ctrl>(__float_as_int(v)));
}
constexpr int kDppQuadSwap1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int kDppQuadSwap2 = 0x4E;  // quad_perm [2, 3, 0, 1]
constexpr int kDppRowMirror = 0x140;
constexpr int kDppRowHalfMirror = 0x141;

__device__ inline double read_lane(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
__device__ inline float read_lane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Sum over each group of eight lanes (lanes 8g .. 8g + 7); every lane of the
// group gets it.
template <typename T>
__device__ inline T octet_sum(T v) {
    v += dpp_move<kDppQuadSwap1>(v);
    v += dpp_move<kDppQuadSwap2>(v);
    v += dpp_move<kDppRowHalfMirror>(v);
    return v;
}

template <typename T>
__device__ inline T wave_sum(T v) {
    v = octet_sum(v);
    v += dpp_move<kDppRowMirror>(v);
    return ((read_lane(v, 0) + read_lane(v, 16)) + read_lane(v, 32)) + read_lane(v, 48);
}

// Block-wide sum of `count` values per thread into out[0..count), fixed order
// (wave butterflies then waves in index order) so the result is reproducible.
template <int Count, int Threads>
__device__ inline void block_sum(const double (&v)[Count], double* lds /* Count * Threads/64 */,
                                 double (&out)[Count]) {
    constexpr int waves = Threads / kWave;
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < Count; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) lds[k * waves + wave] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < Count; ++k) {
        double s = 0.0;
        for (int w = 0; w < waves; ++w) s += lds[k * waves + w];
        out[k] = s;
    }
    __syncthreads();
}

// ------------------------------------------------------ persistent grids ----
//
// The PCG runs as a persistent grid whose groups meet at grid barriers, so
// every group of a launch must be resident at once. It is an ordinary launch
// -- a cooperative launch would serialise it with the work of other streams,
// where the sampling and other jobs run -- and residency is guaranteed by
// construction instead:
//   * a launch never has more groups than the device holds at once for that
//     kernel (occupancy query), nor more than half of the compute units;
//   * launch_persistent admits a launch only while the persistent launches in
//     flight on the device (other jobs' included; each released once the
//     event recorded after it completes) leave a compute unit for every
//     group on every XCD (a launch of g groups is charged ceil(g / xcds) *
//     xcds, the most any XCD receives when groups are dealt to XCDs in turn);
//   * every other kernel runs to completion without waiting on anything
//     outside itself (the sampler is one workgroup; a sort pass's workgroup
//     waits only on workgroups of its own launch that started before it), so
//     the groups of an admitted launch that are not yet resident become
//     resident once the ordinary kernels ahead of them retire.
int persistent_group_budget(const void* kernel, int threads);  // groups one persistent launch may use
// `prelude`, when given, enqueues the work that precedes the launch on the
// stream; it runs once the launch is admitted, so that work and the launch
// reach the device back to back instead of around the admission.
void launch_persistent(const void* kernel, int groups, int threads, void** args, hipStream_t stream,
                       void (*prelude)(void*) = nullptr, void* prelude_arg = nullptr);
// The same admission for work the caller enqueues itself (a graph holding a
// persistent kernel, or a packed launch whose jobs each keep `groups` groups
// resident at most): waits until `groups` groups fit, calls enqueue(arg) and
// records the ledger's event after it on `stream`.
void admit_persistent(int groups, hipStream_t stream, hipError_t (*enqueue)(void*), void* arg);

// --------------------------------------------------------- grid barrier ----

// Grid barrier for persistent grids (launched under the residency rules
// above) on one monotonic arrival counter, zeroed before the launch: a group
// arriving at barrier k draws a ticket in [kP, (k+1)P) and waits until the
// counter reaches (k+1)P -- one atomic and a poll, no reset and no separate
// generation word. The ticket is an agent-scope release (this XCD's L2 is
// written back first); the poll is relaxed, and one agent-scope acquire fence
// after it (it invalidates this XCD's L2 and L1) orders every global write
// before the barrier against every read after it. An acquire on each poll
// would invalidate the L2 once per poll, evicting the lines that the groups
// still computing on the same XCD are reading. grid_arrive_wait is the part
// thread 0 runs between the two workgroup barriers.
__device__ inline void grid_arrive_wait(unsigned* state, unsigned participants) {
    const unsigned ticket = __hip_atomic_fetch_add(state, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = (ticket / participants + 1u) * participants;
    while (static_cast<int>(__hip_atomic_load(state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

__device__ inline void grid_barrier(unsigned* state, unsigned participants) {
    __syncthreads();
    if (threadIdx.x == 0) grid_arrive_wait(state, participants);
    __syncthreads();
}

}  // namespace nricp_hip
