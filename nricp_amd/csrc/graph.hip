// Template-side and node-state kernels: deformation evaluation, retraction,
// farthest-point sampling, vertex binding, node kNN edges and the level
// warm start.
//
// Reference behaviour: src/core/deformation.cpp:19-32,54-84,141-203,
// src/core/sampling.cpp:21-141, src/core/graph.cpp:11-39, and the SO(3) /
// quaternion helpers of src/core/rotation.h, src/core/quaternion.h. Sampling
// and binding compare and minimise squared distances exactly as the CPU route
// does (no contraction), so seeds, bound nodes and weights are bit-identical.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels.hpp"
#include "so3.hpp"

namespace nricp_hip {
namespace {

// ------------------------------------------------------------ rotations ----

struct quat {
    double w, x, y, z;
};

// Shepperd's branch selection.
__device__ inline quat quat_from_matrix(const double* r) {
    const double tr = r[0] + r[4] + r[8];
    if (tr > 0.0) {
        const double s = sqrt(tr + 1.0) * 2.0;
        return {0.25 * s, (r[7] - r[5]) / s, (r[2] - r[6]) / s, (r[3] - r[1]) / s};
    }
    if (r[0] > r[4] && r[0] > r[8]) {
        const double s = sqrt(1.0 + r[0] - r[4] - r[8]) * 2.0;
        return {(r[7] - r[5]) / s, 0.25 * s, (r[1] + r[3]) / s, (r[2] + r[6]) / s};
    }
    if (r[4] > r[8]) {
        const double s = sqrt(1.0 + r[4] - r[0] - r[8]) * 2.0;
        return {(r[2] - r[6]) / s, (r[1] + r[3]) / s, 0.25 * s, (r[5] + r[7]) / s};
    }
    const double s = sqrt(1.0 + r[8] - r[0] - r[4]) * 2.0;
    return {(r[3] - r[1]) / s, (r[2] + r[6]) / s, (r[5] + r[7]) / s, 0.25 * s};
}

__device__ inline void matrix_from_quat(quat q, double* m) {
    const double w = q.w, x = q.x, y = q.y, z = q.z;
    m[0] = 1.0 - 2.0 * (y * y + z * z);
    m[1] = 2.0 * (x * y - z * w);
    m[2] = 2.0 * (x * z + y * w);
    m[3] = 2.0 * (x * y + z * w);
    m[4] = 1.0 - 2.0 * (x * x + z * z);
    m[5] = 2.0 * (y * z - x * w);
    m[6] = 2.0 * (x * z - y * w);
    m[7] = 2.0 * (y * z + x * w);
    m[8] = 1.0 - 2.0 * (x * x + y * y);
}

__device__ inline double qdot(quat a, quat b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }

// ------------------------------------------------------------ evaluate -----

__global__ void evaluate_kernel(const double* __restrict__ p, const int32_t* __restrict__ bn,
                                const double* __restrict__ bw, const int32_t* __restrict__ bc, int64_t nv,
                                const double* __restrict__ R, const double* __restrict__ T,
                                const double* __restrict__ G, int64_t nn, double* __restrict__ out) {
    const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const d3 x = ld3(p, v);
    d3 s{0.0, 0.0, 0.0};
    const int count = bc[v];
    for (int k = 0; k < count; ++k) {
        const int32_t j = bn[4 * v + k];
        if (j < 0 || j >= nn) continue;
        const d3 g = ld3(G, j);
        const d3 moved = (mul3(R + 9 * static_cast<int64_t>(j), x - g) + g) + ld3(T, j);
        s = s + moved * bw[4 * v + k];
    }
    st3(out, v, s);
}

__global__ void retract_kernel(const double* __restrict__ R, const double* __restrict__ T,
                               const double* __restrict__ delta, int64_t nn, double* __restrict__ Ro,
                               double* __restrict__ To) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nn) return;
    const double* d = delta + 6 * j;
    const m3 e = exp_so3(d3{d[0], d[1], d[2]});
    m3 r;
    for (int i = 0; i < 9; ++i) r.m[i] = R[9 * j + i];
    const m3 o = mat_mul(e, r);
    const d3 t = ld3(T, j) + d3{d[3], d[4], d[5]};
    for (int i = 0; i < 9; ++i) Ro[9 * j + i] = o.m[i];
    st3(To, j, t);
}

// ---------------------------------------------------------- sampling -------

constexpr int kFpsThreads = 512;
constexpr int kFpsPerThread = 16;

// A sampling candidate: its running nearest squared distance, its index and
// its position (carried along so the winner's position needs no extra load).
struct ArgMax {
    double value;
    int64_t index;
    d3 point;
};

// The larger value wins; equal values go to the lower index (the first
// maximum of an in-order sweep). A total order, so any reduction order gives
// the same winner.
__device__ inline ArgMax better(ArgMax a, ArgMax b) {
    if (b.value > a.value || (b.value == a.value && b.index < a.index)) return b;
    return a;
}

__device__ inline ArgMax shfl_xor(const ArgMax& v, int m) {
    return ArgMax{__shfl_xor(v.value, m, 64), __shfl_xor(v.index, m, 64),
                  d3{__shfl_xor(v.point.x, m, 64), __shfl_xor(v.point.y, m, 64), __shfl_xor(v.point.z, m, 64)}};
}

__device__ inline ArgMax wave_argmax(ArgMax v) {
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) v = better(v, shfl_xor(v, offset));
    return v;
}

constexpr int kFpsWaves = kFpsThreads / 64;

// Persistent FPS: each workgroup owns a contiguous chunk of points and keeps
// them and their running nearest squared distances in registers. One grid
// barrier per seed: every wave publishes its winner (value, index, position)
// to its slot; after the barrier every wave loads all slots and reduces them
// with shuffles, so every wave of every group picks the same seed (the order
// (value, lower index) is total, so the reduction order cannot matter). A
// launch computes seeds [begin, end); the running distances go back to memory
// at its end, so the next launch resumes where this one stopped.
__global__ void __launch_bounds__(kFpsThreads)
fps_kernel(const double* __restrict__ p, int64_t n, int32_t begin, int32_t end, int32_t first,
           int32_t* __restrict__ out, double* __restrict__ nearest_g, double* partial_v /* 2 x slots x 4 */,
           int64_t* partial_i /* 2 x slots */, unsigned* barrier, int64_t chunk) {
    // The sampling is one long dependent chain running beside the solve's
    // kernels: its waves take issue priority on the units they share.
    __builtin_amdgcn_s_setprio(3);
    const unsigned groups = gridDim.x;
    const unsigned slots = groups * kFpsWaves;
    const unsigned my_slot = blockIdx.x * kFpsWaves + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * chunk;
    double nearest[kFpsPerThread];
    d3 mine[kFpsPerThread];
    const ArgMax none{-1.0, 0x7fffffffffffffffll, d3{0.0, 0.0, 0.0}};
    ArgMax local = none;
    const d3 c0 = ld3(p, begin == 0 ? first : out[begin - 1]);
#pragma unroll
    for (int k = 0; k < kFpsPerThread; ++k) {
        const int64_t i = base + static_cast<int64_t>(k) * kFpsThreads + threadIdx.x;
        if (i < n && i < base + chunk) {
            mine[k] = ld3(p, i);
            nearest[k] = begin == 0 ? norm2(mine[k] - c0) : nearest_g[i];
            local = better(local, ArgMax{nearest[k], i, mine[k]});
        } else {
            mine[k] = d3{0.0, 0.0, 0.0};
            nearest[k] = -1.0;
        }
    }
    if (begin == 0 && blockIdx.x == 0 && threadIdx.x == 0) out[0] = first;
    for (int32_t s = begin > 0 ? begin : 1; s < end; ++s) {
        const ArgMax w = wave_argmax(local);
        double* pv = partial_v + (s & 1) * 4 * slots;
        int64_t* pi = partial_i + (s & 1) * slots;
        if (lane == 0) {
            pv[4 * my_slot] = w.value;
            pv[4 * my_slot + 1] = w.point.x;
            pv[4 * my_slot + 2] = w.point.y;
            pv[4 * my_slot + 3] = w.point.z;
            pi[my_slot] = w.index;
        }
        grid_barrier(barrier, groups);
        ArgMax best = none;
        for (unsigned t = lane; t < slots; t += 64)
            best = better(best, ArgMax{pv[4 * t], pi[t], d3{pv[4 * t + 1], pv[4 * t + 2], pv[4 * t + 3]}});
        best = wave_argmax(best);
        if (blockIdx.x == 0 && threadIdx.x == 0) out[s] = static_cast<int32_t>(best.index);
        const d3 c = best.point;
        local = none;
#pragma unroll
        for (int k = 0; k < kFpsPerThread; ++k) {
            if (nearest[k] >= 0.0) {
                const double d = norm2(mine[k] - c);
                nearest[k] = fmin(nearest[k], d);
                const int64_t i = base + static_cast<int64_t>(k) * kFpsThreads + threadIdx.x;
                local = better(local, ArgMax{nearest[k], i, mine[k]});
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kFpsPerThread; ++k) {
        const int64_t i = base + static_cast<int64_t>(k) * kFpsThreads + threadIdx.x;
        if (nearest[k] >= 0.0) nearest_g[i] = nearest[k];
    }
}

// Sorted insertion of (d, j) into the KMAX best entries kept in ascending
// (d, index) order: the candidate bubbles down the fully unrolled list, so
// every index is a compile-time constant and the list stays in registers.
// Entries start at (+inf, INT32_MAX), which every real candidate beats.
template <int KMAX>
__device__ inline void insert_sorted(double d, int32_t j, double* dist, int32_t* id) {
#pragma unroll
    for (int m = 0; m < KMAX; ++m) {
        const bool beats = d < dist[m] || (d == dist[m] && j < id[m]);
        const double td = dist[m];
        const int32_t tj = id[m];
        dist[m] = beats ? d : td;
        id[m] = beats ? j : tj;
        d = beats ? td : d;
        j = beats ? tj : j;
    }
}

// Nearest 4 nodes by (d^2, index), the order the CPU route's keep_nearest
// produces (candidates visited in increasing index, ties kept by the lower
// index); callers use the first `support` of them. Nodes are staged through
// LDS a tile at a time by the whole block, so every thread of the block must
// call it, including those past the end of the rows.
template <int Threads>
__device__ inline void nearest4(d3 x, const double* __restrict__ nodes, int64_t nn, double* d2, int32_t* id,
                                double* tile) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        d2[m] = __builtin_inf();
        id[m] = 0x7fffffff;
    }
    for (int64_t base = 0; base < nn; base += Threads) {
        __syncthreads();
        const int64_t j = base + threadIdx.x;
        if (j < nn) {
            tile[3 * threadIdx.x] = nodes[3 * j];
            tile[3 * threadIdx.x + 1] = nodes[3 * j + 1];
            tile[3 * threadIdx.x + 2] = nodes[3 * j + 2];
        }
        __syncthreads();
        const int count = static_cast<int>(nn - base < Threads ? nn - base : Threads);
        for (int t = 0; t < count; ++t) {
            const d3 g{tile[3 * t], tile[3 * t + 1], tile[3 * t + 2]};
            const double s = norm2(x - g);
            if (s < d2[3] || (s == d2[3] && base + t < id[3]))
                insert_sorted<4>(s, static_cast<int32_t>(base + t), d2, id);
        }
    }
}

__device__ inline double pick4(const double* v, int k) {
    return k == 0 ? v[0] : (k == 1 ? v[1] : (k == 2 ? v[2] : v[3]));
}

// (1 - d/dmax)^2 over the first `support` entries, normalised (uniform when
// they all vanish); entries past `support` get weight 0.
__device__ inline void distance_weights(const double* d2, int support, double* w) {
    const double dmax = sqrt(pick4(d2, support - 1));
    double total = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        w[k] = 0.0;
        if (k < support) {
            const double d = sqrt(d2[k]);
            const double ratio = dmax > 0.0 ? d / dmax : 0.0;
            const double f = 1.0 - ratio;
            w[k] = f * f;
            total += w[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k < support) w[k] = total > 0.0 ? w[k] / total : 1.0 / support;
}

constexpr int kBindThreads = 128;

__global__ void __launch_bounds__(kBindThreads)
bind_kernel(const double* __restrict__ p, int64_t n, const double* __restrict__ nodes, int64_t nn, int support,
            int32_t* __restrict__ on, double* __restrict__ ow, int32_t* __restrict__ oc) {
    __shared__ double tile[3 * kBindThreads];
    const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const d3 x = v < n ? ld3(p, v) : d3{0.0, 0.0, 0.0};
    double d2[4];
    int32_t id[4];
    nearest4<kBindThreads>(x, nodes, nn, d2, id, tile);
    if (v >= n) return;
    double w[4];
    distance_weights(d2, support, w);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        on[4 * v + k] = k < support ? id[k] : -1;
        ow[4 * v + k] = k < support ? w[k] : 0.0;
    }
    oc[v] = support;
}

// k nearest other nodes by (Euclidean distance, index); emits the (min, max)
// pair of each as a sortable key. KMAX >= k entries are kept (the first k are
// the answer); node positions are staged through LDS a tile at a time.
constexpr int kKnnThreads = 64;
template <int KMAX>
__global__ void __launch_bounds__(kKnnThreads)
knn_kernel(const double* __restrict__ nodes, int64_t nn, int k, uint64_t* __restrict__ keys) {
    __shared__ double tile[3 * kKnnThreads];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double dist[KMAX];
    int32_t id[KMAX];
#pragma unroll
    for (int m = 0; m < KMAX; ++m) {
        dist[m] = __builtin_inf();
        id[m] = 0x7fffffff;
    }
    const d3 x = i < nn ? ld3(nodes, i) : d3{0.0, 0.0, 0.0};
    for (int64_t base = 0; base < nn; base += kKnnThreads) {
        __syncthreads();
        const int64_t j = base + threadIdx.x;
        if (j < nn) {
            tile[3 * threadIdx.x] = nodes[3 * j];
            tile[3 * threadIdx.x + 1] = nodes[3 * j + 1];
            tile[3 * threadIdx.x + 2] = nodes[3 * j + 2];
        }
        __syncthreads();
        const int count = static_cast<int>(nn - base < kKnnThreads ? nn - base : kKnnThreads);
        for (int t = 0; t < count; ++t) {
            if (base + t == i) continue;
            const d3 g{tile[3 * t], tile[3 * t + 1], tile[3 * t + 2]};
            const double s = sqrt(norm2(x - g));
            if (s < dist[KMAX - 1] || (s == dist[KMAX - 1] && base + t < id[KMAX - 1]))
                insert_sorted<KMAX>(s, static_cast<int32_t>(base + t), dist, id);
        }
    }
    if (i >= nn) return;
    const int held = static_cast<int>(nn - 1 < k ? nn - 1 : k);
#pragma unroll
    for (int m = 0; m < KMAX; ++m) {
        if (m >= k) break;
        uint64_t key = ~0ull;
        if (m < held) {
            const uint64_t a = static_cast<uint64_t>(i < id[m] ? i : id[m]);
            const uint64_t b = static_cast<uint64_t>(i < id[m] ? id[m] : i);
            key = a * static_cast<uint64_t>(nn) + b;
        }
        keys[static_cast<int64_t>(k) * i + m] = key;
    }
}

__global__ void unique_edges(const uint64_t* __restrict__ keys, const int32_t* __restrict__ flags,
                             const int32_t* __restrict__ pos, int64_t n, int64_t nn, int32_t* __restrict__ edges) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i] || keys[i] == ~0ull) return;
    const int64_t e = pos[i];
    edges[2 * e] = static_cast<int32_t>(keys[i] / static_cast<uint64_t>(nn));
    edges[2 * e + 1] = static_cast<int32_t>(keys[i] % static_cast<uint64_t>(nn));
}

// Blend of the bound source rotations in quaternion space, signs aligned to
// the heaviest contributor, heaviest-neighbour fallback when the blend nearly
// cancels; translation maps the new rest position onto the field's value.
__global__ void __launch_bounds__(kBindThreads)
transfer_kernel(const double* __restrict__ from, int64_t nf, const double* __restrict__ R,
                                const double* __restrict__ T, const double* __restrict__ to, int64_t nt, int support,
                                double* __restrict__ Ro, double* __restrict__ To) {
    __shared__ double tile[3 * kBindThreads];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const d3 x = i < nt ? ld3(to, i) : d3{0.0, 0.0, 0.0};
    double d2[4];
    int32_t id[4];
    nearest4<kBindThreads>(x, from, nf, d2, id, tile);
    if (i >= nt) return;
    double w[4];
    distance_weights(d2, support, w);
    int32_t heavy = id[0];
    double hw = w[0];
#pragma unroll
    for (int s = 1; s < 4; ++s)
        if (s < support && w[s] > hw) {
            hw = w[s];
            heavy = id[s];
        }
    const quat ref = quat_from_matrix(R + 9 * static_cast<int64_t>(heavy));
    quat blend{0.0, 0.0, 0.0, 0.0};
    d3 pos{0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (s >= support) break;
        const int32_t j = id[s];
        quat q = quat_from_matrix(R + 9 * static_cast<int64_t>(j));
        if (qdot(q, ref) < 0.0) q = quat{-q.w, -q.x, -q.y, -q.z};
        blend = quat{blend.w + q.w * w[s], blend.x + q.x * w[s], blend.y + q.y * w[s], blend.z + q.z * w[s]};
        const d3 g = ld3(from, j);
        const d3 moved = (mul3(R + 9 * static_cast<int64_t>(j), x - g) + g) + ld3(T, j);
        pos = pos + moved * w[s];
    }
    const double len = sqrt(qdot(blend, blend));
    quat q = ref;
    if (len > 1e-8) {
        const double l = sqrt(qdot(blend, blend));
        q = quat{blend.w / l, blend.x / l, blend.y / l, blend.z / l};
    }
    matrix_from_quat(q, Ro + 9 * i);
    st3(To, i, pos - x);
}

}  // namespace

void evaluate_deformation(const double* points, const int32_t* bn, const double* bw, const int32_t* bc, int64_t nv,
                          const double* R, const double* T, const double* G, int64_t nn, double* out,
                          hipStream_t stream) {
    if (nv <= 0) return;
    evaluate_kernel<<<blocks_for(nv, 256), 256, 0, stream>>>(points, bn, bw, bc, nv, R, T, G, nn, out);
    NRICP_LAUNCH_CHECK("evaluate_kernel");
}

void retract(const double* R, const double* T, const double* delta, int64_t nn, double* Ro, double* To,
             hipStream_t stream) {
    if (nn <= 0) return;
    retract_kernel<<<blocks_for(nn, 256), 256, 0, stream>>>(R, T, delta, nn, Ro, To);
    NRICP_LAUNCH_CHECK("retract_kernel");
}

void farthest_point_seeds_range(const double* points, int64_t n, int32_t first, int32_t begin, int32_t end,
                                int32_t* out, FpsWork& w, hipStream_t stream) {
    if (n <= 0 || end <= begin) return;
    if (first < 0 || first >= n) throw InvalidArgument("farthest_point_seeds: the first seed is out of range");
    // Enough groups that each thread holds at most kFpsPerThread points,
    // within the persistent-launch budget.
    const int64_t per_group = static_cast<int64_t>(kFpsThreads) * kFpsPerThread;
    const int64_t groups64 = (n + per_group - 1) / per_group;
    const int64_t budget = persistent_group_budget(reinterpret_cast<const void*>(fps_kernel), kFpsThreads);
    if (groups64 > budget)
        throw InvalidArgument("farthest_point_seeds: " + std::to_string(n) + " points exceed one persistent launch");
    const unsigned groups = static_cast<unsigned>(groups64);
    const int64_t chunk = per_group;
    if (begin == 0) {
        w.nearest.resize(n);
        w.partial.resize(10 * static_cast<int64_t>(groups) * kFpsWaves);  // 2 buffers x slots x (value, x, y, z, index)
        w.barrier.resize(2);
        w.barrier.zero(stream);
    }
    double* pv = w.partial.get();
    int64_t* pi = reinterpret_cast<int64_t*>(w.partial.get() + 8 * static_cast<int64_t>(groups) * kFpsWaves);
    int32_t b = begin, e = end, f = first;
    int64_t nn = n;
    double* near = w.nearest.get();
    unsigned* bar = w.barrier.get();
    void* args[] = {const_cast<double**>(&points), &nn, &b, &e, &f, &out, &near, &pv, &pi, &bar,
                    const_cast<int64_t*>(&chunk)};
    launch_persistent(reinterpret_cast<const void*>(fps_kernel), static_cast<int>(groups), kFpsThreads, args, stream);
}

void farthest_point_seeds(const double* points, int64_t n, int32_t seeds, int32_t first, int32_t* out, FpsWork& w,
                          hipStream_t stream) {
    if (n <= 0 || seeds <= 0) return;
    farthest_point_seeds_range(points, n, first, 0, seeds, out, w, stream);
}

void bind_vertices(const double* points, int64_t n, const double* nodes, int64_t nn, int32_t k, int32_t* on,
                   double* ow, int32_t* oc, hipStream_t stream) {
    if (n <= 0 || nn <= 0) return;
    const int support = std::max(1, std::min<int>({k, static_cast<int>(std::min<int64_t>(nn, 4)), 4}));
    bind_kernel<<<blocks_for(n, kBindThreads), kBindThreads, 0, stream>>>(points, n, nodes, nn, support, on, ow, oc);
    NRICP_LAUNCH_CHECK("bind_kernel");
}

int64_t build_node_edges(const double* nodes, int64_t nn, int32_t k, DeviceBuffer<int32_t>& edges,
                         DeviceBuffer<uint64_t>& keys, DeviceBuffer<int32_t>& flags, DeviceBuffer<int32_t>& pos,
                         SortScratch& scratch, hipStream_t stream) {
    if (nn <= 1 || k <= 0) {
        edges.resize(0);
        return 0;
    }
    if (k > 16) throw InvalidArgument("graph_neighbours must be at most 16");
    const int kk = static_cast<int>(std::min<int64_t>(k, nn - 1));
    const int64_t m = nn * kk;
    keys.resize(m);
    const dim3 grid(static_cast<unsigned>(blocks_for(nn, kKnnThreads)));
    if (kk <= 4)
        knn_kernel<4><<<grid, kKnnThreads, 0, stream>>>(nodes, nn, kk, keys.get());
    else if (kk <= 8)
        knn_kernel<8><<<grid, kKnnThreads, 0, stream>>>(nodes, nn, kk, keys.get());
    else
        knn_kernel<16><<<grid, kKnnThreads, 0, stream>>>(nodes, nn, kk, keys.get());
    NRICP_LAUNCH_CHECK("knn_kernel");
    radix_sort_pairs(keys.get(), nullptr, m, bits_for(static_cast<uint64_t>(nn) * static_cast<uint64_t>(nn)), scratch,
                     stream);
    flags.resize(m + 1);
    pos.resize(m + 1);
    run_starts(keys.get(), m, flags.get(), stream);
    exclusive_scan32(flags.get(), pos.get(), m, scratch, stream);
    int32_t count = 0;
    NRICP_CHECK(hipMemcpyAsync(&count, pos.get() + m, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    NRICP_CHECK(hipStreamSynchronize(stream));
    edges.resize(2 * static_cast<int64_t>(count));
    unique_edges<<<blocks_for(m, 256), 256, 0, stream>>>(keys.get(), flags.get(), pos.get(), m, nn, edges.get());
    NRICP_LAUNCH_CHECK("unique_edges");
    // kNN lists are never short of k entries for kk <= nn - 1, so no ~0 keys
    // were emitted and every run start is a real edge.
    return count;
}

void transfer_level(const double* from, int64_t nf, const double* R, const double* T, const double* to, int64_t nt,
                    int32_t k, double* Ro, double* To, hipStream_t stream) {
    if (nt <= 0) return;
    const int support = std::max(1, std::min<int>({k, static_cast<int>(std::min<int64_t>(nf, 4)), 4}));
    transfer_kernel<<<blocks_for(nt, kBindThreads), kBindThreads, 0, stream>>>(from, nf, R, T, to, nt, support, Ro, To);
    NRICP_LAUNCH_CHECK("transfer_kernel");
}

}  // namespace nricp_hip
