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

namespace nricp_hip {
namespace {

// ------------------------------------------------------------ rotations ----

struct m3 {
    double m[9];
};

__device__ inline m3 mat_mul(const m3& a, const m3& b) {
    m3 o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a.m[r * 3 + k] * b.m[k * 3 + c];
            o.m[r * 3 + c] = s;
        }
    return o;
}

__device__ inline m3 skew(d3 w) {
    return m3{{0.0, -w.z, w.y, w.z, 0.0, -w.x, -w.y, w.x, 0.0}};
}

// Rodrigues with the second-order series below theta^2 = 1e-24.
__device__ inline m3 exp_so3(d3 w) {
    const double t2 = dot(w, w);
    const m3 k = skew(w);
    const m3 k2 = mat_mul(k, k);
    double a, b;
    if (t2 < 1e-24) {
        a = 1.0;
        b = 0.5;
    } else {
        const double t = sqrt(t2);
        a = sin(t) / t;
        b = (1.0 - cos(t)) / t2;
    }
    m3 o;
    for (int i = 0; i < 9; ++i) o.m[i] = (i % 4 == 0 ? 1.0 : 0.0) + (a * k.m[i] + b * k2.m[i]);
    return o;
}

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

constexpr int kFpsThreads = 256;
constexpr int kFpsPerThread = 16;

struct ArgMax {
    double value;
    int64_t index;
};

// The larger value wins; equal values go to the lower index (the first
// maximum of an in-order sweep).
__device__ inline ArgMax better(ArgMax a, ArgMax b) {
    if (b.value > a.value || (b.value == a.value && b.index < a.index)) return b;
    return a;
}

__device__ inline ArgMax block_argmax(ArgMax v, double* lds_v, int64_t* lds_i) {
    for (int offset = 32; offset > 0; offset >>= 1) {
        ArgMax o{__shfl_xor(v.value, offset, 64), __shfl_xor(v.index, offset, 64)};
        v = better(v, o);
    }
    const int wave = threadIdx.x / 64;
    if (threadIdx.x % 64 == 0) {
        lds_v[wave] = v.value;
        lds_i[wave] = v.index;
    }
    __syncthreads();
    ArgMax r{lds_v[0], lds_i[0]};
    for (int w = 1; w < kFpsThreads / 64; ++w) r = better(r, ArgMax{lds_v[w], lds_i[w]});
    __syncthreads();
    return r;
}

// Persistent FPS: each workgroup owns a contiguous chunk of points and keeps
// their running nearest squared distances in registers. One grid barrier per
// seed: every group publishes its chunk's (max, index), and after the barrier
// every group folds the same partials in the same order.
__global__ void __launch_bounds__(kFpsThreads)
fps_kernel(const double* __restrict__ p, int64_t n, int32_t seeds, int32_t first, int32_t* __restrict__ out,
           double* partial_v, int64_t* partial_i, unsigned* barrier, int64_t chunk) {
    __shared__ double lds_v[kFpsThreads / 64];
    __shared__ int64_t lds_i[kFpsThreads / 64];
    const unsigned groups = gridDim.x;
    const int64_t begin = static_cast<int64_t>(blockIdx.x) * chunk;
    double nearest[kFpsPerThread];
    d3 mine[kFpsPerThread];
    const d3 c0 = ld3(p, first);
    ArgMax local{-1.0, 0x7fffffffffffffffll};
#pragma unroll
    for (int k = 0; k < kFpsPerThread; ++k) {
        const int64_t i = begin + static_cast<int64_t>(k) * kFpsThreads + threadIdx.x;
        if (i < n && i < begin + chunk) {
            mine[k] = ld3(p, i);
            nearest[k] = norm2(mine[k] - c0);
            local = better(local, ArgMax{nearest[k], i});
        } else {
            nearest[k] = -1.0;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = first;
    for (int32_t s = 1; s < seeds; ++s) {
        const ArgMax g = block_argmax(local, lds_v, lds_i);
        double* pv = partial_v + (s & 1) * groups;
        int64_t* pi = partial_i + (s & 1) * groups;
        if (threadIdx.x == 0) {
            pv[blockIdx.x] = g.value;
            pi[blockIdx.x] = g.index;
        }
        if (groups > 1) grid_barrier(barrier, groups);
        ArgMax best{-1.0, 0x7fffffffffffffffll};
        if (groups > 1) {
            for (unsigned b = 0; b < groups; ++b) best = better(best, ArgMax{pv[b], pi[b]});
        } else {
            best = g;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) out[s] = static_cast<int32_t>(best.index);
        const d3 c = ld3(p, best.index);
        local = ArgMax{-1.0, 0x7fffffffffffffffll};
#pragma unroll
        for (int k = 0; k < kFpsPerThread; ++k) {
            if (nearest[k] >= 0.0) {
                const double d = norm2(mine[k] - c);
                nearest[k] = fmin(nearest[k], d);
                const int64_t i = begin + static_cast<int64_t>(k) * kFpsThreads + threadIdx.x;
                local = better(local, ArgMax{nearest[k], i});
            }
        }
    }
}

// Nearest `support` nodes by (d^2, index) with the insertion order of the CPU
// route's keep_nearest, then (1 - d/dmax)^2 weights normalised.
__device__ inline void nearest_nodes(d3 x, const double* __restrict__ nodes, int64_t nn, int support,
                                     double* d2, int32_t* id) {
    int held = 0;
    for (int64_t j = 0; j < nn; ++j) {
        const double s = norm2(x - ld3(nodes, j));
        if (held == support) {
            const bool beats = s < d2[support - 1] || (s == d2[support - 1] && j < id[support - 1]);
            if (!beats) continue;
        }
        int at = held < support ? held++ : support - 1;
        while (at > 0 && (d2[at - 1] > s || (d2[at - 1] == s && id[at - 1] > j))) {
            d2[at] = d2[at - 1];
            id[at] = id[at - 1];
            --at;
        }
        d2[at] = s;
        id[at] = static_cast<int32_t>(j);
    }
}

__device__ inline void distance_weights(const double* d2, int support, double* w) {
    const double dmax = sqrt(d2[support - 1]);
    double total = 0.0;
    for (int k = 0; k < support; ++k) {
        const double d = sqrt(d2[k]);
        const double ratio = dmax > 0.0 ? d / dmax : 0.0;
        const double f = 1.0 - ratio;
        w[k] = f * f;
        total += w[k];
    }
    for (int k = 0; k < support; ++k) w[k] = total > 0.0 ? w[k] / total : 1.0 / support;
}

__global__ void bind_kernel(const double* __restrict__ p, int64_t n, const double* __restrict__ nodes, int64_t nn,
                            int support, int32_t* __restrict__ on, double* __restrict__ ow, int32_t* __restrict__ oc) {
    const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    double d2[4];
    int32_t id[4];
    nearest_nodes(ld3(p, v), nodes, nn, support, d2, id);
    double w[4];
    distance_weights(d2, support, w);
    for (int k = 0; k < 4; ++k) {
        on[4 * v + k] = k < support ? id[k] : -1;
        ow[4 * v + k] = k < support ? w[k] : 0.0;
    }
    oc[v] = support;
}

// k nearest other nodes by (Euclidean distance, index); emits the (min, max)
// pair of each as a sortable key.
__global__ void knn_kernel(const double* __restrict__ nodes, int64_t nn, int k, uint64_t* __restrict__ keys) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    double dist[16];
    int32_t id[16];
    int held = 0;
    const d3 x = ld3(nodes, i);
    for (int64_t j = 0; j < nn; ++j) {
        if (j == i) continue;
        const double s = sqrt(norm2(x - ld3(nodes, j)));
        if (held == k) {
            const bool beats = s < dist[k - 1] || (s == dist[k - 1] && j < id[k - 1]);
            if (!beats) continue;
        }
        int at = held < k ? held++ : k - 1;
        while (at > 0 && (dist[at - 1] > s || (dist[at - 1] == s && id[at - 1] > j))) {
            dist[at] = dist[at - 1];
            id[at] = id[at - 1];
            --at;
        }
        dist[at] = s;
        id[at] = static_cast<int32_t>(j);
    }
    for (int m = 0; m < k; ++m) {
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
__global__ void transfer_kernel(const double* __restrict__ from, int64_t nf, const double* __restrict__ R,
                                const double* __restrict__ T, const double* __restrict__ to, int64_t nt, int support,
                                double* __restrict__ Ro, double* __restrict__ To) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const d3 x = ld3(to, i);
    double d2[4];
    int32_t id[4];
    nearest_nodes(x, from, nf, support, d2, id);
    double w[4];
    distance_weights(d2, support, w);
    int32_t heavy = id[0];
    double hw = w[0];
    for (int s = 1; s < support; ++s)
        if (w[s] > hw) {
            hw = w[s];
            heavy = id[s];
        }
    const quat ref = quat_from_matrix(R + 9 * static_cast<int64_t>(heavy));
    quat blend{0.0, 0.0, 0.0, 0.0};
    d3 pos{0.0, 0.0, 0.0};
    for (int s = 0; s < support; ++s) {
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

void farthest_point_seeds(const double* points, int64_t n, int32_t seeds, int32_t first, int32_t* out,
                          DeviceBuffer<double>& partial, DeviceBuffer<unsigned>& barrier, hipStream_t stream) {
    if (n <= 0 || seeds <= 0) return;
    // Enough groups that each thread holds at most kFpsPerThread points, and
    // no more than can be co-resident.
    const int64_t per_group = static_cast<int64_t>(kFpsThreads) * kFpsPerThread;
    const int64_t groups64 = (n + per_group - 1) / per_group;
    int device = 0;
    NRICP_CHECK(hipGetDevice(&device));
    int cus = 0;
    NRICP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    int per_cu = 0;
    NRICP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fps_kernel, kFpsThreads, 0));
    const int64_t resident = static_cast<int64_t>(cus) * std::max(per_cu, 1);
    if (groups64 > resident)
        throw InvalidArgument("farthest_point_seeds: " + std::to_string(n) + " points exceed the resident capacity");
    const unsigned groups = static_cast<unsigned>(groups64);
    const int64_t chunk = per_group;
    partial.resize(4 * static_cast<int64_t>(groups));  // 2 buffers x (value, index)
    barrier.resize(2);
    barrier.zero(stream);
    double* pv = partial.get();
    int64_t* pi = reinterpret_cast<int64_t*>(partial.get() + 2 * groups);
    int32_t s = seeds;
    int32_t f = first;
    int64_t nn = n;
    unsigned* bar = barrier.get();
    void* args[] = {const_cast<double**>(&points), &nn, &s, &f, &out, &pv, &pi, &bar, const_cast<int64_t*>(&chunk)};
    NRICP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(fps_kernel), dim3(groups), dim3(kFpsThreads),
                                           args, 0, stream));
}

void bind_vertices(const double* points, int64_t n, const double* nodes, int64_t nn, int32_t k, int32_t* on,
                   double* ow, int32_t* oc, hipStream_t stream) {
    if (n <= 0 || nn <= 0) return;
    const int support = std::max(1, std::min<int>({k, static_cast<int>(std::min<int64_t>(nn, 4)), 4}));
    bind_kernel<<<blocks_for(n, 128), 128, 0, stream>>>(points, n, nodes, nn, support, on, ow, oc);
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
    knn_kernel<<<blocks_for(nn, 128), 128, 0, stream>>>(nodes, nn, kk, keys.get());
    NRICP_LAUNCH_CHECK("knn_kernel");
    radix_sort_pairs(keys.get(), nullptr, m, 64, scratch, stream);
    flags.resize(m + 1);
    pos.resize(m + 1);
    run_starts(keys.get(), m, flags.get(), stream);
    exclusive_scan32(flags.get(), pos.get(), m, stream);
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
    transfer_kernel<<<blocks_for(nt, 128), 128, 0, stream>>>(from, nf, R, T, to, nt, support, Ro, To);
    NRICP_LAUNCH_CHECK("transfer_kernel");
}

}  // namespace nricp_hip
