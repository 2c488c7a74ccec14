// Target-side kernels: vertex normals, the LBVH over the target triangles, and
// the closest-point query with the correspondence filters and IRLS weights.
//
// Reference behaviour: closest point and filters follow
// src/geometry/closest_point.cpp:24-279 and src/solver/correspondence.cpp:18-96
// (FP64, operation order kept so distances match the CPU route bit for bit);
// the hierarchy replaces src/cuda/lbvh.cu with a 64-bit-key Karras tree with
// outward-rounded FP32 boxes, traversed depth-first with a per-lane LDS stack
// deep enough for any tree over such keys, so no subtree can be dropped.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels.hpp"
#include "packed.hpp"
#include "xcd_map.hpp"

namespace nricp_hip {
namespace {

// --------------------------------------------------------------- normals ---

// Each vertex sums the cross products of its faces in increasing face order
// starting from zero, then normalises: the same sequence of additions the CPU
// route performs per vertex. A face's cross product is recomputed by each of
// its vertices (same operations, same bits) rather than stored by a pass of
// its own, which saves a launch per correspondence pass.
__device__ __forceinline__ void vertex_normal(const double* __restrict__ p, const int32_t* __restrict__ f,
                                     const int32_t* __restrict__ off, const int32_t* __restrict__ faces, int64_t nv,
                                     double* __restrict__ out, int64_t v) {
    if (v >= nv) return;
    d3 s{0.0, 0.0, 0.0};
    for (int32_t k = off[v]; k < off[v + 1]; ++k) {
        const int64_t t = faces[k];
        const d3 a = ld3(p, f[3 * t]), b = ld3(p, f[3 * t + 1]), c = ld3(p, f[3 * t + 2]);
        s = s + cross(b - a, c - a);
    }
    st3(out, v, unit(s));
}

__global__ void vertex_normals_kernel(const double* __restrict__ p, const int32_t* __restrict__ f,
                                      const int32_t* __restrict__ off, const int32_t* __restrict__ faces, int64_t nv,
                                      double* __restrict__ out) {
    vertex_normal(p, f, off, faces, nv, out, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
}

__global__ void vertex_face_entries(const int32_t* __restrict__ f, int64_t nf, uint64_t* __restrict__ keys,
                                    uint32_t* __restrict__ vals) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 3 * nf) return;
    keys[i] = static_cast<uint64_t>(static_cast<uint32_t>(f[i]));
    vals[i] = static_cast<uint32_t>(i / 3);
}

__global__ void copy_u32_to_i32(const uint32_t* __restrict__ in, int64_t n, int32_t* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<int32_t>(in[i]);
}

// ------------------------------------------------------------------ LBVH ---

__global__ void centroid_kernel(const double* __restrict__ p, const int32_t* __restrict__ f, int64_t nt,
                                double* __restrict__ c) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const d3 a = ld3(p, f[3 * t]), b = ld3(p, f[3 * t + 1]), cc = ld3(p, f[3 * t + 2]);
    st3(c, t, (a + b + cc) / 3.0);
}

// Two-stage min/max of the centroids: per-block partials, then one block.
__global__ void __launch_bounds__(256) bounds_partial(const double* __restrict__ c, int64_t n, double* __restrict__ part) {
    __shared__ double lo[3][256], hi[3][256];
    double l[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, h[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
        for (int k = 0; k < 3; ++k) {
            l[k] = fmin(l[k], c[3 * i + k]);
            h[k] = fmax(h[k], c[3 * i + k]);
        }
    }
    for (int k = 0; k < 3; ++k) {
        lo[k][threadIdx.x] = l[k];
        hi[k][threadIdx.x] = h[k];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 3; ++k) {
                lo[k][threadIdx.x] = fmin(lo[k][threadIdx.x], lo[k][threadIdx.x + s]);
                hi[k][threadIdx.x] = fmax(hi[k][threadIdx.x], hi[k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) {
            part[6 * blockIdx.x + k] = lo[k][0];
            part[6 * blockIdx.x + 3 + k] = hi[k][0];
        }
}

// One wave folds the partial boxes (min / max are exact in any order) into
// part[0..6).
__global__ void bounds_final(double* part, int blocks) {
    const int lane = threadIdx.x;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int b = lane; b < blocks; b += 64)
        for (int k = 0; k < 3; ++k) {
            lo[k] = fmin(lo[k], part[6 * b + k]);
            hi[k] = fmax(hi[k], part[6 * b + 3 + k]);
        }
    for (int offset = 32; offset > 0; offset >>= 1)
        for (int k = 0; k < 3; ++k) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], offset, 64));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], offset, 64));
        }
    if (lane == 0)
        for (int k = 0; k < 3; ++k) {
            part[k] = lo[k];
            part[3 + k] = hi[k];
        }
}

__device__ inline uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// 30-bit Morton code of p on a 1024^3 grid over `box` (lo xyz, hi xyz),
// clamped to the grid.
__device__ inline uint64_t morton_code(d3 p, const double* __restrict__ box) {
    const double c[3] = {p.x, p.y, p.z};
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        const double extent = box[3 + k] - box[k];
        const double u = extent > 0.0 ? (c[k] - box[k]) / extent : 0.0;
        double s = u * 1024.0;
        s = s < 0.0 ? 0.0 : (s > 1023.0 ? 1023.0 : s);
        q[k] = static_cast<uint32_t>(s);
    }
    return (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
}

__global__ void morton_kernel(const double* __restrict__ c, int64_t nt, const double* __restrict__ box,
                              int index_bits, uint64_t* __restrict__ keys) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    keys[t] = (morton_code(ld3(c, t), box) << index_bits) | static_cast<uint64_t>(t);
}

__global__ void leaf_kernel(const uint64_t* __restrict__ keys, int64_t nt, uint64_t index_mask,
                            const double* __restrict__ p, const int32_t* __restrict__ f,
                            int32_t* __restrict__ leaf_triangle, double* __restrict__ corners) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const int32_t t = static_cast<int32_t>(keys[i] & index_mask);
    leaf_triangle[i] = t;
    double* c = corners + kLeafStride * i;
    for (int k = 0; k < 3; ++k) {
        const d3 v = ld3(p, f[3 * t + k]);
        c[3 * k] = v.x;
        c[3 * k + 1] = v.y;
        c[3 * k + 2] = v.z;
    }
    c[9] = static_cast<double>(t);  // exact: triangle indices are below 2^31
}

__global__ void key_index_kernel(const uint64_t* __restrict__ keys, int64_t n, uint64_t index_mask,
                                 int32_t* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<int32_t>(keys[i] & index_mask);
}

__device__ inline int delta_keys(const uint64_t* keys, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    return __clzll(static_cast<long long>(keys[i] ^ keys[j]));
}

// Karras (2012), "Maximizing parallelism in the construction of BVHs": the
// keys are unique (triangle index in the low bits), so the range and split of
// internal node i follow from common-prefix lengths alone.
__global__ void karras_kernel(const uint64_t* __restrict__ keys, int64_t n, BvhNode* __restrict__ nodes,
                              int32_t* __restrict__ leaf_parent) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta_keys(keys, n, i, i + 1) - delta_keys(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta_keys(keys, n, i, i - d);
    int64_t lmax = 2;
    while (delta_keys(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta_keys(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = delta_keys(keys, n, i, j);
    int64_t s = 0;
    int64_t t = l;
    do {
        t = (t + 1) >> 1;
        if (delta_keys(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    const int64_t lo = i < j ? i : j;
    const int64_t hi = i < j ? j : i;
    int32_t left, right;
    if (lo == gamma) {
        left = ~static_cast<int32_t>(gamma);
        leaf_parent[gamma] = static_cast<int32_t>(i);
    } else {
        left = static_cast<int32_t>(gamma);
        nodes[gamma].parent = static_cast<int32_t>(i);
    }
    if (hi == gamma + 1) {
        right = ~static_cast<int32_t>(gamma + 1);
        leaf_parent[gamma + 1] = static_cast<int32_t>(i);
    } else {
        right = static_cast<int32_t>(gamma + 1);
        nodes[gamma + 1].parent = static_cast<int32_t>(i);
    }
    nodes[i].child[0] = left;
    nodes[i].child[1] = right;
    if (i == 0) nodes[0].parent = -1;
}

// Bottom-up refit, one launch per binary depth from the deepest up: a node
// at depth d stores its two children's boxes, and its internal children (at
// depth d + 1) were completed by the previous launch, so no node waits on
// another inside a launch. A leaf's box is its triangle's FP64 box rounded
// outward to FP32 (it still contains the triangle); unions of FP32 boxes
// (min / max) are exact from there on.
__device__ inline void child_box(const BvhNode* __restrict__ nodes, const double* __restrict__ corners, int32_t c,
                                 float* lo, float* hi) {
    if (c < 0) {
        const double* t = corners + kLeafStride * static_cast<int64_t>(~c);
        for (int k = 0; k < 3; ++k) {
            lo[k] = __double2float_rd(fmin(fmin(t[k], t[3 + k]), t[6 + k]));
            hi[k] = __double2float_ru(fmax(fmax(t[k], t[3 + k]), t[6 + k]));
        }
    } else {
        const float* b = nodes[c].box;
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(b[k], b[6 + k]);
            hi[k] = fmaxf(b[3 + k], b[9 + k]);
        }
    }
}

__global__ void refit_level_kernel(const uint32_t* __restrict__ order, int64_t begin, int64_t end,
                                   const double* __restrict__ corners, BvhNode* nodes) {
    const int64_t t = begin + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= end) return;
    BvhNode* nd = nodes + order[t];
    for (int side = 0; side < 2; ++side) {
        float lo[3], hi[3];
        child_box(nodes, corners, nd->child[side], lo, hi);
        for (int k = 0; k < 3; ++k) {
            nd->box[6 * side + k] = lo[k];
            nd->box[6 * side + 3 + k] = hi[k];
        }
    }
}

// (depth, node) pairs for the level lists.
__global__ void depth_keys_kernel(const int32_t* __restrict__ depth, int64_t ni, uint64_t* __restrict__ keys,
                                  uint32_t* __restrict__ vals) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    keys[i] = static_cast<uint64_t>(depth[i]);
    vals[i] = static_cast<uint32_t>(i);
}

// [start, end) of each depth's run in the depth-sorted list.
__global__ void level_runs_kernel(const uint64_t* __restrict__ keys, int64_t ni, int32_t* __restrict__ runs) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    const uint64_t d = keys[i];
    if (i == 0 || keys[i - 1] != d) runs[2 * d] = static_cast<int32_t>(i);
    if (i == ni - 1 || keys[i + 1] != d) runs[2 * d + 1] = static_cast<int32_t>(i + 1);
}

// A Karras tree over unique keys of B bits is at most B + 1 levels deep
// (every level consumes a bit of common prefix); keys here have at most 61.
constexpr int kDepthBits = 7;
constexpr int kMaxDepth = 1 << kDepthBits;

// Binary depth of every internal node (parent links up to the root); the
// nodes at depths 0, 3, 6, ... root the wide nodes.
__global__ void depth_kernel(const BvhNode* __restrict__ nodes, int64_t ni, int32_t* __restrict__ depth,
                             int32_t* __restrict__ is_root) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    int32_t d = 0;
    int32_t n = static_cast<int32_t>(i);
    while (n > 0 && d < kMaxDepth - 1) {
        n = nodes[n].parent;
        ++d;
    }
    depth[i] = d;
    is_root[i] = d % 3 == 0 ? 1 : 0;
}

// Wide node of the binary node i (depth % 3 == 0): its descendants down to
// three levels below, stopping at leaves; internal nodes three levels down
// are themselves wide roots and are referenced by their wide index.
__global__ void collapse_kernel(const BvhNode* __restrict__ nodes, int64_t ni, const int32_t* __restrict__ depth,
                                const int32_t* __restrict__ wide_index, WideNode* __restrict__ wide) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ni || depth[i] % 3 != 0) return;
    int32_t ref[kWideArity];
    float box[kWideArity][6];
    int count = 0;
    {
        const BvhNode& root = nodes[i];
        for (int c = 0; c < 2; ++c) {
            ref[count] = root.child[c];
            for (int k = 0; k < 6; ++k) box[count][k] = root.box[6 * c + k];
            ++count;
        }
    }
    for (int level = 1; level < 3; ++level) {
        int32_t nref[kWideArity];
        float nbox[kWideArity][6];
        int m = 0;
        for (int e = 0; e < count; ++e) {
            const int32_t r = ref[e];
            if (r >= 0 && r < ni) {
                const BvhNode& c = nodes[r];
                for (int h = 0; h < 2; ++h) {
                    nref[m] = c.child[h];
                    for (int k = 0; k < 6; ++k) nbox[m][k] = c.box[6 * h + k];
                    ++m;
                }
            } else {
                nref[m] = r;
                for (int k = 0; k < 6; ++k) nbox[m][k] = box[e][k];
                ++m;
            }
        }
        count = m;
        for (int e = 0; e < count; ++e) {
            ref[e] = nref[e];
            for (int k = 0; k < 6; ++k) box[e][k] = nbox[e][k];
        }
    }
    WideNode& w = wide[wide_index[i]];
    for (int e = 0; e < kWideArity; ++e) {
        int32_t r = kWideEmpty;
        if (e < count) {
            r = ref[e];
            if (r >= 0) r = (r < ni && depth[r] % 3 == 0) ? wide_index[r] : kWideEmpty;
        }
        w.child[e] = r;
        w.pad[e] = 0;
        for (int k = 0; k < 3; ++k) {
            w.lo[k][e] = e < count ? box[e][k] : 0.0f;
            w.hi[k][e] = e < count ? box[e][3 + k] : 0.0f;
        }
    }
}

// ------------------------------------------------------------- query -------

// Ericson's Voronoi-region closest point with the CPU route's degenerate
// fallbacks; every operation in the CPU route's order.
__device__ inline d3 closest_on_triangle(d3 p, d3 a, d3 b, d3 c, double* bary) {
    const d3 ab = b - a, ac = c - a, ap = p - a;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        bary[0] = 1.0; bary[1] = 0.0; bary[2] = 0.0;
        return a;
    }
    const d3 bp = p - b;
    const double d3v = dot(ab, bp), d4 = dot(ac, bp);
    if (d3v >= 0.0 && d4 <= d3v) {
        bary[0] = 0.0; bary[1] = 1.0; bary[2] = 0.0;
        return b;
    }
    const double vc = d1 * d4 - d3v * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3v <= 0.0) {
        const double den = d1 - d3v;
        const double v = den != 0.0 ? d1 / den : 0.0;
        bary[0] = 1.0 - v; bary[1] = v; bary[2] = 0.0;
        return a + ab * v;
    }
    const d3 cp = p - c;
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
        bary[0] = 0.0; bary[1] = 0.0; bary[2] = 1.0;
        return c;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double den = d2 - d6;
        const double w = den != 0.0 ? d2 / den : 0.0;
        bary[0] = 1.0 - w; bary[1] = 0.0; bary[2] = w;
        return a + ac * w;
    }
    const double va = d3v * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3v) >= 0.0 && (d5 - d6) >= 0.0) {
        const double den = (d4 - d3v) + (d5 - d6);
        const double w = den != 0.0 ? (d4 - d3v) / den : 0.0;
        bary[0] = 0.0; bary[1] = 1.0 - w; bary[2] = w;
        return b + (c - b) * w;
    }
    const double den = va + vb + vc;
    if (fabs(den) < 1e-300) {
        bary[0] = 1.0; bary[1] = 0.0; bary[2] = 0.0;
        return a;
    }
    const double v = vb / den, w = vc / den;
    bary[0] = 1.0 - v - w; bary[1] = v; bary[2] = w;
    return a + ab * v + ac * w;
}

// Wide traversal, eight lanes per query. Lane j measures child j of the open
// node; the group tests its leaf children in parallel and keeps the
// lexicographic minimum of (distance^2, triangle), then pushes the surviving
// internal children on the group's LDS stack nearest on top, and pops the
// nearest entry still within reach. A box is skipped only when its distance
// exceeds the best squared distance times (1 + 1e-12) -- a margin over the
// rounding of the FP64 distances -- so every triangle that can tie or beat the
// best is tested, and the result is the lexicographic minimum over all
// triangles. Every decision is taken on values the eight lanes share (after
// shuffles), so the group stays converged.
//
// Stack size: an open pushes at most 8 entries and then pops one, so at
// most 7 entries per wide level of the current path stay below the newest
// push: a tree of W wide levels needs at most 7 W - 6 (stack_entries_needed).
// The query is instantiated for a few stack sizes and launched with the
// smallest that the built tree's depth allows. The stack is the query's LDS,
// and LDS is what limits the waves resident per compute unit for this
// latency-bound traversal: the smallest stack keeps the most waves in flight.
// The overflow check and the step budget only guard against a malformed tree
// (the query then reports triangle -1).
constexpr int kQueryThreads = 64;
constexpr int kLanes = kWideArity;
constexpr int kGroupsPerBlock = kQueryThreads / kLanes;
constexpr double kPruneSlack = 1.0 + 1e-12;
// Consecutive query blocks (neighbouring queries in the Morton order) kept
// on one XCD, in runs dealt to the XCDs in turn (xcd_chunked): the runs share
// their L2's tree nodes, and the turns keep the XCDs' loads even. Measured on
// the 200k wrap's queries: runs of 16 blocks 10.88-10.94 ms against
// 11.11-11.19 ms for single blocks in turn (4: 11.23-11.28, 64: 10.93-10.99;
// one contiguous eighth per XCD measured slower than single blocks).
constexpr int64_t kQueryChunk = 16;

// Binary depths 0, 3, 6, ... root the wide nodes.
int64_t stack_entries_needed(int32_t depth_max) {
    const int64_t levels = depth_max / 3 + 1;
    return std::max<int64_t>(8, 7 * levels - 6);
}

struct Nearest {
    double d2;
    int32_t triangle;
    uint32_t opens, tests;  // traversal work, for profiling
};

// Moves within a group of eight lanes by DPP (register to register): lane j
// receives lane j ^ k. Quad permutations give k = 1, 2, 3; the half-row
// mirror gives k = 7; a quad permutation of the mirrored value gives
// k = 4, 5, 6. Groups are converged (all eight lanes active together), so
// every source lane is live.
constexpr int kDppQuadSwap3 = 0x1B;  // quad_perm [3, 2, 1, 0]

template <int K, typename T>
__device__ inline T octet_xor(T v) {
    static_assert(K >= 1 && K <= 7, "partner within the octet");
    if constexpr (K == 1) return dpp_move<kDppQuadSwap1>(v);
    if constexpr (K == 2) return dpp_move<kDppQuadSwap2>(v);
    if constexpr (K == 3) return dpp_move<kDppQuadSwap3>(v);
    if constexpr (K == 7) return dpp_move<kDppRowHalfMirror>(v);
    if constexpr (K == 4) return dpp_move<kDppQuadSwap3>(dpp_move<kDppRowHalfMirror>(v));
    if constexpr (K == 5) return dpp_move<kDppQuadSwap2>(dpp_move<kDppRowHalfMirror>(v));
    if constexpr (K == 6) return dpp_move<kDppQuadSwap1>(dpp_move<kDppRowHalfMirror>(v));
    return v;
}

template <int K>
__device__ inline void group_min_step(double& d2, int32_t& tri) {
    const double od = octet_xor<K>(d2);
    const int32_t ot = octet_xor<K>(tri);
    if (od < d2 || (od == d2 && ot < tri)) {
        d2 = od;
        tri = ot;
    }
}

// Lexicographic (d2, triangle) minimum over the group's eight lanes: pairs,
// then quads, then the two quads (the order is total, so every lane keeps
// the same winner).
__device__ inline void group_min(double& d2, int32_t& tri) {
    group_min_step<1>(d2, tri);
    group_min_step<2>(d2, tri);
    group_min_step<7>(d2, tri);
}

__device__ inline int group_sum(int v) { return octet_sum(v); }

// How many lanes of the group come before lane j in the (key, lane) order
// (non-pushing lanes carry +inf keys, so a pushing lane counts only pushing
// lanes).
template <int K>
__device__ inline void rank_step(float key, int j, int& rank) {
    const float ok = octet_xor<K>(key);
    const int other = j ^ K;
    if (ok < key || (ok == key && other < j)) ++rank;
}

template <int StackDepth>
__device__ __forceinline__ void nearest_wide(d3 p, const WideNode* __restrict__ wide, int64_t wide_count,
                                    const double* __restrict__ corners,
                                    int32_t* stack_node, float* stack_dist, int j, Nearest& best) {
    int sp = 0;
    int32_t node = 0;
    int64_t budget = 2 * wide_count + 64;
    for (;;) {
        if (--budget < 0) {
            best.triangle = -1;
            return;
        }
        ++best.opens;
        const WideNode* w = wide + node;
        // The box is loaded and measured with the reference, not behind a
        // test of it (an empty slot's box is in the same node and its
        // distance is replaced by infinity), so an open costs one memory
        // round trip instead of two.
        const int32_t ref = w->child[j];
        const float b[6] = {w->lo[0][j], w->lo[1][j], w->lo[2][j], w->hi[0][j], w->hi[1][j], w->hi[2][j]};
        const double cx = fmin(fmax(p.x, static_cast<double>(b[0])), static_cast<double>(b[3]));
        const double cy = fmin(fmax(p.y, static_cast<double>(b[1])), static_cast<double>(b[4]));
        const double cz = fmin(fmax(p.z, static_cast<double>(b[2])), static_cast<double>(b[5]));
        const double box_d = norm2(d3{p.x - cx, p.y - cy, p.z - cz});
        const double d = ref != kWideEmpty ? box_d : __builtin_inf();
        // Leaf children within reach: tested in parallel, then merged.
        double cd2 = best.d2;
        int32_t ctri = best.triangle;
        if (ref != kWideEmpty && ref < 0 && !(d > best.d2 * kPruneSlack)) {
            const int32_t leaf = ~ref;
            // The record (corners, then the triangle id) as five 16-byte
            // loads: the id shares one with the last coordinate, so it
            // arrives with the corners instead of a round trip later.
            const double2* r = reinterpret_cast<const double2*>(corners + kLeafStride * static_cast<int64_t>(leaf));
            const double2 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3], r4 = r[4];
            double bary[3];
            const d3 q = closest_on_triangle(p, d3{r0.x, r0.y, r1.x}, d3{r1.y, r2.x, r2.y}, d3{r3.x, r3.y, r4.x},
                                             bary);
            const double d2 = norm2(q - p);
            const int32_t tri = static_cast<int32_t>(r4.y);
            if (d2 < cd2 || (d2 == cd2 && tri < ctri)) {
                cd2 = d2;
                ctri = tri;
            }
            ++best.tests;
        }
        group_min(cd2, ctri);
        best.d2 = cd2;
        best.triangle = ctri;
        // Internal children still within reach, pushed nearest on top.
        const bool push = ref != kWideEmpty && ref >= 0 && !(d > best.d2 * kPruneSlack);
        // The push order (nearest on top) only steers the search, never its
        // result, so it is ranked on the FP32 distance the stack keeps (its
        // rounding down, finite for every box); lanes that do not push carry
        // +inf and rank after every pushing lane.
        const float key = push ? __double2float_rd(d) : __builtin_inff();
        int rank = 0;
        rank_step<1>(key, j, rank);
        rank_step<2>(key, j, rank);
        rank_step<3>(key, j, rank);
        rank_step<4>(key, j, rank);
        rank_step<5>(key, j, rank);
        rank_step<6>(key, j, rank);
        rank_step<7>(key, j, rank);
        const int npush = group_sum(push ? 1 : 0);
        if (sp + npush > StackDepth) {
            best.triangle = -1;
            return;
        }
        if (push) {
            const int at = sp + npush - 1 - rank;
            stack_node[at] = ref;
            stack_dist[at] = key;
        }
        sp += npush;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        node = -1;
        while (sp > 0) {
            --sp;
            const int32_t c = stack_node[sp];
            const double dd = static_cast<double>(stack_dist[sp]);
            if (dd > best.d2 * kPruneSlack) continue;
            node = c;
            break;
        }
        if (node < 0) return;
    }
}


// Eight lanes per query: mask, nearest triangle, interpolated normal, the
// distance / normal / slope filters in the CPU route's order
// (src/solver/correspondence.cpp:33-59), then the IRLS weight of
// src/solver/route_a.cpp:19-38 for accepted rows (written by the group's
// first lane).
template <int StackDepth>
__device__ __forceinline__ void query_block(const WideNode* __restrict__ wide, int64_t wide_count, int64_t nt,
                                   const double* __restrict__ corners, const int32_t* __restrict__ leaf_triangle,
                                   const int32_t* __restrict__ faces, const double* __restrict__ vnormals,
                                   const double* __restrict__ queries, const double* __restrict__ qnormals,
                                   const uint8_t* __restrict__ mask, int64_t count, const QuerySettings& s,
                                   const double* __restrict__ tpoints, const int32_t* __restrict__ warm,
                                   double* __restrict__ out_p, double* __restrict__ out_n, double* __restrict__ out_d,
                                   int32_t* __restrict__ out_t, uint8_t* __restrict__ out_rej,
                                   double* __restrict__ out_u, const int32_t* __restrict__ order,
                                   uint32_t* __restrict__ stats, int64_t block,
                                   int32_t (&stack_node)[kGroupsPerBlock][StackDepth],
                                   float (&stack_dist)[kGroupsPerBlock][StackDepth]) {
    const int j = threadIdx.x % kLanes;
    const int group = threadIdx.x / kLanes;
    const int64_t slot = (block * kQueryThreads + threadIdx.x) / kLanes;
    if (slot >= count) return;  // the whole group leaves together
    const int64_t i = order ? static_cast<int64_t>(order[slot]) : slot;
    const bool writer = j == 0;
    if (mask && mask[i] == 0) {
        if (writer) {
            if (stats) stats[2 * slot] = stats[2 * slot + 1] = 0;
            st3(out_p, i, d3{0.0, 0.0, 0.0});
            st3(out_n, i, d3{0.0, 0.0, 0.0});
            out_d[i] = 0.0;
            out_t[i] = -1;
            out_rej[i] = 4;
            if (out_u) out_u[i] = 1.0;
        }
        return;
    }
    const d3 q = ld3(queries, i);
    Nearest best;
    best.d2 = __builtin_inf();
    best.triangle = 0x7fffffff;
    best.opens = best.tests = 0;
    // Warm start: the triangle this query matched last time is an exact
    // upper bound. The search still returns the lexicographic minimum of
    // (distance^2, triangle) -- boxes at exactly the bound are still visited --
    // so the result is the one a cold search gives, found with less work.
    const int32_t seed = warm ? warm[i] : -1;
    if (seed >= 0 && seed < nt) {
        const int32_t* sf = faces + 3 * static_cast<int64_t>(seed);
        double bary[3];
        const d3 c = closest_on_triangle(q, ld3(tpoints, sf[0]), ld3(tpoints, sf[1]), ld3(tpoints, sf[2]), bary);
        best.d2 = norm2(c - q);
        best.triangle = seed;
    }
    if (wide_count > 0) {
        nearest_wide<StackDepth>(q, wide, wide_count, corners, stack_node[group], stack_dist[group], j, best);
    } else if (nt == 1) {
        const double* c = corners;
        double bary[3];
        const d3 cp = closest_on_triangle(q, d3{c[0], c[1], c[2]}, d3{c[3], c[4], c[5]}, d3{c[6], c[7], c[8]}, bary);
        const double d2 = norm2(cp - q);
        const int32_t tri = leaf_triangle[0];
        if (d2 < best.d2 || (d2 == best.d2 && tri < best.triangle)) {
            best.d2 = d2;
            best.triangle = tri;
        }
    }
    if (stats) {
        const uint32_t tests = static_cast<uint32_t>(group_sum(static_cast<int>(best.tests)));
        if (writer) {
            stats[2 * slot] = best.opens;
            stats[2 * slot + 1] = tests;
        }
    }
    if (!writer) return;
    if (best.triangle < 0 || best.triangle >= nt) {
        // Only a malformed hierarchy gets here: report the row as unusable.
        st3(out_p, i, q);
        st3(out_n, i, d3{0.0, 0.0, 0.0});
        out_d[i] = __builtin_inf();
        out_t[i] = -1;
        out_rej[i] = 1;
        if (out_u) out_u[i] = 1.0;
        return;
    }
    // The winner's point and barycentrics, recomputed from its corners in
    // the original face order (the same operations as the test that chose it).
    const int32_t* f = faces + 3 * static_cast<int64_t>(best.triangle);
    double bary[3];
    const d3 point = closest_on_triangle(q, ld3(tpoints, f[0]), ld3(tpoints, f[1]), ld3(tpoints, f[2]), bary);
    const d3 na = ld3(vnormals, f[0]), nb = ld3(vnormals, f[1]), nc = ld3(vnormals, f[2]);
    const d3 normal = unit((na * bary[0] + nb * bary[1]) + nc * bary[2]);
    const double distance = sqrt(best.d2);
    uint8_t rej = 0;
    if (distance > s.distance_threshold) {
        rej = 1;
    } else if (qnormals && dot(ld3(qnormals, i), normal) < s.min_normal_cosine) {
        rej = 2;
    } else if (distance > 0.0) {
        const double slope = fabs(dot(normal, q - point)) / distance;
        if (slope < s.min_slope) rej = 3;
    }
    st3(out_p, i, point);
    st3(out_n, i, normal);
    out_d[i] = distance;
    out_t[i] = best.triangle;
    out_rej[i] = rej;
    if (out_u) {
        const d3 off = q - point;
        out_u[i] = (rej == 0 && s.robust_threshold > 0.0)
                       ? robust_weight(s.robust_kernel, norm2(off), s.robust_threshold)
                       : 1.0;
    }
}

template <int StackDepth>
__global__ void __launch_bounds__(kQueryThreads)
query_kernel(const WideNode* __restrict__ wide, int64_t wide_count, int64_t nt, const double* __restrict__ corners,
             const int32_t* __restrict__ leaf_triangle, const int32_t* __restrict__ faces,
             const double* __restrict__ vnormals, const double* __restrict__ queries,
             const double* __restrict__ qnormals, const uint8_t* __restrict__ mask, int64_t count,
             QuerySettings s, const double* __restrict__ tpoints, const int32_t* __restrict__ warm,
             double* __restrict__ out_p, double* __restrict__ out_n, double* __restrict__ out_d,
             int32_t* __restrict__ out_t, uint8_t* __restrict__ out_rej, double* __restrict__ out_u,
             const int32_t* __restrict__ order, uint32_t* __restrict__ stats) {
    __shared__ int32_t stack_node[kGroupsPerBlock][StackDepth];
    __shared__ float stack_dist[kGroupsPerBlock][StackDepth];
    query_block<StackDepth>(wide, wide_count, nt, corners, leaf_triangle, faces, vnormals, queries, qnormals, mask,
                            count, s, tpoints, warm, out_p, out_n, out_d, out_t, out_rej, out_u, order, stats,
                            xcd_chunked(blockIdx.x, gridDim.x, kQueryChunk), stack_node, stack_dist);
}

// ---------------------------------------------------------------- packed ---
// Normals and correspondences of the jobs that refresh this tick (packed.hpp).

__global__ void __launch_bounds__(256)
normals_packed(const PackedJob* __restrict__ jobs, const BatchControl* __restrict__ ctl,
               const PackedBlock* __restrict__ map) {
    const PackedBlock m = map[xcd_contiguous(blockIdx.x, gridDim.x)];
    if (!(ctl[m.job].flags & kJobRefresh)) return;
    const PackedJob& J = jobs[m.job];
    vertex_normal(J.deformed, J.faces, J.vf_offsets, J.vf_faces, J.in.vertex_count, J.query_normals,
                  static_cast<int64_t>(m.block) * 256 + threadIdx.x);
}

template <int StackDepth>
__global__ void __launch_bounds__(kQueryThreads)
query_packed(const PackedJob* __restrict__ jobs, const BatchControl* __restrict__ ctl,
             const PackedBlock* __restrict__ map) {
    __shared__ int32_t stack_node[kGroupsPerBlock][StackDepth];
    __shared__ float stack_dist[kGroupsPerBlock][StackDepth];
    // Blocks in launch order, as the single-job query takes them (an
    // XCD-contiguous order measured slower for the traversal).
    const PackedBlock m = map[blockIdx.x];
    if (!(ctl[m.job].flags & kJobRefresh)) return;
    const PackedJob& J = jobs[m.job];
    // Every refresh after the rest-template pass starts from the last
    // triangles (the single-job session's warm start).
    query_block<StackDepth>(J.wide, J.wide_count, J.triangle_count, J.leaf_corners, J.leaf_triangle, J.target_faces,
                            J.target_normals, J.deformed, J.query_normals, J.mask, J.in.vertex_count, J.query,
                            J.target_points, J.corr_triangle, J.corr_points, J.corr_normals, J.corr_distance,
                            J.corr_triangle, J.rejection, J.robust, J.query_order, nullptr, m.block, stack_node,
                            stack_dist);
}

// Traversal statistics of one launch (profiling): per-query (opens, tests)
// folded into running totals [sum opens, sum tests, max opens].
__global__ void __launch_bounds__(1024) query_stats_fold(const uint32_t* __restrict__ rows, int64_t count,
                                                        unsigned long long* __restrict__ totals) {
    __shared__ unsigned long long lds[3][1024];
    unsigned long long o = 0, t = 0, m = 0;
    for (int64_t i = threadIdx.x; i < count; i += 1024) {
        const unsigned long long a = rows[2 * i], b = rows[2 * i + 1];
        o += a;
        t += b;
        m = a > m ? a : m;
    }
    lds[0][threadIdx.x] = o;
    lds[1][threadIdx.x] = t;
    lds[2][threadIdx.x] = m;
    __syncthreads();
    for (int span = 512; span > 0; span >>= 1) {
        if (threadIdx.x < span) {
            lds[0][threadIdx.x] += lds[0][threadIdx.x + span];
            lds[1][threadIdx.x] += lds[1][threadIdx.x + span];
            const unsigned long long other = lds[2][threadIdx.x + span];
            if (other > lds[2][threadIdx.x]) lds[2][threadIdx.x] = other;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        totals[0] += lds[0][0];
        totals[1] += lds[1][0];
        if (lds[2][0] > totals[2]) totals[2] = lds[2][0];
    }
}

__global__ void __launch_bounds__(256)
summary_partial(const uint8_t* __restrict__ rej, const double* __restrict__ dist, int64_t n, double* __restrict__ part) {
    __shared__ double lds[6 * 4];
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
        const int r = rej[i];
        if (r == 0) {
            v[0] += 1.0;
            v[5] += dist[i];
        } else if (r <= 4) {
            v[r] += 1.0;
        }
    }
    double out[6];
    block_sum<6, 256>(v, lds, out);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = out[k];
}

__global__ void summary_final(const double* __restrict__ part, int blocks, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < blocks; ++b)
        for (int k = 0; k < 6; ++k) s[k] += part[6 * b + k];
    for (int k = 0; k < 6; ++k) out[k] = s[k];
}

__global__ void distance_keys(const uint8_t* __restrict__ rej, const double* __restrict__ dist, int64_t n,
                              uint64_t* __restrict__ keys) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // Non-negative doubles order like their bit patterns; rejected rows sort last.
    keys[i] = rej[i] == 0 ? static_cast<uint64_t>(__double_as_longlong(dist[i])) : ~0ull;
}

__global__ void pick_quantiles(const uint64_t* __restrict__ keys, int64_t accepted, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (accepted <= 0) {
        out[0] = 0.0;
        out[1] = 0.0;
        return;
    }
    const double pos = 0.95 * static_cast<double>(accepted - 1);
    int64_t k = static_cast<int64_t>(pos < static_cast<double>(accepted - 1) ? pos : static_cast<double>(accepted - 1));
    out[0] = __longlong_as_double(static_cast<long long>(keys[k]));
    out[1] = __longlong_as_double(static_cast<long long>(keys[accepted - 1]));
}

// The query kernel's IRLS weight, elementwise over given squared residuals.
__global__ void robust_weights_kernel(int kind, const double* __restrict__ squared, int64_t n, double t,
                                      double* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = robust_weight(kind, squared[i], t);
}

}  // namespace

void build_vertex_faces(const int32_t* faces, int64_t nf, int64_t nv, VertexFaces& out, SortScratch& scratch,
                        hipStream_t stream) {
    out.keys.resize(3 * nf);
    out.vals.resize(3 * nf);
    uint64_t* keys = out.keys.get();
    uint32_t* vals = out.vals.get();
    if (nf > 0) {
        vertex_face_entries<<<blocks_for(3 * nf, 256), 256, 0, stream>>>(faces, nf, keys, vals);
        NRICP_LAUNCH_CHECK("vertex_face_entries");
        radix_sort_pairs(keys, vals, 3 * nf, bits_for(static_cast<uint64_t>(nv)), scratch, stream);
    }
    out.offsets.resize(nv + 1);
    out.faces.resize(3 * nf);
    offsets_from_sorted(keys, 3 * nf, 0, ~0ull, nv, out.offsets.get(), stream);
    if (nf > 0) {
        copy_u32_to_i32<<<blocks_for(3 * nf, 256), 256, 0, stream>>>(vals, 3 * nf, out.faces.get());
        NRICP_LAUNCH_CHECK("copy_u32_to_i32");
    }
}

void vertex_normals(const double* points, const int32_t* faces, const VertexFaces& vf, int64_t nv, double* normals,
                    hipStream_t stream) {
    if (nv <= 0) return;
    vertex_normals_kernel<<<blocks_for(nv, 256), 256, 0, stream>>>(points, faces, vf.offsets.get(), vf.faces.get(), nv,
                                                                    normals);
    NRICP_LAUNCH_CHECK("vertex_normals_kernel");
}

void build_lbvh(Lbvh& bvh, int64_t nv, SortScratch& scratch, VertexFaces& vf, hipStream_t stream) {
    const int64_t nt = bvh.triangle_count;
    if (nt <= 0) throw InvalidArgument("the target has no triangles");
    if (nt >= (1ll << 31)) throw InvalidArgument("too many target triangles");
    // Target vertex normals (face order per vertex, as the CPU route).
    build_vertex_faces(bvh.faces.get(), nt, nv, vf, scratch, stream);
    bvh.vertex_normals.resize(3 * nv);
    vertex_normals(bvh.points.get(), bvh.faces.get(), vf, nv, bvh.vertex_normals.get(), stream);
    // Morton keys of the centroids, triangle index in the low bits.
    bvh.centroids.resize(3 * nt);
    centroid_kernel<<<blocks_for(nt, 256), 256, 0, stream>>>(bvh.points.get(), bvh.faces.get(), nt, bvh.centroids.get());
    NRICP_LAUNCH_CHECK("centroid_kernel");
    const int parts = static_cast<int>(std::min<int64_t>(blocks_for(nt, 256), 256));
    bvh.bounds.resize(6 * parts);
    bounds_partial<<<parts, 256, 0, stream>>>(bvh.centroids.get(), nt, bvh.bounds.get());
    NRICP_LAUNCH_CHECK("bounds_partial");
    bounds_final<<<1, 64, 0, stream>>>(bvh.bounds.get(), parts);
    NRICP_LAUNCH_CHECK("bounds_final");
    const int index_bits = std::max(1, bits_for(static_cast<uint64_t>(nt - 1)));
    bvh.index_bits = index_bits;
    bvh.keys.resize(nt);
    morton_kernel<<<blocks_for(nt, 256), 256, 0, stream>>>(bvh.centroids.get(), nt, bvh.bounds.get(), index_bits,
                                                            bvh.keys.get());
    NRICP_LAUNCH_CHECK("morton_kernel");
    radix_sort_pairs(bvh.keys.get(), nullptr, nt, 30 + index_bits, scratch, stream);
    bvh.leaf_triangle.resize(nt);
    bvh.leaf_corners.resize(kLeafStride * nt);
    leaf_kernel<<<blocks_for(nt, 256), 256, 0, stream>>>(bvh.keys.get(), nt, (1ull << index_bits) - 1,
                                                          bvh.points.get(), bvh.faces.get(), bvh.leaf_triangle.get(),
                                                          bvh.leaf_corners.get());
    NRICP_LAUNCH_CHECK("leaf_kernel");
    bvh.leaf_parent.resize(nt);
    bvh.nodes.resize(std::max<int64_t>(nt - 1, 1));
    if (nt > 1) {
        karras_kernel<<<blocks_for(nt - 1, 256), 256, 0, stream>>>(bvh.keys.get(), nt, bvh.nodes.get(),
                                                                    bvh.leaf_parent.get());
        NRICP_LAUNCH_CHECK("karras_kernel");
        // Depths (from the parent links) give the wide roots (depth % 3 == 0)
        // and the refit's level lists; one read-back brings the wide count
        // and the level runs.
        const int64_t ni = nt - 1;
        bvh.flags.resize(ni);
        bvh.depth.resize(ni);
        bvh.wide_index.resize(ni + 1);
        depth_kernel<<<blocks_for(ni, 256), 256, 0, stream>>>(bvh.nodes.get(), ni, bvh.depth.get(), bvh.flags.get());
        NRICP_LAUNCH_CHECK("depth_kernel");
        exclusive_scan32(bvh.flags.get(), bvh.wide_index.get(), ni, scratch, stream);
        bvh.level_order.resize(ni);
        bvh.level_keys.resize(ni);
        depth_keys_kernel<<<blocks_for(ni, 256), 256, 0, stream>>>(bvh.depth.get(), ni, bvh.level_keys.get(),
                                                                    bvh.level_order.get());
        NRICP_LAUNCH_CHECK("depth_keys_kernel");
        radix_sort_pairs(bvh.level_keys.get(), bvh.level_order.get(), ni, kDepthBits, scratch, stream);
        bvh.level_runs.resize(2 * kMaxDepth);
        bvh.level_runs.zero(stream);
        level_runs_kernel<<<blocks_for(ni, 256), 256, 0, stream>>>(bvh.level_keys.get(), ni, bvh.level_runs.get());
        NRICP_LAUNCH_CHECK("level_runs_kernel");
        std::vector<int32_t> runs(2 * kMaxDepth + 1);
        NRICP_CHECK(hipMemcpyAsync(runs.data(), bvh.level_runs.get(), 2 * kMaxDepth * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, stream));
        NRICP_CHECK(hipMemcpyAsync(runs.data() + 2 * kMaxDepth, bvh.wide_index.get() + ni, sizeof(int32_t),
                                   hipMemcpyDeviceToHost, stream));
        NRICP_CHECK(hipStreamSynchronize(stream));
        for (int d = kMaxDepth - 1; d >= 0; --d) {
            const int64_t begin = runs[2 * d], end = runs[2 * d + 1];
            if (end <= begin) continue;
            refit_level_kernel<<<blocks_for(end - begin, 256), 256, 0, stream>>>(
                bvh.level_order.get(), begin, end, bvh.leaf_corners.get(), bvh.nodes.get());
            NRICP_LAUNCH_CHECK("refit_level_kernel");
        }
        const int32_t wide_count = runs[2 * kMaxDepth];
        bvh.wide_count = wide_count;
        bvh.depth_max = 0;
        for (int d = kMaxDepth - 1; d >= 0; --d) {
            if (runs[2 * d + 1] > runs[2 * d]) {
                bvh.depth_max = d;
                break;
            }
        }
        bvh.wide.resize(std::max<int64_t>(wide_count, 1));
        collapse_kernel<<<blocks_for(ni, 128), 128, 0, stream>>>(bvh.nodes.get(), ni, bvh.depth.get(),
                                                                 bvh.wide_index.get(), bvh.wide.get());
        NRICP_LAUNCH_CHECK("collapse_kernel");
    } else {
        bvh.wide_count = 0;
        bvh.depth_max = 0;
    }
}

void morton_order(const double* points, int64_t n, int32_t* out, SortScratch& scratch, DeviceBuffer<uint64_t>& keys,
                  DeviceBuffer<double>& bounds, hipStream_t stream) {
    if (n <= 0) return;
    if (n >= (1ll << 31)) throw InvalidArgument("too many points to order");
    const int parts = static_cast<int>(std::min<int64_t>(blocks_for(n, 256), 256));
    bounds.resize(6 * parts);
    bounds_partial<<<parts, 256, 0, stream>>>(points, n, bounds.get());
    NRICP_LAUNCH_CHECK("bounds_partial");
    bounds_final<<<1, 64, 0, stream>>>(bounds.get(), parts);
    NRICP_LAUNCH_CHECK("bounds_final");
    const int index_bits = std::max(1, bits_for(static_cast<uint64_t>(n - 1)));
    keys.resize(n);
    morton_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(points, n, bounds.get(), index_bits, keys.get());
    NRICP_LAUNCH_CHECK("morton_kernel");
    radix_sort_pairs(keys.get(), nullptr, n, 30 + index_bits, scratch, stream);
    key_index_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(keys.get(), n, (1ull << index_bits) - 1, out);
    NRICP_LAUNCH_CHECK("key_index_kernel");
}

void closest_points(const Lbvh& bvh, const double* queries, const double* query_normals, const uint8_t* mask,
                    int64_t count, const QuerySettings& s, double* out_points, double* out_normals,
                    double* out_distance, int32_t* out_triangle, uint8_t* out_rejection, double* out_robust,
                    hipStream_t stream, const int32_t* warm, const int32_t* order, QueryStats* stats) {
    if (count <= 0) return;
    const int64_t need = stack_entries_needed(bvh.depth_max);
    auto launch = [&](auto kernel) {
        kernel<<<blocks_for(count * kLanes, kQueryThreads), kQueryThreads, 0, stream>>>(
            bvh.wide.get(), bvh.wide_count, bvh.triangle_count, bvh.leaf_corners.get(), bvh.leaf_triangle.get(),
            bvh.faces.get(), bvh.vertex_normals.get(), queries, query_normals, mask, count, s, bvh.points.get(), warm,
            out_points, out_normals, out_distance, out_triangle, out_rejection, out_robust, order,
            stats ? stats->rows.get() : nullptr);
        NRICP_LAUNCH_CHECK("query_kernel");
    };
    // The depth cap of the tree build (kMaxDepth - 1) needs 7 * 43 - 6 = 295.
    if (need <= 64)
        launch(query_kernel<64>);
    else if (need <= 96)
        launch(query_kernel<96>);
    else if (need <= 128)
        launch(query_kernel<128>);
    else
        launch(query_kernel<296>);
}

int32_t packed_query_stack(int32_t depth_max) {
    const int64_t need = stack_entries_needed(depth_max);
    return need <= 64 ? 64 : need <= 96 ? 96 : need <= 128 ? 128 : 296;
}

void packed_normals(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, KernelSink& sink) {
    if (maps.vertex_blocks <= 0) return;
    sink_launch(sink, normals_packed, static_cast<unsigned>(maps.vertex_blocks), 256, jobs, ctl, maps.vertex);
}

void packed_query(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, KernelSink& sink) {
    if (maps.query_blocks <= 0) return;
    const unsigned blocks = static_cast<unsigned>(maps.query_blocks);
    switch (maps.query_stack) {
        case 64:
            sink_launch(sink, query_packed<64>, blocks, kQueryThreads, jobs, ctl, maps.query);
            break;
        case 96:
            sink_launch(sink, query_packed<96>, blocks, kQueryThreads, jobs, ctl, maps.query);
            break;
        case 128:
            sink_launch(sink, query_packed<128>, blocks, kQueryThreads, jobs, ctl, maps.query);
            break;
        default:
            sink_launch(sink, query_packed<296>, blocks, kQueryThreads, jobs, ctl, maps.query);
            break;
    }
}

void fold_query_stats(QueryStats& stats, int64_t count, hipStream_t stream) {
    if (count <= 0) return;
    query_stats_fold<<<1, 1024, 0, stream>>>(stats.rows.get(), count,
                                             reinterpret_cast<unsigned long long*>(stats.totals.get()));
    NRICP_LAUNCH_CHECK("query_stats_fold");
}

void summarize_correspondences(const uint8_t* rejection, const double* distance, int64_t count,
                               DeviceBuffer<double>& partials, double* out, hipStream_t stream) {
    const int blocks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(blocks_for(count, 256), 240)));
    partials.resize(6 * blocks);
    summary_partial<<<blocks, 256, 0, stream>>>(rejection, distance, count, partials.get());
    NRICP_LAUNCH_CHECK("summary_partial");
    summary_final<<<1, 64, 0, stream>>>(partials.get(), blocks, out);
    NRICP_LAUNCH_CHECK("summary_final");
}

void distance_quantiles(const uint8_t* rejection, const double* distance, int64_t count, DeviceBuffer<uint64_t>& keys,
                        SortScratch& scratch, double* out, int64_t accepted, hipStream_t stream) {
    if (count <= 0) return;
    keys.resize(count);
    distance_keys<<<blocks_for(count, 256), 256, 0, stream>>>(rejection, distance, count, keys.get());
    NRICP_LAUNCH_CHECK("distance_keys");
    radix_sort_pairs(keys.get(), nullptr, count, 64, scratch, stream);
    pick_quantiles<<<1, 64, 0, stream>>>(keys.get(), accepted, out);
    NRICP_LAUNCH_CHECK("pick_quantiles");
}

void robust_weights(int kind, const double* squared, int64_t count, double threshold, double* out,
                    hipStream_t stream) {
    if (count <= 0) return;
    robust_weights_kernel<<<blocks_for(count, 256), 256, 0, stream>>>(kind, squared, count, threshold, out);
    NRICP_LAUNCH_CHECK("robust_weights_kernel");
}

}  // namespace nricp_hip
