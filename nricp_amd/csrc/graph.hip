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
#include "packed.hpp"
#include "so3.hpp"
#include "xcd_map.hpp"

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

__device__ __forceinline__ void evaluate_vertex(const double* __restrict__ p, const int32_t* __restrict__ bn,
                                       const double* __restrict__ bw, const int32_t* __restrict__ bc, int64_t nv,
                                       const double* __restrict__ R, const double* __restrict__ T,
                                       const double* __restrict__ G, int64_t nn, double* __restrict__ out, int64_t v) {
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

__global__ void evaluate_kernel(const double* __restrict__ p, const int32_t* __restrict__ bn,
                                const double* __restrict__ bw, const int32_t* __restrict__ bc, int64_t nv,
                                const double* __restrict__ R, const double* __restrict__ T,
                                const double* __restrict__ G, int64_t nn, double* __restrict__ out) {
    evaluate_vertex(p, bn, bw, bc, nv, R, T, G, nn, out, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
}

__device__ __forceinline__ void retract_node(const double* __restrict__ R, const double* __restrict__ T,
                                    const double* __restrict__ delta, int64_t nn, double* __restrict__ Ro,
                                    double* __restrict__ To, int64_t j) {
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

__global__ void retract_kernel(const double* __restrict__ R, const double* __restrict__ T,
                               const double* __restrict__ delta, int64_t nn, double* __restrict__ Ro,
                               double* __restrict__ To) {
    retract_node(R, T, delta, nn, Ro, To, static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------- packed ---
// The same per-vertex and per-node work for a batch of jobs (packed.hpp):
// workgroup b takes the block the stage's map gives it (in the XCD-contiguous
// order), and returns at once when its job has no work in the stage.

__global__ void __launch_bounds__(256)
evaluate_packed(const PackedJob* __restrict__ jobs, const BatchControl* __restrict__ ctl,
                const PackedBlock* __restrict__ map) {
    const PackedBlock m = map[xcd_contiguous(blockIdx.x, gridDim.x)];
    if (!(ctl[m.job].flags & kJobRefresh)) return;
    const PackedJob& J = jobs[m.job];
    evaluate_vertex(J.in.points, J.in.bind_nodes, J.in.bind_weights, J.in.bind_counts, J.in.vertex_count, J.rot,
                    J.trans, J.in.rest, J.in.node_count, J.deformed, static_cast<int64_t>(m.block) * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(128)
retract_packed(const PackedJob* __restrict__ jobs, const BatchControl* __restrict__ ctl,
               const PackedBlock* __restrict__ map) {
    const PackedBlock m = map[xcd_contiguous(blockIdx.x, gridDim.x)];
    if (!(ctl[m.job].flags & kJobActive)) return;
    const PackedJob& J = jobs[m.job];
    retract_node(J.rot, J.trans, J.delta, J.in.node_count, J.rot_cand, J.trans_cand,
                 static_cast<int64_t>(m.block) * 128 + threadIdx.x);
}

// An accepted candidate becomes the state.
__global__ void __launch_bounds__(128)
commit_packed(const PackedJob* __restrict__ jobs, const BatchControl* __restrict__ ctl,
              const PackedBlock* __restrict__ map) {
    const PackedBlock m = map[xcd_contiguous(blockIdx.x, gridDim.x)];
    if (!(ctl[m.job].flags & kJobAccepted)) return;
    const PackedJob& J = jobs[m.job];
    const int64_t j = static_cast<int64_t>(m.block) * 128 + threadIdx.x;
    if (j >= J.in.node_count) return;
    for (int i = 0; i < 9; ++i) J.rot[9 * j + i] = J.rot_cand[9 * j + i];
    for (int i = 0; i < 3; ++i) J.trans[3 * j + i] = J.trans_cand[3 * j + i];
}

// ---------------------------------------------------------- sampling -------
//
// Farthest-point sampling is one long dependent chain: each seed is the point
// farthest from all earlier seeds, so seed s + 1 can only be chosen once seed
// s has updated every running distance. The chain runs in ONE workgroup, so a
// step costs a workgroup barrier instead of a grid barrier, and it only
// touches the points a new seed can change:
//
//   * The points are stored in the Morton order of their positions, in blocks
//     of 64 x ppl. Each block keeps an outward-rounded box and the maximum of
//     its points' running distances (with that point's index and position).
//   * Lane l of wave w owns block l * 16 + w (neighbouring blocks go to
//     different waves) and keeps its summary in registers.
//   * Step: every wave reduces its lanes' block maxima and publishes the
//     winner; after one __syncthreads every wave reduces the 16 winners and
//     so knows the seed. Each lane then tests its block: if the seed's squared
//     distance to the box is at least the block's maximum (with a relative
//     margin far above rounding), no point of the block can get closer, so
//     the block is skipped. The wave sweeps the blocks that remain, updates
//     their points' distances and recomputes their maxima.
//   * The arg-max reductions run on FP32 keys (each maximum rounded down,
//     which keeps the order) and compare exactly only the lanes that tie on
//     the key, so a reduction step is one DPP move and one FP32 max rather
//     than an FP64 compare-and-select of value and index.
//
// Late in the run a seed reaches a few blocks, so a step is one sweep of a few
// hundred points plus two reductions. Distances are the CPU route's own
// expression (no contraction), ties go to the lower index, so the seeds are
// the CPU route's seeds bit for bit.

constexpr int kSampleThreads = 1024;
constexpr int kSampleWaves = kSampleThreads / 64;
constexpr int kSampleBlocks = kSampleThreads;  // one block per lane
constexpr double kSampleMargin = 1.0 - 1e-9;

// The larger value wins; equal values go to the lower index (the first
// maximum of an in-order sweep). A total order, so any reduction order gives
// the same winner.
__device__ inline bool beats(double v, int32_t i, double bv, int32_t bi) {
    return v > bv || (v == bv && i < bi);
}

// One arg-max step against the lane a DPP move pairs with (pairs are
// symmetric, and the order is total, so both lanes keep the same winner).
template <int Ctrl>
__device__ inline void argmax_step(double& v, int32_t& i) {
    const double ov = dpp_move<Ctrl>(v);
    const int32_t oi = dpp_move<Ctrl>(i);
    if (beats(ov, oi, v, i)) {
        v = ov;
        i = oi;
    }
}

// Arg-max over each row of sixteen lanes; every lane of the row gets it.
__device__ inline void row_argmax(double& v, int32_t& i) {
    argmax_step<kDppQuadSwap1>(v, i);
    argmax_step<kDppQuadSwap2>(v, i);
    argmax_step<kDppRowHalfMirror>(v, i);
    argmax_step<kDppRowMirror>(v, i);
}

// Arg-max over the wave: rows by DPP moves, then the four row winners read
// as scalars. Every lane ends with the winner.
__device__ inline void wave_argmax(double& v, int32_t& i) {
    row_argmax(v, i);
    double bv = read_lane(v, 0);
    int32_t bi = __builtin_amdgcn_readlane(i, 0);
#pragma unroll
    for (int r = 1; r < 4; ++r) {
        const double ov = read_lane(v, 16 * r);
        const int32_t oi = __builtin_amdgcn_readlane(i, 16 * r);
        if (beats(ov, oi, bv, bi)) {
            bv = ov;
            bi = oi;
        }
    }
    v = bv;
    i = bi;
}

// The lane holding index `i` (indices are unique within a reduction); the
// same for every lane.
__device__ inline int owner_lane(int32_t mine, int32_t i) {
    return __ffsll(static_cast<unsigned long long>(__ballot(mine == i))) - 1;
}

// p of lane `lane` (the same lane for every lane of the wave).
__device__ inline d3 lane_d3(d3 p, int lane) {
    return d3{read_lane(p.x, lane), read_lane(p.y, lane), read_lane(p.z, lane)};
}

// Maximum of `key` over the first `rows` rows of sixteen lanes, by FP32 max
// over DPP moves: every row reduces to its maximum, then the rows are read as
// scalars.
__device__ inline float keys_max(float key, int rows) {
    key = fmaxf(key, dpp_move<kDppQuadSwap1>(key));
    key = fmaxf(key, dpp_move<kDppQuadSwap2>(key));
    key = fmaxf(key, dpp_move<kDppRowHalfMirror>(key));
    key = fmaxf(key, dpp_move<kDppRowMirror>(key));
    float m = read_lane(key, 0);
    for (int r = 1; r < rows; ++r) m = fmaxf(m, read_lane(key, 16 * r));
    return m;
}

// The lane holding the (value, index) maximum of the lanes in `lanes` (the
// whole wave, or its first row), each lane's value given with its key, the
// value rounded down to FP32. Rounding is monotone, so the maximum's key is
// the largest key: the keys are reduced in FP32 (one DPP move and one max per
// step), and only the lanes whose key equals the largest -- usually one -- are
// compared exactly. The same lane for every lane of the wave.
__device__ inline int argmax_lane(double v, int32_t i, float key, uint64_t lanes) {
    const float top = keys_max(key, lanes == ~0ull ? 4 : 1);
    uint64_t cand = __ballot(key == top) & lanes;
    int lane = __ffsll(static_cast<unsigned long long>(cand)) - 1;
    cand &= cand - 1;
    if (cand) {
        double bv = read_lane(v, lane);
        int32_t bi = __builtin_amdgcn_readlane(i, lane);
        while (cand) {
            const int l = __ffsll(static_cast<unsigned long long>(cand)) - 1;
            cand &= cand - 1;
            const double ov = read_lane(v, l);
            const int32_t oi = __builtin_amdgcn_readlane(i, l);
            if (beats(ov, oi, bv, bi)) {
                bv = ov;
                bi = oi;
                lane = l;
            }
        }
    }
    return lane;
}

// Builds the blocked layout: records (x, y, z, running distance to the first
// seed) in Morton order, the original indices, and each block's box and
// maximum. One wave per block; padding slots hold distance -1 (never chosen,
// never changed) and index INT32_MAX.
__global__ void __launch_bounds__(64)
sample_init_kernel(const double* __restrict__ p, int64_t n, const int32_t* __restrict__ order, int32_t first,
                   int ppl, double4* __restrict__ rec, int32_t* __restrict__ ids, float* __restrict__ box,
                   double4* __restrict__ best, int32_t* __restrict__ best_index, int32_t* __restrict__ out) {
    const int lane = threadIdx.x;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * 64 * ppl;
    const d3 c0 = ld3(p, first);
    d3 lo{DBL_MAX, DBL_MAX, DBL_MAX}, hi{-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double v = -1.0;
    int32_t vi = 0x7fffffff;
    d3 vp{0.0, 0.0, 0.0};
    for (int k = 0; k < ppl; ++k) {
        const int64_t j = base + static_cast<int64_t>(k) * 64 + lane;
        if (j < n) {
            const int32_t i = order[j];
            const d3 q = ld3(p, i);
            const double d = norm2(q - c0);
            rec[j] = make_double4(q.x, q.y, q.z, d);
            ids[j] = i;
            lo = d3{fmin(lo.x, q.x), fmin(lo.y, q.y), fmin(lo.z, q.z)};
            hi = d3{fmax(hi.x, q.x), fmax(hi.y, q.y), fmax(hi.z, q.z)};
            if (beats(d, i, v, vi)) {
                v = d;
                vi = i;
                vp = q;
            }
        } else {
            rec[j] = make_double4(0.0, 0.0, 0.0, -1.0);
            ids[j] = 0x7fffffff;
        }
    }
    const int32_t own = vi;
    wave_argmax(v, vi);
    vp = lane_d3(vp, owner_lane(own, vi));
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) {
        lo = d3{fmin(lo.x, __shfl_xor(lo.x, offset, 64)), fmin(lo.y, __shfl_xor(lo.y, offset, 64)),
                fmin(lo.z, __shfl_xor(lo.z, offset, 64))};
        hi = d3{fmax(hi.x, __shfl_xor(hi.x, offset, 64)), fmax(hi.y, __shfl_xor(hi.y, offset, 64)),
                fmax(hi.z, __shfl_xor(hi.z, offset, 64))};
    }
    if (lane == 0) {
        float* b = box + 6 * static_cast<int64_t>(blockIdx.x);
        b[0] = __double2float_rd(lo.x);
        b[1] = __double2float_rd(lo.y);
        b[2] = __double2float_rd(lo.z);
        b[3] = __double2float_ru(hi.x);
        b[4] = __double2float_ru(hi.y);
        b[5] = __double2float_ru(hi.z);
        best[blockIdx.x] = make_double4(v, vp.x, vp.y, vp.z);
        best_index[blockIdx.x] = vi;
        if (blockIdx.x == 0) out[0] = first;
    }
}

// Sweeps NB blocks (bases base[0..NB)) for seed c together, so their loads
// are in flight at once: U records per lane per block per round, `rounds`
// rounds. Returns each block's new maximum (every lane holds them).
template <int U, int NB>
__device__ inline void sample_sweep(double4* __restrict__ rec, const int32_t* __restrict__ ids,
                                    const int64_t (&base)[NB], int rounds, d3 c, int lane, double (&v)[NB],
                                    int32_t (&vi)[NB], d3 (&vp)[NB]) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        v[b] = -1.0;
        vi[b] = 0x7fffffff;
        vp[b] = d3{0.0, 0.0, 0.0};
    }
    for (int r = 0; r < rounds; ++r) {
        double4 q[NB][U];
        int32_t id[NB][U];
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t j = base[b] + static_cast<int64_t>(r * U + u) * 64 + lane;
                q[b][u] = rec[j];
                id[b][u] = ids[j];
            }
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (q[b][u].w >= 0.0) {
                    const double d = norm2(d3{q[b][u].x, q[b][u].y, q[b][u].z} - c);
                    if (d < q[b][u].w) {
                        q[b][u].w = d;
                        reinterpret_cast<double*>(rec + base[b] + static_cast<int64_t>(r * U + u) * 64 + lane)[3] = d;
                    }
                    if (beats(q[b][u].w, id[b][u], v[b], vi[b])) {
                        v[b] = q[b][u].w;
                        vi[b] = id[b][u];
                        vp[b] = d3{q[b][u].x, q[b][u].y, q[b][u].z};
                    }
                }
            }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int at = argmax_lane(v[b], vi[b], __double2float_rd(v[b]), ~0ull);
        v[b] = read_lane(v[b], at);
        vi[b] = __builtin_amdgcn_readlane(vi[b], at);
        vp[b] = lane_d3(vp[b], at);
    }
}

// Seeds [begin, end) in one workgroup; the running state (records and block
// summaries) stays in memory between launches, so ranges resume exactly.
template <int U>
__global__ void __launch_bounds__(kSampleThreads)
sample_kernel(double4* __restrict__ rec, const int32_t* __restrict__ ids, const float* __restrict__ box,
              double4* __restrict__ best, int32_t* __restrict__ best_index, int blocks, int rounds, int32_t begin,
              int32_t end, int32_t* __restrict__ out) {
    // A long dependent chain beside the solve's kernels: take issue priority
    // on the unit it shares.
    __builtin_amdgcn_s_setprio(3);
    __shared__ double slot_v[2][kSampleWaves];
    __shared__ int32_t slot_i[2][kSampleWaves];
    __shared__ float slot_k[2][kSampleWaves];
    __shared__ double slot_p[2][kSampleWaves][3];
    const int wave = threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    const int block = lane * kSampleWaves + wave;
    const bool owned = block < blocks;
    const int64_t points_per_block = static_cast<int64_t>(64) * U * rounds;
    d3 lo{0.0, 0.0, 0.0}, hi{0.0, 0.0, 0.0}, bp{0.0, 0.0, 0.0};
    double bv = -2.0;  // below the padding's -1: an unowned lane never wins
    int32_t bi = 0x7fffffff;
    if (owned) {
        const float* b = box + 6 * static_cast<int64_t>(block);
        lo = d3{b[0], b[1], b[2]};
        hi = d3{b[3], b[4], b[5]};
        const double4 q = best[block];
        bv = q.x;
        bp = d3{q.y, q.z, q.w};
        bi = best_index[block];
    }
    float bk = __double2float_rd(bv);  // key of the block maximum
    for (int32_t s = begin; s < end; ++s) {
        const int buf = s & 1;
        const int at = argmax_lane(bv, bi, bk, ~0ull);
        const double v = read_lane(bv, at);
        const int32_t vi = __builtin_amdgcn_readlane(bi, at);
        const d3 vp = lane_d3(bp, at);
        if (lane == 0) {
            slot_v[buf][wave] = v;
            slot_i[buf][wave] = vi;
            slot_k[buf][wave] = __double2float_rd(v);
            slot_p[buf][wave][0] = vp.x;
            slot_p[buf][wave][1] = vp.y;
            slot_p[buf][wave][2] = vp.z;
        }
        __syncthreads();
        // The first row of lanes holds the sixteen wave winners.
        const int w16 = lane % kSampleWaves;
        const int from = argmax_lane(slot_v[buf][w16], slot_i[buf][w16], slot_k[buf][w16], 0xffffull);
        const int32_t si = slot_i[buf][from];
        const d3 c{slot_p[buf][from][0], slot_p[buf][from][1], slot_p[buf][from][2]};
        if (threadIdx.x == 0) out[s] = si;
        // Blocks the seed may reach: squared distance from c to the box below
        // the block's maximum.
        const double dx = fmax(fmax(lo.x - c.x, c.x - hi.x), 0.0);
        const double dy = fmax(fmax(lo.y - c.y, c.y - hi.y), 0.0);
        const double dz = fmax(fmax(lo.z - c.z, c.z - hi.z), 0.0);
        const double reach = dx * dx + dy * dy + dz * dz;
        unsigned long long todo = __ballot(owned && reach * kSampleMargin < bv);
        // Two blocks at a time while at least two remain (where the registers
        // allow: up to two records per lane per block).
        while (todo) {
            const int l0 = __ffsll(todo) - 1;
            todo &= todo - 1;
            if (U <= 2 && todo) {
                const int l1 = __ffsll(todo) - 1;
                todo &= todo - 1;
                const int64_t base[2] = {static_cast<int64_t>(l0 * kSampleWaves + wave) * points_per_block,
                                         static_cast<int64_t>(l1 * kSampleWaves + wave) * points_per_block};
                double nv[2];
                int32_t ni[2];
                d3 np[2];
                sample_sweep<U, 2>(rec, ids, base, rounds, c, lane, nv, ni, np);
                if (lane == l0) {
                    bv = nv[0];
                    bi = ni[0];
                    bp = np[0];
                    bk = __double2float_rd(bv);
                }
                if (lane == l1) {
                    bv = nv[1];
                    bi = ni[1];
                    bp = np[1];
                    bk = __double2float_rd(bv);
                }
            } else {
                const int64_t base[1] = {static_cast<int64_t>(l0 * kSampleWaves + wave) * points_per_block};
                double nv[1];
                int32_t ni[1];
                d3 np[1];
                sample_sweep<U, 1>(rec, ids, base, rounds, c, lane, nv, ni, np);
                if (lane == l0) {
                    bv = nv[0];
                    bi = ni[0];
                    bp = np[0];
                    bk = __double2float_rd(bv);
                }
            }
        }
    }
    if (owned) {
        best[block] = make_double4(bv, bp.x, bp.y, bp.z);
        best_index[block] = bi;
    }
}

// ------------------------------------------------------------ node grid ----
//
// Exact nearest-node searches over a uniform grid of cells. Nodes are sorted
// by cell; a search visits the cells in cubic shells around the query's cell
// and stops once its k-th best is closer than the nearest face of the visited
// cube that still has cells beyond it (less a margin far above the rounding
// of the cell assignment), or every cell has been visited. The visit order
// does not matter: candidates are kept in the total (distance, index) order,
// so the answer is the brute-force answer.

struct GridView {
    const double* frame;  // origin xyz, cell size
    int cells;            // per axis
    const int32_t* runs;  // per cell [start, end) in the sorted order
    const double4* sorted;  // (x, y, z, node index) in cell order
};

__device__ inline int grid_axis_cell(double x, double origin, double h, int cells) {
    const double t = floor((x - origin) / h);
    return t < 0.0 ? 0 : (t > static_cast<double>(cells - 1) ? cells - 1 : static_cast<int>(t));
}

// Visits every node of the shells around x's cell, shell by shell, calling
// visit(index, position) for each, until done(bound) holds after a shell
// (bound: every node not yet visited is at least that far from x) or no cell
// is left.
template <class Visit, class Done>
__device__ inline void grid_search(const GridView& g, d3 x, Visit visit, Done done) {
    const double o[3] = {g.frame[0], g.frame[1], g.frame[2]};
    const double h = g.frame[3];
    const double xv[3] = {x.x, x.y, x.z};
    int c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = grid_axis_cell(xv[k], o[k], h, g.cells);
    const double margin = 1e-9 * h;
    for (int r = 0;; ++r) {
        const int z0 = max(c[2] - r, 0), z1 = min(c[2] + r, g.cells - 1);
        const int y0 = max(c[1] - r, 0), y1 = min(c[1] + r, g.cells - 1);
        const int x0 = max(c[0] - r, 0), x1 = min(c[0] + r, g.cells - 1);
        for (int cz = z0; cz <= z1; ++cz) {
            const bool zface = cz == c[2] - r || cz == c[2] + r;
            for (int cy = y0; cy <= y1; ++cy) {
                const bool yface = zface || cy == c[1] - r || cy == c[1] + r;
                for (int cx = x0; cx <= x1; ++cx) {
                    if (!yface && cx != c[0] - r && cx != c[0] + r) continue;  // inside: an earlier shell
                    const int64_t cell = (static_cast<int64_t>(cz) * g.cells + cy) * g.cells + cx;
                    const int32_t e = g.runs[2 * cell + 1];
                    for (int32_t t = g.runs[2 * cell]; t < e; ++t) {
                        const double4 q = g.sorted[t];
                        visit(static_cast<int32_t>(q.w), d3{q.x, q.y, q.z});
                    }
                }
            }
        }
        // Distance from x to the nearest face of the visited cube that has
        // cells beyond it.
        double bound = __builtin_inf();
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (c[k] - r > 0) bound = fmin(bound, xv[k] - (o[k] + static_cast<double>(c[k] - r) * h));
            if (c[k] + r < g.cells - 1) bound = fmin(bound, (o[k] + static_cast<double>(c[k] + r + 1) * h) - xv[k]);
        }
        if (bound == __builtin_inf()) return;
        if (done(bound - margin)) return;
    }
}

// Box of the nodes and the grid frame: origin = box low corner, cell size =
// largest extent / cells (a little larger, so the far corner is in the last
// cell; 1 for a degenerate box). One block.
__global__ void __launch_bounds__(256) grid_frame_kernel(const double* __restrict__ p, int64_t n, int cells,
                                                         double* __restrict__ frame) {
    __shared__ double lo[3][256], hi[3][256];
    double l[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, u[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int64_t i = threadIdx.x; i < n; i += 256)
        for (int k = 0; k < 3; ++k) {
            l[k] = fmin(l[k], p[3 * i + k]);
            u[k] = fmax(u[k], p[3 * i + k]);
        }
    for (int k = 0; k < 3; ++k) {
        lo[k][threadIdx.x] = l[k];
        hi[k][threadIdx.x] = u[k];
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
    if (threadIdx.x == 0) {
        double extent = 0.0;
        for (int k = 0; k < 3; ++k) extent = fmax(extent, hi[k][0] - lo[k][0]);
        for (int k = 0; k < 3; ++k) frame[k] = lo[k][0];
        frame[3] = extent > 0.0 ? extent / cells * (1.0 + 1e-6) : 1.0;
    }
}

__global__ void grid_keys_kernel(const double* __restrict__ p, int64_t n, const double* __restrict__ frame,
                                 int cells, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double h = frame[3];
    const int cx = grid_axis_cell(p[3 * i], frame[0], h, cells);
    const int cy = grid_axis_cell(p[3 * i + 1], frame[1], h, cells);
    const int cz = grid_axis_cell(p[3 * i + 2], frame[2], h, cells);
    keys[i] = (static_cast<uint64_t>(cz) * cells + cy) * cells + cx;
    vals[i] = static_cast<uint32_t>(i);
}

__global__ void grid_fill_kernel(const double* __restrict__ p, int64_t n, const uint64_t* __restrict__ keys,
                                 const uint32_t* __restrict__ vals, int32_t* __restrict__ runs,
                                 double4* __restrict__ sorted) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t c = keys[i];
    if (i == 0 || keys[i - 1] != c) runs[2 * c] = static_cast<int32_t>(i);
    if (i == n - 1 || keys[i + 1] != c) runs[2 * c + 1] = static_cast<int32_t>(i + 1);
    const uint32_t j = vals[i];
    sorted[i] = make_double4(p[3 * j], p[3 * j + 1], p[3 * j + 2], static_cast<double>(j));
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

__device__ inline double pick4(const double* v, int k) {
    return k == 0 ? v[0] : (k == 1 ? v[1] : (k == 2 ? v[2] : v[3]));
}

// Nearest 4 nodes by (d^2, index), the order the CPU route's keep_nearest
// produces (candidates visited in increasing index, ties kept by the lower
// index); callers use the first `need` (at most the node count) of them.
__device__ inline void nearest4(d3 x, const GridView& g, int need, double* d2, int32_t* id) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        d2[m] = __builtin_inf();
        id[m] = 0x7fffffff;
    }
    grid_search(
        g, x,
        [&](int32_t j, d3 p) {
            const double s = norm2(x - p);
            if (s < d2[3] || (s == d2[3] && j < id[3])) insert_sorted<4>(s, j, d2, id);
        },
        [&](double bound) { return bound > 0.0 && pick4(d2, need - 1) < bound * bound; });
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

// ------------------------------------------------------ brute force ----
//
// The k nearest of a few thousand points, for a few thousand queries: one
// wave per query, kBruteWaves queries per block sharing each tile of
// candidate points in LDS. Each lane keeps its own sorted KMAX best among the
// candidates it visits; the wave then merges the lanes' lists in KMAX rounds
// of a butterfly arg-min. The order is the grid search's (the key, then the
// index), so the answer is the same bits: Root orders by sqrt(d^2) (the
// Euclidean distance, taken only when a candidate's d^2 is within rounding of
// the lane's current KMAX-th), otherwise by d^2.
constexpr int kBruteWaves = 16;
constexpr int kBruteThreads = kBruteWaves * kWave;
constexpr int kBruteTile = kBruteThreads;  // candidate points per LDS tile
// Query-candidate pairs up to which the brute-force searches replace the
// grid's. Measured on the 200k template's levels (kNN of 2,000 / 4,000 /
// 8,000 nodes: 45 / 88 / 438 us by brute force against 165 / 154 / 183 us on
// the grid; the level transfer onto 4,000 / 8,000 nodes: 60 / 95 against
// 91 / 80 us): the grid's dependent cell visits leave a small level's
// launch latency-bound, and every pair costs more beyond a few million.
constexpr int64_t kBruteMaxPairs = int64_t{1} << 24;

template <int KMAX, bool Root>
__device__ inline void brute_nearest(d3 x, int64_t self, const double* __restrict__ pts, int64_t np, bool active,
                                     double* tile /* 3 x kBruteTile */, double (&best)[KMAX],
                                     int32_t (&best_id)[KMAX]) {
    const int lane = threadIdx.x % kWave;
    double dist[KMAX];
    int32_t id[KMAX];
#pragma unroll
    for (int m = 0; m < KMAX; ++m) {
        dist[m] = __builtin_inf();
        id[m] = 0x7fffffff;
    }
    for (int64_t base = 0; base < np; base += kBruteTile) {
        const int count = static_cast<int>(np - base < kBruteTile ? np - base : kBruteTile);
        __syncthreads();
        for (int t = threadIdx.x; t < 3 * count; t += kBruteThreads) tile[t] = pts[3 * base + t];
        __syncthreads();
        if (!active) continue;
        for (int t = lane; t < count; t += kWave) {
            const int64_t j = base + t;
            if (j == self) continue;
            const double d2 = norm2(x - d3{tile[3 * t], tile[3 * t + 1], tile[3 * t + 2]});
            const double last = dist[KMAX - 1];
            // sqrt is monotone and correctly rounded: a d^2 this far above
            // the KMAX-th one's square cannot tie or beat it.
            if (Root ? !(d2 <= last * last * (1.0 + 1e-15)) : !(d2 <= last)) continue;
            const double key = Root ? sqrt(d2) : d2;
            const int32_t jj = static_cast<int32_t>(j);
            if (key < last || (key == last && jj < id[KMAX - 1])) insert_sorted<KMAX>(key, jj, dist, id);
        }
    }
    if (!active) return;
#pragma unroll
    for (int m = 0; m < KMAX; ++m) {
        double bd = dist[0];
        int32_t bi = id[0];
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
            const double od = __shfl_xor(bd, off);
            const int32_t oi = __shfl_xor(bi, off);
            if (od < bd || (od == bd && oi < bi)) {
                bd = od;
                bi = oi;
            }
        }
        best[m] = bd;
        best_id[m] = bi;
        if (dist[0] == bd && id[0] == bi) {
#pragma unroll
            for (int q = 0; q + 1 < KMAX; ++q) {
                dist[q] = dist[q + 1];
                id[q] = id[q + 1];
            }
            dist[KMAX - 1] = __builtin_inf();
            id[KMAX - 1] = 0x7fffffff;
        }
    }
}

constexpr int kBindThreads = 128;

__global__ void __launch_bounds__(kBindThreads)
bind_kernel(const double* __restrict__ p, int64_t n, const int32_t* __restrict__ order, GridView grid, int support,
            int32_t* __restrict__ on, double* __restrict__ ow, int32_t* __restrict__ oc) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t v = order ? order[t] : t;
    const d3 x = ld3(p, v);
    double d2[4];
    int32_t id[4];
    nearest4(x, grid, support, d2, id);
    double w[4];
    distance_weights(d2, support, w);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        on[4 * v + k] = k < support ? id[k] : -1;
        ow[4 * v + k] = k < support ? w[k] : 0.0;
    }
    oc[v] = support;
}

// Node i's first k neighbours as sortable (min, max) pair keys; entries past
// the nn - 1 other nodes are all ones.
template <int KMAX>
__device__ inline void emit_knn_keys(int64_t i, int64_t nn, int k, const int32_t (&id)[KMAX],
                                     uint64_t* __restrict__ keys) {
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

// The grid search of one query spread over a wave: each shell's cells are
// dealt to the lanes, each lane keeps its own sorted KMAX best, and after
// each shell the wave merges the lanes' lists (without consuming them) to test
// the stop condition on the need-th best, which is grid_search's test on the
// same candidates. Root orders by sqrt(d^2) and stops when the need-th best
// is closer than the bound; otherwise by d^2, stopping when it is below the
// bound's square. Every lane of the wave returns the merged list.
constexpr int kWaveQueryThreads = 256;
constexpr int kWaveQueries = kWaveQueryThreads / kWave;

template <int KMAX, bool Root>
__device__ inline void wave_grid_nearest(const GridView& g, d3 x, int64_t self, int need, double (&best)[KMAX],
                                         int32_t (&best_id)[KMAX]) {
    const int lane = threadIdx.x % kWave;
    const double o[3] = {g.frame[0], g.frame[1], g.frame[2]};
    const double h = g.frame[3];
    const double xv[3] = {x.x, x.y, x.z};
    int c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = grid_axis_cell(xv[k], o[k], h, g.cells);
    const double margin = 1e-9 * h;
    double dist[KMAX];
    int32_t id[KMAX];
#pragma unroll
    for (int m = 0; m < KMAX; ++m) {
        dist[m] = __builtin_inf();
        id[m] = 0x7fffffff;
    }
    for (int r = 0;; ++r) {
        const int side = 2 * r + 1;
        const int cube = side * side * side;
        for (int q = lane; q < cube; q += kWave) {
            const int dz = q / (side * side) - r, dy = (q / side) % side - r, dx = q % side - r;
            if (abs(dx) != r && abs(dy) != r && abs(dz) != r) continue;  // inside: an earlier shell
            const int cx = c[0] + dx, cy = c[1] + dy, cz = c[2] + dz;
            if (cx < 0 || cy < 0 || cz < 0 || cx >= g.cells || cy >= g.cells || cz >= g.cells) continue;
            const int64_t cell = (static_cast<int64_t>(cz) * g.cells + cy) * g.cells + cx;
            const int32_t e = g.runs[2 * cell + 1];
            for (int32_t t = g.runs[2 * cell]; t < e; ++t) {
                const double4 p = g.sorted[t];
                const int32_t j = static_cast<int32_t>(p.w);
                if (j == self) continue;
                const double d2 = norm2(x - d3{p.x, p.y, p.z});
                const double key = Root ? sqrt(d2) : d2;
                if (key < dist[KMAX - 1] || (key == dist[KMAX - 1] && j < id[KMAX - 1]))
                    insert_sorted<KMAX>(key, j, dist, id);
            }
        }
        // The merged best so far, the lanes' lists left intact.
        int head = 0;
#pragma unroll
        for (int m = 0; m < KMAX; ++m) {
            double hd = __builtin_inf();
            int32_t hi = 0x7fffffff;
#pragma unroll
            for (int q = 0; q < KMAX; ++q)
                if (q == head) {
                    hd = dist[q];
                    hi = id[q];
                }
            double bd = hd;
            int32_t bi = hi;
#pragma unroll
            for (int off = kWave / 2; off > 0; off >>= 1) {
                const double od = __shfl_xor(bd, off);
                const int32_t oi = __shfl_xor(bi, off);
                if (od < bd || (od == bd && oi < bi)) {
                    bd = od;
                    bi = oi;
                }
            }
            best[m] = bd;
            best_id[m] = bi;
            if (hd == bd && hi == bi && head < KMAX) ++head;
        }
        double bound = __builtin_inf();
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (c[k] - r > 0) bound = fmin(bound, xv[k] - (o[k] + static_cast<double>(c[k] - r) * h));
            if (c[k] + r < g.cells - 1) bound = fmin(bound, (o[k] + static_cast<double>(c[k] + r + 1) * h) - xv[k]);
        }
        if (bound == __builtin_inf()) return;
        const double b = bound - margin;
        double kth = best[0];
#pragma unroll
        for (int m = 1; m < KMAX; ++m) kth = m == need - 1 ? best[m] : kth;
        if (Root ? kth < b : (b > 0.0 && kth < b * b)) return;
    }
}

// kNN edges of a large level: one wave per node.
template <int KMAX>
__global__ void __launch_bounds__(kWaveQueryThreads)
knn_wave(const double* __restrict__ nodes, int64_t nn, GridView grid, int k, uint64_t* __restrict__ keys) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kWaveQueries + threadIdx.x / kWave;
    if (i >= nn) return;  // the whole wave
    const int held = static_cast<int>(nn - 1 < k ? nn - 1 : k);
    double dist[KMAX];
    int32_t id[KMAX];
    wave_grid_nearest<KMAX, true>(grid, ld3(nodes, i), i, held, dist, id);
    if (threadIdx.x % kWave == 0) emit_knn_keys<KMAX>(i, nn, k, id, keys);
}

// The same k nearest, by brute force (few nodes): one wave per node.
template <int KMAX>
__global__ void __launch_bounds__(kBruteThreads)
knn_brute(const double* __restrict__ nodes, int64_t nn, int k, uint64_t* __restrict__ keys) {
    __shared__ double tile[3 * kBruteTile];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBruteWaves + threadIdx.x / kWave;
    const bool active = i < nn;
    const d3 x = active ? ld3(nodes, i) : d3{0.0, 0.0, 0.0};
    double dist[KMAX];
    int32_t id[KMAX];
    brute_nearest<KMAX, true>(x, i, nodes, nn, active, tile, dist, id);
    if (active && threadIdx.x % kWave == 0) emit_knn_keys<KMAX>(i, nn, k, id, keys);
}

__global__ void unique_edges(const uint64_t* __restrict__ keys, const int32_t* __restrict__ flags,
                             const int32_t* __restrict__ pos, int64_t n, int64_t nn, int32_t* __restrict__ edges) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i] || keys[i] == ~0ull) return;
    const int64_t e = pos[i];
    edges[2 * e] = static_cast<int32_t>(keys[i] / static_cast<uint64_t>(nn));
    edges[2 * e + 1] = static_cast<int32_t>(keys[i] % static_cast<uint64_t>(nn));
}

__device__ inline void transfer_node(const double* __restrict__ from, const double* __restrict__ R,
                                     const double* __restrict__ T, int64_t i, d3 x, int support, const double* d2,
                                     const int32_t* id, double* __restrict__ Ro, double* __restrict__ To);

// Blend of the bound source rotations in quaternion space, signs aligned to
// the heaviest contributor, heaviest-neighbour fallback when the blend nearly
// cancels; translation maps the new rest position onto the field's value.
__global__ void __launch_bounds__(kWaveQueryThreads)
transfer_wave(const double* __restrict__ from, GridView grid, const double* __restrict__ R,
              const double* __restrict__ T, const double* __restrict__ to, int64_t nt, int support,
              double* __restrict__ Ro, double* __restrict__ To) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kWaveQueries + threadIdx.x / kWave;
    if (i >= nt) return;  // the whole wave
    const d3 x = ld3(to, i);
    double d2[4];
    int32_t id[4];
    wave_grid_nearest<4, false>(grid, x, -1, support, d2, id);
    if (threadIdx.x % kWave == 0) transfer_node(from, R, T, i, x, support, d2, id, Ro, To);
}

// The same transfer with the 4 nearest coarse nodes by brute force (few
// coarse nodes): one wave per fine node.
__global__ void __launch_bounds__(kBruteThreads)
transfer_brute(const double* __restrict__ from, int64_t nf, const double* __restrict__ R,
               const double* __restrict__ T, const double* __restrict__ to, int64_t nt, int support,
               double* __restrict__ Ro, double* __restrict__ To) {
    __shared__ double tile[3 * kBruteTile];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBruteWaves + threadIdx.x / kWave;
    const bool active = i < nt;
    const d3 x = active ? ld3(to, i) : d3{0.0, 0.0, 0.0};
    double d2[4];
    int32_t id[4];
    brute_nearest<4, false>(x, -1, from, nf, active, tile, d2, id);
    if (active && threadIdx.x % kWave == 0) transfer_node(from, R, T, i, x, support, d2, id, Ro, To);
}

// A fine node's rotation and translation from its (up to) 4 nearest coarse
// nodes: distance weights, rotations blended as quaternions in the
// hemisphere of the heaviest one, positions moved by each and averaged.
__device__ inline void transfer_node(const double* __restrict__ from, const double* __restrict__ R,
                                     const double* __restrict__ T, int64_t i, d3 x, int support, const double* d2,
                                     const int32_t* id, double* __restrict__ Ro, double* __restrict__ To) {
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

void packed_evaluate(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, KernelSink& sink) {
    if (maps.vertex_blocks <= 0) return;
    sink_launch(sink, evaluate_packed, static_cast<unsigned>(maps.vertex_blocks), 256, jobs, ctl, maps.vertex);
}

void packed_retract(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, KernelSink& sink) {
    if (maps.node_blocks <= 0) return;
    sink_launch(sink, retract_packed, static_cast<unsigned>(maps.node_blocks), 128, jobs, ctl, maps.node);
}

void packed_commit(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, KernelSink& sink) {
    if (maps.node_blocks <= 0) return;
    sink_launch(sink, commit_packed, static_cast<unsigned>(maps.node_blocks), 128, jobs, ctl, maps.node);
}

namespace {
// Points per lane per block, so that at most kSampleBlocks blocks cover the
// points; beyond 4 per lane, whole rounds of 4.
int sample_ppl(int64_t n) {
    int ppl = static_cast<int>((n + 64 * kSampleBlocks - 1) / (64 * kSampleBlocks));
    if (ppl > 4) ppl = (ppl + 3) / 4 * 4;
    return ppl;
}
}  // namespace

void farthest_point_prepare(const double* points, int64_t n, int32_t first, int32_t* out, FpsWork& w,
                            hipStream_t stream) {
    if (n <= 0) return;
    if (first < 0 || first >= n) throw InvalidArgument("farthest_point_seeds: the first seed is out of range");
    if (n >= (1ll << 31) - 1) throw InvalidArgument("farthest_point_seeds: too many points");
    const int ppl = sample_ppl(n);
    const int64_t per_block = static_cast<int64_t>(64) * ppl;
    const int blocks = static_cast<int>((n + per_block - 1) / per_block);
    w.order.resize(n);
    morton_order(points, n, w.order.get(), w.sort, w.keys, w.bounds, stream);
    w.records.resize(static_cast<int64_t>(blocks) * per_block);
    w.ids.resize(static_cast<int64_t>(blocks) * per_block);
    w.box.resize(6 * static_cast<int64_t>(blocks));
    w.best.resize(blocks);
    w.best_index.resize(blocks);
    w.ppl = ppl;
    sample_init_kernel<<<blocks, 64, 0, stream>>>(points, n, w.order.get(), first, ppl, w.records.get(), w.ids.get(),
                                                  w.box.get(), w.best.get(), w.best_index.get(), out);
    NRICP_LAUNCH_CHECK("sample_init_kernel");
}

void farthest_point_seeds_range(const double* points, int64_t n, int32_t first, int32_t begin, int32_t end,
                                int32_t* out, FpsWork& w, hipStream_t stream) {
    if (n <= 0 || end <= begin) return;
    if (end > n) throw InvalidArgument("farthest_point_seeds: more seeds than points");
    if (begin == 0) {
        farthest_point_prepare(points, n, first, out, w, stream);
        begin = 1;
        if (end <= begin) return;
    }
    const int ppl = sample_ppl(n);
    const int64_t per_block = static_cast<int64_t>(64) * ppl;
    const int blocks = static_cast<int>((n + per_block - 1) / per_block);
    if (w.ppl != ppl || static_cast<int64_t>(w.best.size()) != blocks)
        throw InvalidArgument("farthest_point_seeds: a range resumed with different points");
    const int unroll = ppl < 4 ? ppl : 4;
    const int rounds = ppl / unroll;
    switch (unroll) {
        case 1:
            sample_kernel<1><<<1, kSampleThreads, 0, stream>>>(w.records.get(), w.ids.get(), w.box.get(),
                                                               w.best.get(), w.best_index.get(), blocks, rounds,
                                                               begin, end, out);
            break;
        case 2:
            sample_kernel<2><<<1, kSampleThreads, 0, stream>>>(w.records.get(), w.ids.get(), w.box.get(),
                                                               w.best.get(), w.best_index.get(), blocks, rounds,
                                                               begin, end, out);
            break;
        case 3:
            sample_kernel<3><<<1, kSampleThreads, 0, stream>>>(w.records.get(), w.ids.get(), w.box.get(),
                                                               w.best.get(), w.best_index.get(), blocks, rounds,
                                                               begin, end, out);
            break;
        default:
            sample_kernel<4><<<1, kSampleThreads, 0, stream>>>(w.records.get(), w.ids.get(), w.box.get(),
                                                               w.best.get(), w.best_index.get(), blocks, rounds,
                                                               begin, end, out);
            break;
    }
    NRICP_LAUNCH_CHECK("sample_kernel");
}

void farthest_point_seeds(const double* points, int64_t n, int32_t seeds, int32_t first, int32_t* out, FpsWork& w,
                          hipStream_t stream) {
    if (n <= 0 || seeds <= 0) return;
    farthest_point_seeds_range(points, n, first, 0, seeds, out, w, stream);
}

namespace {
GridView view_of(const NodeGrid& g) { return GridView{g.frame.get(), g.cells, g.runs.get(), g.sorted.get()}; }
}  // namespace

void build_node_grid(const double* nodes, int64_t nn, NodeGrid& g, SortScratch& scratch, hipStream_t stream) {
    g.count = nn;
    if (nn <= 0) return;
    if (nn >= (1ll << 31)) throw InvalidArgument("too many nodes");
    // About one cell per node: a surface's nodes fill ~ cells^(2/3) of them,
    // a few per occupied cell.
    const int cells = std::clamp(static_cast<int>(std::ceil(std::cbrt(static_cast<double>(nn)))), 1, 128);
    const int64_t total = static_cast<int64_t>(cells) * cells * cells;
    g.cells = cells;
    g.frame.resize(4);
    grid_frame_kernel<<<1, 256, 0, stream>>>(nodes, nn, cells, g.frame.get());
    NRICP_LAUNCH_CHECK("grid_frame_kernel");
    g.keys.resize(nn);
    g.vals.resize(nn);
    grid_keys_kernel<<<blocks_for(nn, 256), 256, 0, stream>>>(nodes, nn, g.frame.get(), cells, g.keys.get(),
                                                               g.vals.get());
    NRICP_LAUNCH_CHECK("grid_keys_kernel");
    radix_sort_pairs(g.keys.get(), g.vals.get(), nn, std::max(1, bits_for(static_cast<uint64_t>(total - 1))), scratch,
                     stream);
    g.runs.resize(2 * total);
    g.runs.zero(stream);
    g.sorted.resize(nn);
    grid_fill_kernel<<<blocks_for(nn, 256), 256, 0, stream>>>(nodes, nn, g.keys.get(), g.vals.get(), g.runs.get(),
                                                               g.sorted.get());
    NRICP_LAUNCH_CHECK("grid_fill_kernel");
}

void bind_vertices(const double* points, int64_t n, const double* nodes, int64_t nn, const NodeGrid& grid, int32_t k,
                   int32_t* on, double* ow, int32_t* oc, hipStream_t stream, const int32_t* order) {
    if (n <= 0 || nn <= 0) return;
    if (grid.count != nn) throw InvalidArgument("bind_vertices: the grid is not over these nodes");
    (void)nodes;
    const int support = std::max(1, std::min<int>({k, static_cast<int>(std::min<int64_t>(nn, 4)), 4}));
    bind_kernel<<<blocks_for(n, kBindThreads), kBindThreads, 0, stream>>>(points, n, order, view_of(grid), support, on,
                                                                           ow, oc);
    NRICP_LAUNCH_CHECK("bind_kernel");
}

int64_t build_node_edges(const double* nodes, int64_t nn, const NodeGrid& grid, int32_t k, DeviceBuffer<int32_t>& edges,
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
    if (grid.count != nn) throw InvalidArgument("build_node_edges: the grid is not over these nodes");
    if (nn * nn <= kBruteMaxPairs) {
        // Few nodes: every pair, one wave per node.
        const dim3 blocks(static_cast<unsigned>(blocks_for(nn, kBruteWaves)));
        if (kk <= 4)
            knn_brute<4><<<blocks, kBruteThreads, 0, stream>>>(nodes, nn, kk, keys.get());
        else if (kk <= 8)
            knn_brute<8><<<blocks, kBruteThreads, 0, stream>>>(nodes, nn, kk, keys.get());
        else
            knn_brute<16><<<blocks, kBruteThreads, 0, stream>>>(nodes, nn, kk, keys.get());
        NRICP_LAUNCH_CHECK("knn_brute");
    } else {
        const dim3 blocks(static_cast<unsigned>(blocks_for(nn, kWaveQueries)));
        const GridView view = view_of(grid);
        if (kk <= 4)
            knn_wave<4><<<blocks, kWaveQueryThreads, 0, stream>>>(nodes, nn, view, kk, keys.get());
        else if (kk <= 8)
            knn_wave<8><<<blocks, kWaveQueryThreads, 0, stream>>>(nodes, nn, view, kk, keys.get());
        else
            knn_wave<16><<<blocks, kWaveQueryThreads, 0, stream>>>(nodes, nn, view, kk, keys.get());
        NRICP_LAUNCH_CHECK("knn_wave");
    }
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

void transfer_level(const double* from, int64_t nf, const NodeGrid& from_grid, const double* R, const double* T,
                    const double* to, int64_t nt, int32_t k, double* Ro, double* To, hipStream_t stream) {
    if (nt <= 0) return;
    if (nf <= 0) throw InvalidArgument("transfer_level: the coarser level has no nodes");
    if (from_grid.count != nf) throw InvalidArgument("transfer_level: the grid is not over the coarser nodes");
    const int support = std::max(1, std::min<int>({k, static_cast<int>(std::min<int64_t>(nf, 4)), 4}));
    if (nf * nt <= kBruteMaxPairs) {
        transfer_brute<<<static_cast<unsigned>(blocks_for(nt, kBruteWaves)), kBruteThreads, 0, stream>>>(
            from, nf, R, T, to, nt, support, Ro, To);
        NRICP_LAUNCH_CHECK("transfer_brute");
        return;
    }
    transfer_wave<<<static_cast<unsigned>(blocks_for(nt, kWaveQueries)), kWaveQueryThreads, 0, stream>>>(
        from, view_of(from_grid), R, T, to, nt, support, Ro, To);
    NRICP_LAUNCH_CHECK("transfer_wave");
}

}  // namespace nricp_hip
