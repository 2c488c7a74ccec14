// Block Gauss-Newton assembly of the 6-DOF-per-node normal equations and the
// five block energies, atomics-free.
//
// Reference behaviour: the residual rows and Jacobians of
// src/energy/energy.cpp:124-271 accumulated as in
// src/solver/fused_assembly.cpp:159-300 (J^T J per node, -J^T r, J_a^T J_b per
// interacting pair). The per-row sums are taken in closed form: for a row
// direction e_k with arm a = R (v - g) and scale c the rows of a 3-row offset
// block give
//     sum_k J_k J_k^T = c^2 [ |a|^2 I - a a^T   [a]x ]
//                           [ [a]x^T             I   ]
//     -sum_k J_k r_k  = -c [ a x r ; r ]
// which is the same quantity with a fraction of the arithmetic.
//
// Passes: (1) one thread per vertex resolves its rows (deformed point,
// residuals, per-slot arms, scales) and their energies; (2) one thread per
// edge/node adds smoothness and anchor energies; (3) one wave per node gathers
// its diagonal block and right-hand side from its vertex rows, edges and
// anchor; (4) eight lanes per interacting pair gather the off-diagonal block
// ((3) and (4) share one launch, system_blocks).
// Energies are reduced in a fixed order, so the energy of a state is the same
// bits whether it comes from an assembly or an energy-only pass.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "packed.hpp"
#include "so3.hpp"
#include "xcd_map.hpp"

namespace nricp_hip {
namespace {

constexpr int kThreads = 256;

__device__ inline d3 apply_node(const double* R, const double* T, const double* G, int64_t j, d3 p) {
    const d3 g = ld3(G, j);
    return (mul3(R + 9 * j, p - g) + g) + ld3(T, j);
}

// J_l^{-1}(w) = I - [w]/2 + c [w]^2 (src/core/rotation.h:69-103).
__device__ inline void left_jacobian_inverse(d3 w, double* J) {
    const double t2 = dot(w, w);
    double c;
    if (t2 < 1e-12) {
        c = 1.0 / 12.0;
    } else {
        const double t = sqrt(t2);
        c = t > 3.14159265358979323846 - 1e-6 ? 0.0 : 1.0 / t2 - (1.0 + cos(t)) / (2.0 * t * sin(t));
    }
    const double k[9] = {0.0, -w.z, w.y, w.z, 0.0, -w.x, -w.y, w.x, 0.0};
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 3; ++col) {
            double k2 = 0.0;
            for (int m = 0; m < 3; ++m) k2 += k[r * 3 + m] * k[m * 3 + col];
            J[r * 3 + col] = (r == col ? 1.0 : 0.0) - 0.5 * k[r * 3 + col] + c * k2;
        }
}

// ------------------------------------------------------------ pass 1 -------

// Block `block` of the vertex rows: its vertices' residual rows (and, with
// Records, the scales, residuals and arms the assembly reuses), and the
// block's energy partials. lds: 3 x waves doubles.
template <bool Records>
__device__ __forceinline__ void vertex_rows_block(const AssemblyInputs& in, double* __restrict__ scale,
                                         double* __restrict__ resid, double* __restrict__ arm,
                                         double* __restrict__ partials, int64_t block, double* lds) {
    const int64_t v = block * kThreads + threadIdx.x;
    double e[3] = {0.0, 0.0, 0.0};
    if (v < in.vertex_count) {
        const d3 x = ld3(in.points, v);
        d3 def{0.0, 0.0, 0.0};
        d3 arms[4];
        bool use[4];
        int usable = 0;
        const int count = in.bind_counts[v];
        for (int s = 0; s < 4; ++s) {
            const int32_t j = in.bind_nodes[4 * v + s];
            const double w = in.bind_weights[4 * v + s];
            use[s] = s < count && j >= 0 && j < in.node_count && w != 0.0;
            arms[s] = d3{0.0, 0.0, 0.0};
            if (!use[s]) continue;
            ++usable;
            const d3 g = ld3(in.rest, j);
            arms[s] = mul3(in.rotations + 9 * static_cast<int64_t>(j), x - g);
            def = def + ((arms[s] + g) + ld3(in.translations, j)) * w;
        }
        double sp = -1.0, sl = -1.0, sm = -1.0;
        double r[7] = {0, 0, 0, 0, 0, 0, 0};
        if (usable > 0) {
            if (in.active ? in.active[v] != 0 : in.rejection[v] == 0) {
                const double u = fmax(0.0, in.robust[v]);
                const d3 y = ld3(in.corr_points, v);
                if (in.w.p2p > 0.0) {
                    sp = sqrt(in.w.p2p * u);
                    const d3 off = (def - y) * sp;
                    r[0] = off.x; r[1] = off.y; r[2] = off.z;
                    e[0] = (off.x * off.x + off.y * off.y) + off.z * off.z;
                }
                if (in.w.plane > 0.0) {
                    sl = sqrt(in.w.plane * u);
                    r[3] = sl * dot(ld3(in.corr_normals, v), def - y);
                    e[1] = r[3] * r[3];
                }
            }
            if (in.landmark_flag[v] && in.w.landmark > 0.0) {
                sm = sqrt(in.w.landmark);
                const d3 off = (def - ld3(in.landmark_target, v)) * sm;
                r[4] = off.x; r[5] = off.y; r[6] = off.z;
                e[2] = (off.x * off.x + off.y * off.y) + off.z * off.z;
            }
        }
        if (Records) {
            scale[3 * v] = sp;
            scale[3 * v + 1] = sl;
            scale[3 * v + 2] = sm;
            for (int k = 0; k < 7; ++k) resid[7 * v + k] = r[k];
            for (int s = 0; s < 4; ++s) st3(arm, 4 * v + s, arms[s]);
        }
    }
    double out[3];
    block_sum<3, kThreads>(e, lds, out);
    if (threadIdx.x == 0) {
        double* p = partials + 5 * block;
        p[0] = out[0];
        p[1] = out[1];
        p[2] = 0.0;
        p[3] = 0.0;
        p[4] = out[2];
    }
}

// ------------------------------------------------------------ pass 2 -------

// Block `block` of the graph terms (smoothness per edge, anchor per node) and
// its energy partials. lds: 2 x waves doubles.
__device__ __forceinline__ void graph_energy_block(const AssemblyInputs& in, double* __restrict__ partials, int64_t block,
                                          double* lds) {
    const int64_t i = block * kThreads + threadIdx.x;
    double e[2] = {0.0, 0.0};
    if (i < in.edge_count && in.w.smooth > 0.0) {
        const int64_t a = in.edges[2 * i], b = in.edges[2 * i + 1];
        if (a >= 0 && b >= 0 && a < in.node_count && b < in.node_count) {
            const double sc = sqrt(in.w.smooth);
            const d3 gb = ld3(in.rest, b);
            const d3 diff = (apply_node(in.rotations, in.translations, in.rest, a, gb) -
                             apply_node(in.rotations, in.translations, in.rest, b, gb)) * sc;
            e[0] = (diff.x * diff.x + diff.y * diff.y) + diff.z * diff.z;
        }
    }
    if (i < in.node_count && in.w.anchor > 0.0) {
        const double sc = sqrt(in.w.anchor);
        const d3 w = log_so3(in.rotations + 9 * i) * sc;
        const d3 t = ld3(in.translations, i) * sc;
        e[1] = (((((w.x * w.x + w.y * w.y) + w.z * w.z) + t.x * t.x) + t.y * t.y) + t.z * t.z);
    }
    double out[2];
    block_sum<2, kThreads>(e, lds, out);
    if (threadIdx.x == 0) {
        double* p = partials + 5 * block;
        p[0] = 0.0;
        p[1] = 0.0;
        p[2] = out[0];
        p[3] = out[1];
        p[4] = 0.0;
    }
}

// The five energies from every block's partials, in a fixed order (thread t
// adds blocks t, t + 256, ...; then the block sum). With `host` (pinned host
// memory as the device addresses it), also writes `before` (nb values), the
// five energies and `after` (na values) there, in that order: the readback
// that follows an energy pass, without a launch of its own.
__device__ __forceinline__ void fold_partials(const double* __restrict__ partials, int blocks, double* lds, double (&s)[5]) {
    double e[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < blocks; b += kThreads)
        for (int k = 0; k < 5; ++k) e[k] += partials[5 * b + k];
    block_sum<5, kThreads>(e, lds, s);
}

__global__ void __launch_bounds__(kThreads) reduce_partials(const double* __restrict__ partials, int blocks,
                                                            double* __restrict__ out, double* __restrict__ host,
                                                            const double* __restrict__ before, int nb,
                                                            const double* __restrict__ after, int na) {
    __shared__ double lds[5 * (kThreads / kWave)];
    double s[5];
    fold_partials(partials, blocks, lds, s);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 5; ++k) out[k] = s[k];
        if (host) {
            for (int i = 0; i < nb; ++i) host[i] = before[i];
            for (int k = 0; k < 5; ++k) host[nb + k] = s[k];
            for (int i = 0; i < na; ++i) host[nb + 5 + i] = after[i];
        }
    }
}

// The rows and graph parts of an energy pass in one launch: blocks
// [0, vblocks) are the vertex rows, the rest the graph terms, each writing
// its own partials (reduce_partials folds them). Folding in the last block to
// finish instead was measured slower: each block's agent-scope arrival writes
// back and invalidates its XCD's L2.
template <bool Records>
__global__ void __launch_bounds__(kThreads)
energy_rows(AssemblyInputs in, double* __restrict__ scale, double* __restrict__ resid, double* __restrict__ arm,
            double* __restrict__ partials, int vblocks) {
    __shared__ double lds[3 * (kThreads / kWave)];
    if (static_cast<int>(blockIdx.x) < vblocks)
        vertex_rows_block<Records>(in, scale, resid, arm, partials, blockIdx.x, lds);
    else
        graph_energy_block(in, partials + 5 * static_cast<int64_t>(vblocks), blockIdx.x - vblocks, lds);
}

// ------------------------------------------------------------ pass 3 -------

// Upper triangle of a symmetric 6x6 in row order, plus 6 rhs entries.
__device__ constexpr int U(int r, int c) { return r * 6 - (r * (r - 1)) / 2 + (c - r); }
constexpr int kUpper = 21;

// Block part (Acc arithmetic) and right-hand-side part (FP64) of a 3-row
// offset block with arm a, scale c and residual r.
template <typename Acc>
__device__ inline void add_offset_block(Acc* blk, Acc c, const Acc (&av)[3]) {
    const Acc c2 = c * c;
    const Acc aa = (av[0] * av[0] + av[1] * av[1]) + av[2] * av[2];
    for (int x = 0; x < 3; ++x)
        for (int y = x; y < 3; ++y) blk[U(x, y)] += c2 * ((x == y ? aa : Acc(0)) - av[x] * av[y]);
    // [a]x = [[0,-az,ay],[az,0,-ax],[-ay,ax,0]]
    blk[U(0, 4)] += c2 * -av[2];
    blk[U(0, 5)] += c2 * av[1];
    blk[U(1, 3)] += c2 * av[2];
    blk[U(1, 5)] += c2 * -av[0];
    blk[U(2, 3)] += c2 * -av[1];
    blk[U(2, 4)] += c2 * av[0];
    blk[U(3, 3)] += c2;
    blk[U(4, 4)] += c2;
    blk[U(5, 5)] += c2;
}

__device__ inline void add_offset_rhs(double* rhs, double c, d3 a, const double* r) {
    const d3 rv{r[0], r[1], r[2]};
    const d3 m = cross(a, rv);
    rhs[0] -= c * m.x;
    rhs[1] -= c * m.y;
    rhs[2] -= c * m.z;
    rhs[3] -= c * rv.x;
    rhs[4] -= c * rv.y;
    rhs[5] -= c * rv.z;
}

template <typename Acc>
__device__ inline void add_offset_rows(Acc* blk, double* rhs, double c, d3 a, const double* r) {
    const Acc av[3] = {static_cast<Acc>(a.x), static_cast<Acc>(a.y), static_cast<Acc>(a.z)};
    add_offset_block<Acc>(blk, static_cast<Acc>(c), av);
    add_offset_rhs(rhs, c, a, r);
}

template <typename Acc>
__device__ inline void add_plane_row(Acc* blk, double* rhs, double c, d3 a, d3 n, double value) {
    const d3 an = cross(a, n);
    const double q[6] = {c * an.x, c * an.y, c * an.z, c * n.x, c * n.y, c * n.z};
    if constexpr (sizeof(Acc) == sizeof(double)) {
        for (int x = 0; x < 6; ++x)
            for (int y = x; y < 6; ++y) blk[U(x, y)] += q[x] * q[y];
    } else {
        // The Jacobian row in FP32: arm, normal and scale rounded, then the
        // cross product and the outer product in FP32.
        const Acc cf = static_cast<Acc>(c);
        const Acc af[3] = {static_cast<Acc>(a.x), static_cast<Acc>(a.y), static_cast<Acc>(a.z)};
        const Acc nf[3] = {static_cast<Acc>(n.x), static_cast<Acc>(n.y), static_cast<Acc>(n.z)};
        const Acc qf[6] = {cf * (af[1] * nf[2] - af[2] * nf[1]), cf * (af[2] * nf[0] - af[0] * nf[2]),
                           cf * (af[0] * nf[1] - af[1] * nf[0]), cf * nf[0], cf * nf[1], cf * nf[2]};
        for (int x = 0; x < 6; ++x)
            for (int y = x; y < 6; ++y) blk[U(x, y)] += qf[x] * qf[y];
    }
    for (int x = 0; x < 6; ++x) rhs[x] -= q[x] * value;
}

// Blocks = false: the right-hand side only, with its rotation half written as
// zero (the translation system of the local/global route). The block is
// accumulated in Acc (FP64, or FP32 for the all-FP32 assembly), the
// right-hand side always in FP64. `copy32`, when given, receives an FP32
// copy of the (FP64) block.
template <typename Scalar, bool Blocks, typename Acc>
__device__ __forceinline__ void node_block_rows(const AssemblyInputs& in, const double* __restrict__ scale,
                                       const double* __restrict__ resid, const double* __restrict__ arm,
                                       const int32_t* __restrict__ goff, const int32_t* __restrict__ gent,
                                       const int32_t* __restrict__ eoff, const int32_t* __restrict__ eent,
                                       const int32_t* __restrict__ diag_pos, Scalar* __restrict__ bsr,
                                       double* __restrict__ rhs, float* __restrict__ copy32, int64_t block,
                                       const int32_t* __restrict__ order) {
    const int lane = threadIdx.x % kWave;
    const int64_t slot = block * (kThreads / kWave) + threadIdx.x / kWave;
    if (slot >= in.node_count) return;
    const int64_t node = order ? order[slot] : slot;
    Acc blk[kUpper];
    double g[6];
    for (int k = 0; k < kUpper; ++k) blk[k] = Acc(0);
    for (int k = 0; k < 6; ++k) g[k] = 0.0;
    for (int32_t k = goff[node] + lane; k < goff[node + 1]; k += kWave) {
        const int32_t ent = gent[k];
        const int64_t v = ent >> 2;
        const int s = ent & 3;
        const double w = in.bind_weights[4 * v + s];
        const d3 a = ld3(arm, 4 * v + s);
        const double sp = scale[3 * v], sl = scale[3 * v + 1], sm = scale[3 * v + 2];
        if (sp >= 0.0) add_offset_rows<Acc>(blk, g, sp * w, a, resid + 7 * v);
        if (sl >= 0.0) add_plane_row<Acc>(blk, g, sl * w, a, ld3(in.corr_normals, v), resid[7 * v + 3]);
        if (sm >= 0.0) add_offset_rows<Acc>(blk, g, sm * w, a, resid + 7 * v + 4);
    }
    if (in.w.smooth > 0.0) {
        const double sc = sqrt(in.w.smooth);
        for (int32_t k = eoff[node] + lane; k < eoff[node + 1]; k += kWave) {
            const int32_t ent = eent[k];
            const int64_t e = ent >> 1;
            const int64_t a = in.edges[2 * e], b = in.edges[2 * e + 1];
            if (a < 0 || b < 0 || a >= in.node_count || b >= in.node_count) continue;
            const d3 ga = ld3(in.rest, a), gb = ld3(in.rest, b);
            const d3 diff = (apply_node(in.rotations, in.translations, in.rest, a, gb) -
                             apply_node(in.rotations, in.translations, in.rest, b, gb)) * sc;
            const double r[3] = {diff.x, diff.y, diff.z};
            if ((ent & 1) == 0) {
                add_offset_rows<Acc>(blk, g, sc, mul3(in.rotations + 9 * a, gb - ga), r);
            } else {
                const Acc s2 = static_cast<Acc>(sc * sc);
                blk[U(3, 3)] += s2;
                blk[U(4, 4)] += s2;
                blk[U(5, 5)] += s2;
                g[3] += sc * r[0];
                g[4] += sc * r[1];
                g[5] += sc * r[2];
            }
        }
    }
    if (lane == 0 && in.w.anchor > 0.0) {
        const double sc2 = in.w.anchor;
        const double* R = in.rotations + 9 * node;
        const d3 w = log_so3(R);
        double J[9];
        left_jacobian_inverse(w, J);
        const double wv[3] = {w.x, w.y, w.z};
        for (int x = 0; x < 3; ++x) {
            for (int y = x; y < 3; ++y) {
                double s = 0.0;
                for (int r = 0; r < 3; ++r) s += J[r * 3 + x] * J[r * 3 + y];
                blk[U(x, y)] += static_cast<Acc>(sc2 * s);
            }
            double gr = 0.0;
            for (int r = 0; r < 3; ++r) gr += J[r * 3 + x] * wv[r];
            g[x] -= sc2 * gr;
        }
        const d3 t = ld3(in.translations, node);
        blk[U(3, 3)] += static_cast<Acc>(sc2);
        blk[U(4, 4)] += static_cast<Acc>(sc2);
        blk[U(5, 5)] += static_cast<Acc>(sc2);
        g[3] -= sc2 * t.x;
        g[4] -= sc2 * t.y;
        g[5] -= sc2 * t.z;
    }
    for (int k = 0; k < kUpper; ++k) blk[k] = wave_sum(blk[k]);
    for (int k = 0; k < 6; ++k) g[k] = wave_sum(g[k]);
    if (lane == 0) {
        if (Blocks) {
            Scalar* out = bsr + 36 * static_cast<int64_t>(diag_pos[node]);
            for (int x = 0; x < 6; ++x)
                for (int y = 0; y < 6; ++y)
                    out[6 * x + y] = static_cast<Scalar>(x <= y ? blk[U(x, y)] : blk[U(y, x)]);
            if (copy32) {
                float* c = copy32 + 36 * static_cast<int64_t>(diag_pos[node]);
                for (int x = 0; x < 6; ++x)
                    for (int y = 0; y < 6; ++y) c[6 * x + y] = static_cast<float>(x <= y ? blk[U(x, y)] : blk[U(y, x)]);
            }
        }
        for (int x = 0; x < 6; ++x) rhs[6 * node + x] = (Blocks || x >= 3) ? g[x] : 0.0;
    }
}

template <typename Scalar, bool Blocks, typename Acc>
__global__ void __launch_bounds__(kThreads)
node_blocks(AssemblyInputs in, const double* __restrict__ scale, const double* __restrict__ resid,
            const double* __restrict__ arm, const int32_t* __restrict__ goff, const int32_t* __restrict__ gent,
            const int32_t* __restrict__ eoff, const int32_t* __restrict__ eent, const int32_t* __restrict__ diag_pos,
            Scalar* __restrict__ bsr, double* __restrict__ rhs, float* __restrict__ copy32,
            const int32_t* __restrict__ order) {
    node_block_rows<Scalar, Blocks, Acc>(in, scale, resid, arm, goff, gent, eoff, eent, diag_pos, bsr, rhs, copy32,
                                         xcd_contiguous(blockIdx.x, gridDim.x), order);
}

// ------------------------------------------------------------ pass 4 -------

template <typename Acc>
__device__ inline void add_pair_offset(Acc* acc, Acc clo, const Acc (&lo)[3], Acc chi, const Acc (&hi)[3]) {
    const Acc cc = clo * chi;
    const Acc d = (lo[0] * hi[0] + lo[1] * hi[1]) + lo[2] * hi[2];
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y) acc[6 * x + y] += cc * ((x == y ? d : Acc(0)) - hi[x] * lo[y]);
    // top right [alo]x, bottom left [ahi]x^T
    acc[6 * 0 + 4] += cc * -lo[2];
    acc[6 * 0 + 5] += cc * lo[1];
    acc[6 * 1 + 3] += cc * lo[2];
    acc[6 * 1 + 5] += cc * -lo[0];
    acc[6 * 2 + 3] += cc * -lo[1];
    acc[6 * 2 + 4] += cc * lo[0];
    acc[6 * 4 + 0] += cc * -hi[2];
    acc[6 * 5 + 0] += cc * hi[1];
    acc[6 * 3 + 1] += cc * hi[2];
    acc[6 * 5 + 1] += cc * -hi[0];
    acc[6 * 3 + 2] += cc * -hi[1];
    acc[6 * 4 + 2] += cc * hi[0];
    acc[6 * 3 + 3] += cc;
    acc[6 * 4 + 4] += cc;
    acc[6 * 5 + 5] += cc;
}

template <typename Acc>
__device__ inline void acc3(Acc (&out)[3], d3 v) {
    out[0] = static_cast<Acc>(v.x);
    out[1] = static_cast<Acc>(v.y);
    out[2] = static_cast<Acc>(v.z);
}

// Eight lanes per interacting pair: the lanes split the pair's entries, then
// sum their 36 partial block entries with shuffles in a fixed order. The
// block is accumulated in Acc (FP64, or FP32 for the all-FP32 assembly);
// `copy32`, when given, receives an FP32 copy of it.
constexpr int kPairLanes = 8;

template <typename Scalar, typename Acc>
__device__ __forceinline__ void pair_block_rows(const AssemblyInputs& in, const double* __restrict__ scale,
                                       const double* __restrict__ arm, int64_t groups,
                                       const int32_t* __restrict__ goff, const int32_t* __restrict__ gent,
                                       const int32_t* __restrict__ pos, Scalar* __restrict__ bsr,
                                       int32_t* __restrict__ touched, float* __restrict__ copy32, int64_t block,
                                       const int32_t* __restrict__ order) {
    const int sub = threadIdx.x % kPairLanes;
    const int64_t slot = (block * kThreads + threadIdx.x) / kPairLanes;
    const bool valid = slot < groups;
    const int64_t g = valid && order ? order[slot] : slot;
    Acc acc[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) acc[k] = Acc(0);
    int hit = 0;
    const int32_t end = valid ? goff[g + 1] : 0;
    for (int32_t k = valid ? goff[g] + sub : 0; k < end; k += kPairLanes) {
        const int32_t ent = gent[k];
        if (ent >= 0) {
            const int64_t v = ent >> 4;
            const int slo = (ent >> 2) & 3, shi = ent & 3;
            const double wlo = in.bind_weights[4 * v + slo], whi = in.bind_weights[4 * v + shi];
            Acc alo[3], ahi[3];
            acc3(alo, ld3(arm, 4 * v + slo));
            acc3(ahi, ld3(arm, 4 * v + shi));
            const double sp = scale[3 * v], sl = scale[3 * v + 1], sm = scale[3 * v + 2];
            if (sp >= 0.0) {
                hit = 1;
                add_pair_offset<Acc>(acc, static_cast<Acc>(sp * wlo), alo, static_cast<Acc>(sp * whi), ahi);
            }
            if (sl >= 0.0) {
                hit = 1;
                Acc n[3];
                acc3(n, ld3(in.corr_normals, v));
                const Acc cl = static_cast<Acc>(sl * wlo), ch = static_cast<Acc>(sl * whi);
                const Acc ql[6] = {cl * (alo[1] * n[2] - alo[2] * n[1]), cl * (alo[2] * n[0] - alo[0] * n[2]),
                                   cl * (alo[0] * n[1] - alo[1] * n[0]), cl * n[0], cl * n[1], cl * n[2]};
                const Acc qh[6] = {ch * (ahi[1] * n[2] - ahi[2] * n[1]), ch * (ahi[2] * n[0] - ahi[0] * n[2]),
                                   ch * (ahi[0] * n[1] - ahi[1] * n[0]), ch * n[0], ch * n[1], ch * n[2]};
#pragma unroll
                for (int x = 0; x < 6; ++x)
#pragma unroll
                    for (int y = 0; y < 6; ++y) acc[6 * x + y] += ql[x] * qh[y];
            }
            if (sm >= 0.0) {
                hit = 1;
                add_pair_offset<Acc>(acc, static_cast<Acc>(sm * wlo), alo, static_cast<Acc>(sm * whi), ahi);
            }
        } else if (in.w.smooth > 0.0) {
            const int64_t e = ~ent;
            const int64_t a = in.edges[2 * e], b = in.edges[2 * e + 1];
            hit = 1;
            const Acc s2 = static_cast<Acc>(in.w.smooth);
            Acc rot[3];
            acc3(rot, mul3(in.rotations + 9 * a, ld3(in.rest, b) - ld3(in.rest, a)));
            // J_a J_b^T: columns 3..5 hold -s^2 [ [rot]x ; I ]; transposed
            // into the (lower, higher) block when a > b.
            const bool fwd = a < b;
#define NRICP_PAIR_ADD(r, c, v)                  \
    do {                                         \
        if (fwd)                                 \
            acc[6 * (r) + (c)] += (v);           \
        else                                     \
            acc[6 * (c) + (r)] += (v);           \
    } while (0)
            NRICP_PAIR_ADD(0, 4, -s2 * -rot[2]);
            NRICP_PAIR_ADD(0, 5, -s2 * rot[1]);
            NRICP_PAIR_ADD(1, 3, -s2 * rot[2]);
            NRICP_PAIR_ADD(1, 5, -s2 * -rot[0]);
            NRICP_PAIR_ADD(2, 3, -s2 * -rot[1]);
            NRICP_PAIR_ADD(2, 4, -s2 * rot[0]);
            NRICP_PAIR_ADD(3, 3, -s2);
            NRICP_PAIR_ADD(4, 4, -s2);
            NRICP_PAIR_ADD(5, 5, -s2);
#undef NRICP_PAIR_ADD
        }
    }
#pragma unroll
    for (int k = 0; k < 36; ++k) acc[k] = octet_sum(acc[k]);
    hit = octet_sum(hit);
    if (!valid || sub != 0) return;
    Scalar* ab = bsr + 36 * static_cast<int64_t>(pos[2 * g]);
    Scalar* ba = bsr + 36 * static_cast<int64_t>(pos[2 * g + 1]);
#pragma unroll
    for (int x = 0; x < 6; ++x)
#pragma unroll
        for (int y = 0; y < 6; ++y) {
            ab[6 * x + y] = static_cast<Scalar>(acc[6 * x + y]);
            ba[6 * y + x] = static_cast<Scalar>(acc[6 * x + y]);
        }
    if (copy32) {
        float* cab = copy32 + 36 * static_cast<int64_t>(pos[2 * g]);
        float* cba = copy32 + 36 * static_cast<int64_t>(pos[2 * g + 1]);
#pragma unroll
        for (int x = 0; x < 6; ++x)
#pragma unroll
            for (int y = 0; y < 6; ++y) {
                cab[6 * x + y] = static_cast<float>(acc[6 * x + y]);
                cba[6 * y + x] = static_cast<float>(acc[6 * x + y]);
            }
    }
    if (touched) touched[g] = hit != 0 ? 1 : 0;
}

// The node blocks (node_groups workgroups) and the pair blocks (the rest) of
// one assembly in one launch: both only read the vertex records, so their
// gathers overlap instead of running one launch after the other. Nodes and
// groups are taken in the layout's space order, and each XCD takes one
// contiguous eighth of the nodes and of the groups, so the records one region
// of the mesh needs are gathered through one L2.
template <typename Scalar>
__global__ void __launch_bounds__(kThreads)
system_blocks(AssemblyInputs in, const double* __restrict__ scale, const double* __restrict__ resid,
              const double* __restrict__ arm, const int32_t* __restrict__ goff, const int32_t* __restrict__ gent,
              const int32_t* __restrict__ eoff, const int32_t* __restrict__ eent, const int32_t* __restrict__ diag_pos,
              int64_t node_groups, int64_t pairs, const int32_t* __restrict__ poff, const int32_t* __restrict__ pent,
              const int32_t* __restrict__ ppos, Scalar* __restrict__ bsr, double* __restrict__ rhs,
              int32_t* __restrict__ touched, float* __restrict__ copy32, const int32_t* __restrict__ node_order,
              const int32_t* __restrict__ group_order) {
    int kind = 0;
    const int64_t b = xcd_split(blockIdx.x, gridDim.x, node_groups, &kind);
    if (kind == 0)
        node_block_rows<Scalar, true, Scalar>(in, scale, resid, arm, goff, gent, eoff, eent, diag_pos, bsr, rhs,
                                              copy32, b, node_order);
    else
        pair_block_rows<Scalar, Scalar>(in, scale, arm, pairs, poff, pent, ppos, bsr, touched, copy32, b,
                                        group_order);
}

// ---------------------------------------------------------------- packed ---
// The assembly and the energy passes for a batch of jobs (packed.hpp). A
// job's blocks are its single-job launch's blocks, so its records, blocks and
// energies are the same bits packed or alone.

// Records: the current state's rows (jobs that refresh or assemble this
// tick); otherwise the candidate state's energies (jobs that solve).
template <bool Records>
__global__ void __launch_bounds__(kThreads)
rows_packed(const PackedJob* __restrict__ jobs, const BatchControl* __restrict__ ctl,
            const PackedBlock* __restrict__ map) {
    __shared__ double lds[3 * (kThreads / kWave)];
    const PackedBlock m = map[xcd_contiguous(blockIdx.x, gridDim.x)];
    const int32_t flags = ctl[m.job].flags;
    if (!(flags & (Records ? (kJobRefresh | kJobAssemble) : kJobActive))) return;
    const PackedJob& J = jobs[m.job];
    AssemblyInputs in = J.in;
    if (!Records) {
        in.rotations = J.rot_cand;
        in.translations = J.trans_cand;
    }
    if (m.block < J.vertex_blocks)
        vertex_rows_block<Records>(in, J.scale, J.resid, J.arm, J.partials, m.block, lds);
    else
        graph_energy_block(in, J.partials + 5 * static_cast<int64_t>(J.vertex_blocks), m.block - J.vertex_blocks,
                           lds);
}

// Scalar: the blocks' type; Copy32: the mixed form (FP64 blocks into bsr64,
// their FP32 copy into bsr).
template <typename Scalar, bool Copy32>
__global__ void __launch_bounds__(kThreads)
system_packed(const PackedJob* __restrict__ jobs, const BatchControl* __restrict__ ctl,
              const PackedBlock* __restrict__ map) {
    const PackedBlock m = map[xcd_contiguous(blockIdx.x, gridDim.x)];
    if (!(ctl[m.job].flags & (kJobRefresh | kJobAssemble))) return;
    const PackedJob& J = jobs[m.job];
    Scalar* bsr = Copy32 ? reinterpret_cast<Scalar*>(J.bsr64) : static_cast<Scalar*>(J.bsr);
    float* copy32 = Copy32 ? static_cast<float*>(J.bsr) : nullptr;
    if (m.block < J.node_blocks)
        node_block_rows<Scalar, true, Scalar>(J.in, J.scale, J.resid, J.arm, J.gather_offsets, J.gather_entries,
                                              J.edge_offsets, J.edge_entries, J.diag_position, bsr, J.rhs, copy32,
                                              m.block, J.node_order);
    else
        pair_block_rows<Scalar, Scalar>(J.in, J.scale, J.arm, J.group_count, J.group_offsets, J.group_entries,
                                        J.pair_position, bsr, J.touched, copy32, m.block - J.node_blocks,
                                        J.group_order);
}

// One workgroup per job: the current state's energies from its partials.
__global__ void __launch_bounds__(kThreads)
reduce_current_packed(const PackedJob* __restrict__ jobs, const BatchControl* __restrict__ ctl) {
    __shared__ double lds[5 * (kThreads / kWave)];
    const int job = blockIdx.x;
    if (!(ctl[job].flags & (kJobRefresh | kJobAssemble))) return;
    const PackedJob& J = jobs[job];
    double e[5];
    fold_partials(J.partials, J.vertex_blocks + J.graph_blocks, lds, e);
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; ++k) J.energy_cur[k] = e[k];
}

__device__ inline double total5(const double* e) { return (((e[0] + e[1]) + e[2]) + e[3]) + e[4]; }

__device__ inline double relative_change_dev(double before, double after) {
    return (before - after) / fmax(fabs(before), 1e-12);
}

// One workgroup per job: the candidate's energies, then the decision the
// single-job session takes after a damping attempt (route A semantics,
// src/solver/route_a.cpp:131-169, with the outer loop's close): accept iff
// the solve converged and the candidate's energy is below the current one;
// damping times the decrease (floor 1e-12) on accept, times the increase on
// reject or a failed solve; a new inner step after an accept, a retry while
// attempts remain, else the outer iteration closes, and the level ends when
// its outer loop converges or runs out.
__global__ void __launch_bounds__(kThreads)
decide_packed(const PackedJob* __restrict__ jobs, BatchControl* __restrict__ ctl, int32_t* __restrict__ host_flags) {
    __shared__ double lds[5 * (kThreads / kWave)];
    const int job = blockIdx.x;
    BatchControl& c = ctl[job];
    const int32_t flags = c.flags;
    if (!(flags & kJobActive)) {
        if (threadIdx.x == 0) {
            c.flags = flags & ~kJobAccepted;
            host_flags[job] = c.flags;
        }
        return;
    }
    const PackedJob& J = jobs[job];
    double cand_e[5];
    fold_partials(J.partials, J.vertex_blocks + J.graph_blocks, lds, cand_e);
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 5; ++k) J.energy_cand[k] = cand_e[k];
    const double cur = total5(J.energy_cur);
    if (c.attempt == 0) {
        for (int k = 0; k < 5; ++k) c.last_energy[k] = J.energy_cur[k];
        c.current_energy = cur;
        if (c.outer == 0 && c.inner == 0) c.energy_start = cur;
    }
    const double iterations = J.outcome[0];
    c.pcg_launches += 1.0;
    c.pcg_iterations += iterations;
    c.pcg_most = fmax(c.pcg_most, iterations);
    c.gn_steps += 1.0;
    bool accepted = false;
    double cand = 0.0;
    if (J.outcome[2] != 0.0) {
        cand = total5(cand_e);
        accepted = cand < cur;
    }
    int32_t next = 0;
    bool close = false;
    if (accepted) {
        for (int k = 0; k < 5; ++k) {
            c.last_energy[k] = cand_e[k];
            J.energy_cur[k] = cand_e[k];
        }
        c.current_energy = cand;
        c.damping = fmax(c.damping * c.damping_decrease, 1e-12);
        c.inner += 1;
        c.view_inner += 1;
        c.attempt = 0;
        next = kJobAccepted;
        if (relative_change_dev(cur, cand) <= c.inner_tolerance || c.inner >= max(1, c.max_inner))
            close = true;
        else
            next |= kJobActive | kJobAssemble;
    } else {
        c.damping *= c.damping_increase;
        c.attempt += 1;
        if (c.attempt < max(1, c.max_attempts))
            next = kJobActive;
        else
            close = true;
    }
    if (close) {
        const double level_energy = total5(c.last_energy);
        c.energy_end = level_energy;
        c.view_outer += 1;
        bool converged = false;
        if (c.outer > 0 && relative_change_dev(c.previous_level_energy, level_energy) <= c.outer_tolerance) {
            c.converged = 1;
            converged = true;
        }
        c.previous_level_energy = level_energy;
        c.outer += 1;
        c.attempt = 0;
        c.inner = 0;
        if (!converged && c.outer < max(1, c.max_outer)) next |= kJobActive | kJobRefresh;
    }
    c.flags = next;
    host_flags[job] = next;
}

void launch_energies(const AssemblyInputs& in, AssemblyScratch& s, bool records, double* out, hipStream_t stream,
                     const HostReadback* readback = nullptr) {
    const int vblocks = std::max(1, blocks_for(in.vertex_count, kThreads));
    const int gblocks = std::max(1, blocks_for(std::max(in.edge_count, in.node_count), kThreads));
    s.partials.resize(5 * static_cast<int64_t>(vblocks + gblocks));
    const int blocks = vblocks + gblocks;
    if (records) {
        s.vertex_scale.resize(3 * std::max<int64_t>(in.vertex_count, 1));
        s.vertex_resid.resize(7 * std::max<int64_t>(in.vertex_count, 1));
        s.vertex_arm.resize(12 * std::max<int64_t>(in.vertex_count, 1));
        energy_rows<true><<<blocks, kThreads, 0, stream>>>(in, s.vertex_scale.get(), s.vertex_resid.get(),
                                                            s.vertex_arm.get(), s.partials.get(), vblocks);
    } else {
        energy_rows<false><<<blocks, kThreads, 0, stream>>>(in, nullptr, nullptr, nullptr, s.partials.get(), vblocks);
    }
    NRICP_LAUNCH_CHECK("energy_rows");
    const HostReadback none{};
    const HostReadback& r = readback ? *readback : none;
    reduce_partials<<<1, kThreads, 0, stream>>>(s.partials.get(), blocks, out, r.host, r.before, r.before_count,
                                                r.after, r.after_count);
    NRICP_LAUNCH_CHECK("reduce_partials");
}

}  // namespace

void energies(const AssemblyInputs& in, AssemblyScratch& s, double* out, hipStream_t stream,
              const HostReadback* readback) {
    launch_energies(in, s, false, out, stream, readback);
}

// The block arithmetic follows the matrix type: FP64 blocks (optionally with
// an FP32 copy for the mixed form's iterations) from FP64 accumulation, FP32
// blocks from FP32 arithmetic and accumulation. The right-hand side and the
// energies are FP64 either way.
template <typename Scalar>
void assemble(const AssemblyInputs& in, const LevelLayout& L, AssemblyScratch& s, Scalar* bsr, double* rhs,
              double* energies_out, hipStream_t stream, float* copy32) {
    if (sizeof(Scalar) == sizeof(float) && copy32) throw InvalidArgument("assemble: an FP32 matrix takes no copy");
    launch_energies(in, s, true, energies_out, stream);
    const int per_block = kThreads / kWave;
    const int64_t node_groups = blocks_for(in.node_count, per_block);
    const int64_t pair_groups = L.group_count > 0 ? blocks_for(L.group_count * kPairLanes, kThreads) : 0;
    s.touched.resize(std::max<int64_t>(L.group_count, 1));
    if (node_groups + pair_groups <= 0) return;
    system_blocks<Scalar><<<static_cast<unsigned>(node_groups + pair_groups), kThreads, 0, stream>>>(
        in, s.vertex_scale.get(), s.vertex_resid.get(), s.vertex_arm.get(), L.gather_offsets.get(),
        L.gather_entries.get(), L.node_edge_offsets.get(), L.node_edge_entries.get(), L.diag_position.get(),
        node_groups, L.group_count, L.group_offsets.get(), L.group_entries.get(), L.pair_position.get(), bsr, rhs,
        s.touched.get(), copy32, L.node_order.get(), L.group_order.get());
    NRICP_LAUNCH_CHECK("system_blocks");
}

void translation_rhs(const AssemblyInputs& in, const LevelLayout& L, AssemblyScratch& s, double* rhs,
                     double* energies_out, hipStream_t stream) {
    launch_energies(in, s, true, energies_out, stream);
    const int per_block = kThreads / kWave;
    node_blocks<double, false, double><<<blocks_for(in.node_count, per_block), kThreads, 0, stream>>>(
        in, s.vertex_scale.get(), s.vertex_resid.get(), s.vertex_arm.get(), L.gather_offsets.get(),
        L.gather_entries.get(), L.node_edge_offsets.get(), L.node_edge_entries.get(), L.diag_position.get(), nullptr,
        rhs, nullptr, L.node_order.get());
    NRICP_LAUNCH_CHECK("node_blocks");
}

void packed_rows(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, bool records,
                 KernelSink& sink) {
    if (maps.rows_blocks <= 0) return;
    const unsigned blocks = static_cast<unsigned>(maps.rows_blocks);
    if (records)
        sink_launch(sink, rows_packed<true>, blocks, kThreads, jobs, ctl, maps.rows);
    else
        sink_launch(sink, rows_packed<false>, blocks, kThreads, jobs, ctl, maps.rows);
}

void packed_system(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, int precision,
                   KernelSink& sink) {
    if (maps.system_blocks <= 0) return;
    const unsigned blocks = static_cast<unsigned>(maps.system_blocks);
    if (precision == NRICP_HIP_PRECISION_FLOAT)
        sink_launch(sink, system_packed<float, false>, blocks, kThreads, jobs, ctl, maps.system);
    else if (precision == NRICP_HIP_PRECISION_MIXED)
        sink_launch(sink, system_packed<double, true>, blocks, kThreads, jobs, ctl, maps.system);
    else
        sink_launch(sink, system_packed<double, false>, blocks, kThreads, jobs, ctl, maps.system);
}

void packed_reduce_current(const PackedJob* jobs, const BatchControl* ctl, int32_t job_count, KernelSink& sink) {
    if (job_count <= 0) return;
    sink_launch(sink, reduce_current_packed, static_cast<unsigned>(job_count), kThreads, jobs, ctl);
}

void packed_decide(const PackedJob* jobs, BatchControl* ctl, int32_t job_count, int32_t* host_flags,
                   KernelSink& sink) {
    if (job_count <= 0) return;
    sink_launch(sink, decide_packed, static_cast<unsigned>(job_count), kThreads, jobs, ctl, host_flags);
}

template void assemble<double>(const AssemblyInputs&, const LevelLayout&, AssemblyScratch&, double*, double*, double*,
                               hipStream_t, float*);
template void assemble<float>(const AssemblyInputs&, const LevelLayout&, AssemblyScratch&, float*, double*, double*,
                              hipStream_t, float*);

}  // namespace nricp_hip
