// Block-Jacobi preconditioned conjugate gradients on the damped block system,
// as one persistent launch per solve.
//
// Reference behaviour: replaces src/cuda/pcg_solver.cu:15-326 (block-Jacobi
// PCG from x0 = 0, stop when |r|/|b| <= tol, ok <=> converged). Differences,
// chosen for gfx950:
//   * pipelined CG recurrence (Ghysels and Vanroose): the inner products and
//     the matrix product of an iteration are independent, so an iteration
//     costs one grid barrier, not four;
//   * the matrix is block-sparse rows of the full symmetric system, so the
//     product is a gather with no atomics: the eight lanes of a node split its
//     blocks and sum their partial products with wave shuffles;
//   * each lane keeps its rows of every CG vector and of the preconditioner in
//     registers for the whole solve; only the vector the product needs
//     crosses lanes (through memory, for the neighbours' products);
//   * damping is the CPU route's (src/solver/normal_equations.cpp:81-96):
//     lambda * max(A_ii, 1e-12) with diagonal scaling, lambda otherwise;
//   * inner products and the refinement residual are accumulated in FP64 and
//     reduced in a fixed order, so a solve is reproducible bit for bit;
//   * convergence is decided on the true residual b - A x, evaluated in FP64;
//     mixed mode runs the iteration in FP32 on an FP32 matrix and refines in
//     FP64 inside the same launch: x += e, r = b - A x, restart, until the
//     FP64 relative residual reaches the tolerance.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "packed.hpp"
#include "xcd_map.hpp"

namespace nricp_hip {
namespace {

constexpr int kPcgThreads = 512;
constexpr int kNodesPerPass = kPcgThreads / 8;  // 8 lanes per node, 6 active
constexpr int kBarrierGroups = 32;

// Inverse of the symmetric 6x6 block `a` (lower triangle read) plus `jitter`
// on its diagonal, through its Cholesky factor: L L^T = a, X = L^-1,
// inverse = X^T X. Fully unrolled, so the block never leaves registers.
// Returns false when a pivot is not positive.
__device__ inline bool spd_inverse6(const double (&a)[36], double jitter, double (&inv)[36]) {
    double L[36];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double d = a[6 * j + j] + jitter;
#pragma unroll
        for (int k = 0; k < j; ++k) d -= L[6 * j + k] * L[6 * j + k];
        ok = ok && d > 0.0;
        const double ljj = sqrt(d > 0.0 ? d : 1.0);
        L[6 * j + j] = ljj;
        const double r = 1.0 / ljj;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double t = a[6 * i + j];
#pragma unroll
            for (int k = 0; k < j; ++k) t -= L[6 * i + k] * L[6 * j + k];
            L[6 * i + j] = t * r;
        }
    }
    double X[36];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        X[6 * j + j] = 1.0 / L[6 * j + j];
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double t = 0.0;
#pragma unroll
            for (int k = j; k < i; ++k) t -= L[6 * i + k] * X[6 * k + j];
            X[6 * i + j] = t / L[6 * i + i];
        }
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) {
#pragma unroll
        for (int c = r; c < 6; ++c) {
            double t = 0.0;
#pragma unroll
            for (int k = c; k < 6; ++k) t += X[6 * k + r] * X[6 * k + c];
            inv[6 * r + c] = t;
            inv[6 * c + r] = t;
        }
    }
    return ok;
}

// Damping of the system (the CPU route's lambda * max(A_ii, 1e-12) with
// diagonal scaling, lambda otherwise) and the block-Jacobi preconditioner:
// the inverse of each damped diagonal block by Cholesky, retried with a
// growing diagonal jitter and replaced by the identity when the block is not
// positive definite (the recovery of the reference's factor kernel,
// src/cuda/pcg_solver.cu:15-46). The jitter only shapes the preconditioner;
// the system keeps its damping.
template <typename S>
__device__ __forceinline__ void precondition_node(const S* __restrict__ bsr, const double* __restrict__ bsr64,
                                         const int32_t* __restrict__ diag_pos, int64_t nn, double lambda, int scaling,
                                         S* __restrict__ damp, double* __restrict__ damp64, S* __restrict__ minv,
                                         unsigned* __restrict__ barrier, int64_t i) {
    // The solve's grid-barrier counter starts at zero (the solve runs next on
    // the same stream).
    if (i == 0) barrier[0] = barrier[1] = 0u;
    if (i >= nn) return;
    // With an FP64 copy of the matrix (the mixed form) the damping and the
    // preconditioner come from its diagonal block, and the FP64 damping is
    // kept for the refinement residual.
    double a[36];
    if (bsr64) {
        const double* blk = bsr64 + 36 * static_cast<int64_t>(diag_pos[i]);
#pragma unroll
        for (int k = 0; k < 36; ++k) a[k] = blk[k];
    } else {
        const S* blk = bsr + 36 * static_cast<int64_t>(diag_pos[i]);
#pragma unroll
        for (int k = 0; k < 36; ++k) a[k] = static_cast<double>(blk[k]);
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        double d = 0.0;
        if (lambda > 0.0) d = lambda * (scaling ? fmax(a[6 * r + r], 1e-12) : 1.0);
        damp[6 * i + r] = static_cast<S>(d);
        if (damp64) damp64[6 * i + r] = d;
        a[6 * r + r] += d;
    }
    double inv[36];
    bool ok = spd_inverse6(a, 0.0, inv);
    double jitter = 1e-6;
#pragma unroll 1
    for (int attempt = 1; attempt < 8 && !ok; ++attempt) {
        ok = spd_inverse6(a, jitter, inv);
        jitter *= 10.0;
    }
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c)
            minv[36 * i + 6 * r + c] = static_cast<S>(ok ? inv[6 * r + c] : (r == c ? 1.0 : 0.0));
}

template <typename S>
__global__ void precondition_kernel(const S* __restrict__ bsr, const double* __restrict__ bsr64,
                                    const int32_t* __restrict__ diag_pos, int64_t nn, double lambda, int scaling,
                                    S* __restrict__ damp, double* __restrict__ damp64, S* __restrict__ minv,
                                    unsigned* __restrict__ barrier) {
    precondition_node<S>(bsr, bsr64, diag_pos, nn, lambda, scaling, damp, damp64, minv, barrier,
                         static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
}

// The solve's groups: a whole launch (group blockIdx.x of gridDim.x), or one
// job's share of a packed launch, whose group index and count the packed
// kernel keeps in LDS (read where used, so the long solve holds no register
// for them).
struct LaunchGroups {
    __device__ static unsigned group() { return blockIdx.x; }
    __device__ static unsigned groups() { return gridDim.x; }
};
__shared__ unsigned packed_group_set[2];
struct PackedGroups {
    __device__ static unsigned group() { return packed_group_set[0]; }
    __device__ static unsigned groups() { return packed_group_set[1]; }
};

template <bool Grid, typename GS>
__device__ inline void sync_all(unsigned* barrier, GS gs) {
    if (Grid)
        grid_barrier(barrier, gs.groups());
    else
        __syncthreads();
}

// Reproducible sums over the whole launch. Each wave sums its threads'
// partials, thread 0 adds the group's wave sums in wave order (through LDS)
// and publishes them in the group's slot just before its barrier arrival, so
// the partial sums ride on the barrier. After it, fold(): every wave loads the
// group slots lane-strided and sums them, so every wave of every group
// obtains the same bits.
constexpr int kPcgWaves = kPcgThreads / kWave;

template <int Count, typename GS>
__device__ inline void publish_and_wait(const double (&v)[Count], double* lds /* Count x waves */, double* slots,
                                        unsigned* barrier, GS gs) {
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < Count; ++k) {
        const double w = wave_sum(v[k]);
        if (lane == 0) lds[k * kPcgWaves + wave] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < Count; ++k) {
            double s = lds[k * kPcgWaves];
            for (int w = 1; w < kPcgWaves; ++w) s += lds[k * kPcgWaves + w];
            slots[Count * gs.group() + k] = s;
        }
        grid_arrive_wait(barrier, gs.groups());
    }
    __syncthreads();
}
template <int Count, typename GS>
__device__ inline void fold(const double* slots, double (&out)[Count], GS gs) {
    double acc[Count];
#pragma unroll
    for (int k = 0; k < Count; ++k) acc[k] = 0.0;
    for (unsigned t = threadIdx.x % kWave; t < gs.groups(); t += kWave)
#pragma unroll
        for (int k = 0; k < Count; ++k) acc[k] += slots[Count * t + k];
#pragma unroll
    for (int k = 0; k < Count; ++k) out[k] = wave_sum(acc[k]);
}

// One launch-wide sum with its own synchronisation (used where no vector
// exchange shares the barrier).
template <int Count, bool Grid, typename GS>
__device__ inline void launch_sum(const double (&v)[Count], double* slots, double* lds, unsigned* barrier,
                                  double (&out)[Count], GS gs) {
    if (!Grid) {
        block_sum<Count, kPcgThreads>(v, lds, out);
        return;
    }
    publish_and_wait<Count>(v, lds, slots, barrier, gs);
    fold<Count>(slots, out, gs);
}

// Row `lane` of the node's preconditioner block times the node's 6-vector,
// which is held one row per lane: the six values come from the node's lanes
// by shuffles. Every lane must call it.
template <typename S>
__device__ inline S precondition_row(const S (&mrow)[6], S v, int lane_base) {
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c)
        acc += static_cast<double>(mrow[c]) * static_cast<double>(__shfl(v, lane_base + c, 64));
    return static_cast<S>(acc);
}

// This lane's share of one node's block row, fixed for the whole solve and
// kept in registers: its first block (the row's block `lane`, or -1 past the
// row's end) with that block's column, the column of its second block (block
// first + 8, when the row has one), and the row's end (further blocks of the
// lane are every eighth after the second).
struct RowShare {
    int32_t first, column, second_column, end;
};

__device__ inline RowShare row_share(const int32_t* __restrict__ row_off, const int32_t* __restrict__ cols,
                                     int64_t node, bool valid, int lane) {
    RowShare r{-1, 0, 0, 0};
    if (valid) {
        r.end = row_off[node + 1];
        const int32_t b = row_off[node] + lane;
        if (b < r.end) {
            r.first = b;
            r.column = cols[b];
            if (b + 8 < r.end) r.second_column = cols[b + 8];
        }
    }
    return r;
}

// The matrix a workgroup reads, staged in LDS. A group's nodes are
// consecutive, so their block rows are one contiguous range of the block
// array; the group copies as much of it as the LDS holds once per solve (the
// matrix is fixed for the solve) and reads the rest from memory.
constexpr int kMatrixLdsBytes = 156 * 1024;

template <typename S>
struct MatrixView {
    const S* global;  // all blocks
    const S* local;   // LDS copy of blocks [begin, begin + staged)
    int32_t begin, staged;
};

// One 6x6 block as whole 16-byte loads (blocks are 16-byte aligned in both
// copies), then its product with the column node's 6-vector into acc. An
// FP32 block is nine loads; an FP64 block goes row by row (three loads each)
// to bound the registers it holds.
template <typename S, typename V, typename Acc>
__device__ inline void add_block(const S* blk, const V* __restrict__ x, Acc (&acc)[6]) {
    Acc xv[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) xv[c] = static_cast<Acc>(x[c]);
    if constexpr (sizeof(S) == 4) {
        S v[36];
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const float4 q = reinterpret_cast<const float4*>(blk)[c];
            v[4 * c] = q.x;
            v[4 * c + 1] = q.y;
            v[4 * c + 2] = q.z;
            v[4 * c + 3] = q.w;
        }
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            Acc t = Acc(0);
#pragma unroll
            for (int c = 0; c < 6; ++c) t += static_cast<Acc>(v[6 * r + c]) * xv[c];
            acc[r] += t;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const double2* row = reinterpret_cast<const double2*>(blk + 6 * r);
            const double2 a = row[0], b = row[1], c = row[2];
            const S v[6] = {a.x, a.y, b.x, b.y, c.x, c.y};
            Acc t = Acc(0);
#pragma unroll
            for (int k = 0; k < 6; ++k) t += static_cast<Acc>(v[k]) * xv[k];
            acc[r] += t;
        }
    }
}

template <typename S, typename V, typename Acc>
__device__ inline void add_block_at(const MatrixView<S>& m, int32_t b, const V* __restrict__ x, Acc (&acc)[6]) {
    const int32_t rel = b - m.begin;
    if (rel >= 0 && rel < m.staged)
        add_block<S, V, Acc>(m.local + 36 * rel, x, acc);
    else
        add_block<S, V, Acc>(m.global + 36 * static_cast<int64_t>(b), x, acc);
}

// (A + diag(damp)) u for the six rows of one node, evaluated by the node's
// eight lanes together: lane j multiplies blocks j, j + 8, ... of the node's
// block row (the full 6x6 block times the column node's 6-vector), and the
// eight partial 6-vectors are summed across the lanes in a fixed order. The
// column of the lane's first block (for an FP32 matrix, of its first two)
// comes from registers, so a row of at most eight (sixteen) blocks costs no
// dependent load. `self` is this lane's entry of u
// (the row's own entry, which the lane holds) and `damp` its damping. Every
// lane of the wave must call it (the sum crosses lanes); `valid` false adds
// zeros. Returns row `lane` of the product (lanes 6 and 7 get row 0). Acc is
// the accumulation precision (the matrix precision for the iteration, FP64
// for the refinement residual).
template <typename S, typename V, typename Acc>
__device__ inline Acc node_product(const RowShare& share, const int32_t* __restrict__ cols, const MatrixView<S>& m,
                                   S damp, V self, const V* __restrict__ u, bool valid, int lane) {
    Acc acc[6] = {Acc(0), Acc(0), Acc(0), Acc(0), Acc(0), Acc(0)};
    if (share.first >= 0) {
        add_block_at<S, V, Acc>(m, share.first, u + 6 * static_cast<int64_t>(share.column), acc);
        int32_t b = share.first + 8;
        // The FP64 kernels have no registers to spare for the second column.
        if constexpr (sizeof(S) == 4) {
            if (b < share.end)
                add_block_at<S, V, Acc>(m, b, u + 6 * static_cast<int64_t>(share.second_column), acc);
            b += 8;
        }
        for (; b < share.end; b += 8)
            add_block_at<S, V, Acc>(m, b, u + 6 * static_cast<int64_t>(cols[b]), acc);
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) acc[r] = octet_sum(acc[r]);
    Acc out = acc[0];
#pragma unroll
    for (int r = 1; r < 6; ++r) out = lane == r ? acc[r] : out;
    if (valid && lane < 6) out += static_cast<Acc>(damp) * static_cast<Acc>(self);
    return out;
}

// Row `lane` of (|A| + |D|) |u| for one node: the lane walks every block of
// the node's row itself (one accumulator, no lane exchanges), so this rarely
// needed product holds few registers. Lanes past the node's six rows, and
// rows of absent nodes, give 0.
template <typename S>
__device__ inline double abs_row_product(const int32_t* __restrict__ row_off, const int32_t* __restrict__ cols,
                                         const MatrixView<S>& m, S damp, double self, const double* __restrict__ u,
                                         int64_t node, bool own, int lane) {
    if (!own) return 0.0;
    double acc = fabs(static_cast<double>(damp)) * fabs(self);
    const int32_t end = row_off[node + 1];
    for (int32_t b = row_off[node]; b < end; ++b) {
        const int32_t rel = b - m.begin;
        const S* blk = (rel >= 0 && rel < m.staged) ? m.local + 36 * rel : m.global + 36 * static_cast<int64_t>(b);
        const double* x = u + 6 * static_cast<int64_t>(cols[b]);
#pragma unroll
        for (int c = 0; c < 6; ++c) acc += fabs(static_cast<double>(blk[6 * lane + c])) * fabs(x[c]);
    }
    return acc;
}

// Rounding-error factor of an FP64 residual row b - (A + D) x: a row sums at
// most a few hundred products, each off by at most one unit roundoff, so
// |fl(r) - r| <= gamma (|b| + (|A| + |D|) |x|) componentwise with this gamma
// (n u / (1 - n u) for n = 256, u = 2^-53).
constexpr double kResidualGamma = 256.0 * 1.1102230246251565e-16;

// Preconditioned pipelined CG (Ghysels and Vanroose, 2014): the iteration's
// two inner products and its matrix product are independent, so the
// preconditioned vector m = M w and the partial sums are published together
// and a single grid barrier per iteration serves both. Each inner cycle runs
// in S arithmetic (FP32 or FP64) on the S matrix until the recurrence residual
// falls below max(inner_tol * |r_cycle|, tol * |b|); then x accumulates into
// FP64 and the true residual b - A x is evaluated in FP64. The solve ends when
// that true residual meets tol * |b|; otherwise the next cycle restarts from
// it (iterative refinement in mixed precision, residual replacement in FP64).
//
// Exchange buffers are double-buffered by barrier parity: data written before
// barrier j is read after it, and the same half is next written after barrier
// j + 1, when every block has finished reading it.
//
// The body serves the whole launch (group = blockIdx.x of gridDim.x) or one
// job's groups of a packed launch; `lds` holds 3 x waves doubles and
// `matrix_lds` kMatrixLdsBytes.
constexpr int kPcgLdsDoubles = 3 * (kPcgThreads / kWave);

template <int K, typename S, bool Grid, typename GS>
__device__ __forceinline__ void pcg_body(const int32_t* __restrict__ row_off, const int32_t* __restrict__ cols,
                                const S* __restrict__ bsr, const S* __restrict__ damp, const S* __restrict__ minv,
                                const double* __restrict__ bsr64, const double* __restrict__ damp64,
                                const double* __restrict__ rhs, int64_t nn, double tol, double inner_tol, int max_iter,
                                S* __restrict__ vbuf /* 2 x 6N */, double* __restrict__ xbuf /* 6N */,
                                double* __restrict__ slots /* 2 x 3 x groups */, unsigned* __restrict__ barrier,
                                double* __restrict__ solution, double* __restrict__ outcome, GS gs,
                                double* lds, float4* matrix_lds) {
    const int lane = threadIdx.x & 7;
    const int lane_base = (threadIdx.x & 63) & ~7;
    const bool row_lane = lane < 6;
    const int64_t stride = 6 * nn;
    int64_t node[K];
    bool own[K], live[K];
    RowShare share[K];
    S dampk[K];
    double damp64k[K];  // FP64 damping of the refinement residual (with bsr64)
    S mrow[K][6];
    double b64[K], x64[K];
    S x[K], r[K], u[K], w[K], p[K], s[K], q[K], z[K];
    double bb = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        node[k] = (static_cast<int64_t>(gs.group()) * K + k) * kNodesPerPass + (threadIdx.x >> 3);
        live[k] = node[k] < nn;
        own[k] = row_lane && live[k];
        share[k] = row_share(row_off, cols, node[k], live[k], lane);
        dampk[k] = own[k] ? damp[6 * node[k] + lane] : S(0);
        damp64k[k] = (bsr64 && own[k]) ? damp64[6 * node[k] + lane] : 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) mrow[k][c] = own[k] ? minv[36 * node[k] + 6 * lane + c] : S(0);
        b64[k] = own[k] ? rhs[6 * node[k] + lane] : 0.0;
        x64[k] = 0.0;
        r[k] = static_cast<S>(b64[k]);
        x[k] = u[k] = w[k] = p[k] = s[k] = q[k] = z[k] = S(0);
        bb += b64[k] * b64[k];
    }
    // The group's block rows into LDS, as far as they fit.
    MatrixView<S> mat;
    {
        const int64_t first = static_cast<int64_t>(gs.group()) * K * kNodesPerPass;
        const int64_t last = first + K * kNodesPerPass < nn ? first + K * kNodesPerPass : nn;
        mat.global = bsr;
        mat.local = reinterpret_cast<const S*>(matrix_lds);
        mat.begin = first < nn ? row_off[first] : 0;
        const int32_t blocks = first < nn ? row_off[last] - mat.begin : 0;
        constexpr int32_t capacity = kMatrixLdsBytes / (36 * static_cast<int32_t>(sizeof(S)));
        mat.staged = blocks < capacity ? blocks : capacity;
        const float4* src = reinterpret_cast<const float4*>(bsr + 36 * static_cast<int64_t>(mat.begin));
        const int32_t chunks = mat.staged * 36 * static_cast<int32_t>(sizeof(S)) / 16;
        for (int32_t i = threadIdx.x; i < chunks; i += kPcgThreads) matrix_lds[i] = src[i];
        __syncthreads();
    }
    int phase = 0;
    double red[3];
    {
        const double v[1] = {bb};
        double o[1];
        launch_sum<1, Grid>(v, slots + phase * 3 * gs.groups(), lds, barrier, o, gs);
        phase ^= 1;
        red[0] = o[0];
    }
    const double bnorm = sqrt(red[0]);
    const double target = tol * bnorm;
    double rel = bnorm > 0.0 ? 1.0 : 0.0;
    double cycle_norm = bnorm;
    bool converged = bnorm == 0.0, breakdown = false;
    int it = 0;
    while (!converged && it < max_iter) {
        const double inner_target = fmax(inner_tol * cycle_norm, target);
        // Cycle start: u = M r, w = A u.
        S* ex = vbuf + phase * stride;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const S v = precondition_row(mrow[k], r[k], lane_base);
            u[k] = own[k] ? v : S(0);
            if (own[k]) ex[6 * node[k] + lane] = u[k];
            x[k] = p[k] = s[k] = q[k] = z[k] = S(0);
        }
        sync_all<Grid>(barrier, gs);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const S v = node_product<S, S, S>(share[k], cols, mat, dampk[k], u[k], ex, live[k], lane);
            w[k] = own[k] ? v : S(0);
        }
        phase ^= 1;
        double gamma_prev = 1.0, alpha_prev = 1.0;
        int cycle_it = 0;
        bool inner_done = false;
        while (it < max_iter) {
            S m[K];
            double part[3] = {0.0, 0.0, 0.0};
            ex = vbuf + phase * stride;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const S v = precondition_row(mrow[k], w[k], lane_base);
                m[k] = own[k] ? v : S(0);
                if (own[k]) ex[6 * node[k] + lane] = m[k];
                const double rk = static_cast<double>(r[k]), uk = static_cast<double>(u[k]);
                part[0] += rk * uk;
                part[1] += static_cast<double>(w[k]) * uk;
                part[2] += rk * rk;
            }
            if (Grid) {
                publish_and_wait<3>(part, lds, slots + phase * 3 * gs.groups(), barrier, gs);
                fold<3>(slots + phase * 3 * gs.groups(), red, gs);
            } else {
                block_sum<3, kPcgThreads>(part, lds, red);
            }
            const double gamma = red[0], delta = red[1], rnorm = sqrt(red[2]);
            rel = bnorm > 0.0 ? rnorm / bnorm : 0.0;
            if (rnorm <= inner_target) {
                phase ^= 1;
                inner_done = true;
                break;
            }
            S n[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const S v = node_product<S, S, S>(share[k], cols, mat, dampk[k], m[k], ex, live[k], lane);
                n[k] = own[k] ? v : S(0);
            }
            phase ^= 1;
            double beta, alpha;
            if (cycle_it == 0) {
                beta = 0.0;
                alpha = delta != 0.0 ? gamma / delta : 0.0;
            } else {
                beta = gamma / gamma_prev;
                alpha = gamma / (delta - beta * gamma / alpha_prev);
            }
            if (!(alpha > 0.0) || !isfinite(alpha)) {
                // The recurrences have lost their consistency (the pipelined
                // form's extra recurrences amplify rounding, in FP32 most): end
                // the cycle and restart from the true residual, unless no step
                // was taken from it.
                if (cycle_it == 0) {
                    breakdown = true;
                } else {
                    inner_done = true;
                }
                break;
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                z[k] = static_cast<S>(static_cast<double>(n[k]) + beta * static_cast<double>(z[k]));
                q[k] = static_cast<S>(static_cast<double>(m[k]) + beta * static_cast<double>(q[k]));
                s[k] = static_cast<S>(static_cast<double>(w[k]) + beta * static_cast<double>(s[k]));
                p[k] = static_cast<S>(static_cast<double>(u[k]) + beta * static_cast<double>(p[k]));
                x[k] = static_cast<S>(static_cast<double>(x[k]) + alpha * static_cast<double>(p[k]));
                r[k] = static_cast<S>(static_cast<double>(r[k]) - alpha * static_cast<double>(s[k]));
                u[k] = static_cast<S>(static_cast<double>(u[k]) - alpha * static_cast<double>(q[k]));
                w[k] = static_cast<S>(static_cast<double>(w[k]) - alpha * static_cast<double>(z[k]));
            }
            gamma_prev = gamma;
            alpha_prev = alpha;
            ++it;
            ++cycle_it;
        }
        if (!inner_done) break;
        // True residual of the accumulated FP64 solution.
#pragma unroll
        for (int k = 0; k < K; ++k) {
            x64[k] += static_cast<double>(x[k]);
            x[k] = S(0);
            if (own[k]) xbuf[6 * node[k] + lane] = x64[k];
        }
        sync_all<Grid>(barrier, gs);
        double rr = 0.0;
        double r64[K];
        // The residual of the FP64 system when its matrix is given (the mixed
        // form: FP32 iterations refine the FP64 solution), else of the
        // iteration's own matrix.
        const MatrixView<double> m64{bsr64, nullptr, 0, 0};
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double v =
                bsr64 ? node_product<double, double, double>(share[k], cols, m64, damp64k[k], x64[k], xbuf, live[k],
                                                             lane)
                      : node_product<S, double, double>(share[k], cols, mat, dampk[k], x64[k], xbuf, live[k], lane);
            r64[k] = own[k] ? b64[k] - v : 0.0;
            rr += r64[k] * r64[k];
        }
        {
            const double v[1] = {rr};
            double o[1];
            launch_sum<1, Grid>(v, slots + phase * 3 * gs.groups(), lds, barrier, o, gs);
            phase ^= 1;
            red[0] = o[0];
        }
        const double tnorm = sqrt(red[0]);
        rel = bnorm > 0.0 ? tnorm / bnorm : 0.0;
        if (tnorm <= target) {
            converged = true;
            break;
        }
        // A cycle that no longer reduces the true residual ends the solve. If
        // the residual is then within its own rounding error,
        // gamma (|b| + (|A| + |D|) |x|), the solution is as accurate as FP64
        // evaluates it (a tolerance below that floor cannot be met by any
        // solver); the solve counts as converged, as a direct solve of the same
        // system would. Every group takes the same branch (the sums are the
        // same bits everywhere), so the extra sum's barrier is matched.
        if (__builtin_expect(!(tnorm < cycle_norm), 0)) {
            double ee = 0.0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double a =
                    bsr64 ? abs_row_product<double>(row_off, cols, m64, damp64k[k], x64[k], xbuf, node[k], own[k], lane)
                          : abs_row_product<S>(row_off, cols, mat, dampk[k], x64[k], xbuf, node[k], own[k], lane);
                const double e = own[k] ? fabs(b64[k]) + a : 0.0;
                ee += e * e;
            }
            const double v[1] = {ee};
            double o[1];
            launch_sum<1, Grid>(v, slots + phase * 3 * gs.groups(), lds, barrier, o, gs);
            phase ^= 1;
            converged = tnorm <= kResidualGamma * sqrt(o[0]);
            break;
        }
        cycle_norm = tnorm;
#pragma unroll
        for (int k = 0; k < K; ++k) r[k] = static_cast<S>(r64[k]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (!own[k]) continue;
        solution[6 * node[k] + lane] = x64[k] + static_cast<double>(x[k]);
    }
    if (gs.group() == 0 && threadIdx.x == 0) {
        outcome[0] = it;
        outcome[1] = rel;
        outcome[2] = converged ? 1.0 : 0.0;
        outcome[3] = breakdown ? 1.0 : 0.0;
    }
}

template <int K, typename S, bool Grid>
__global__ void __launch_bounds__(kPcgThreads)
pcg_kernel(const int32_t* __restrict__ row_off, const int32_t* __restrict__ cols, const S* __restrict__ bsr,
           const S* __restrict__ damp, const S* __restrict__ minv, const double* __restrict__ bsr64,
           const double* __restrict__ damp64, const double* __restrict__ rhs, int64_t nn,
           double tol, double inner_tol, int max_iter, S* __restrict__ vbuf /* 2 x 6N */,
           double* __restrict__ xbuf /* 6N */, double* __restrict__ slots /* 2 x 3 x groups */,
           unsigned* __restrict__ barrier, double* __restrict__ solution, double* __restrict__ outcome) {
    __shared__ double lds[kPcgLdsDoubles];
    __shared__ float4 matrix_lds[kMatrixLdsBytes / 16];
    pcg_body<K, S, Grid>(row_off, cols, bsr, damp, minv, bsr64, damp64, rhs, nn, tol, inner_tol, max_iter, vbuf, xbuf,
                         slots, barrier, solution, outcome, LaunchGroups{}, lds, matrix_lds);
}

// ---------------------------------------------------------------- packed ---
// The solves of a batch of jobs (packed.hpp) in one launch. Each job has its
// own preconditioner, damping (its control record's), barrier counter and
// partial-sum slots, and its groups meet only each other. Groups take their
// job and group index by ticket, in the order they start: a job's groups are
// consecutive tickets, so every job whose tickets have all been taken has all
// its groups running and finishes, and the one job still taking tickets
// waits only for groups that start as others finish. With no job holding
// more groups than the device runs at once, every wait ends, whatever the
// launch's total.

template <typename S, bool Mixed>
__global__ void __launch_bounds__(128)
precondition_packed(const PackedJob* __restrict__ jobs, const BatchControl* __restrict__ ctl,
                    const PackedBlock* __restrict__ map, unsigned* __restrict__ ticket) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 0u;
    const PackedBlock m = map[xcd_contiguous(blockIdx.x, gridDim.x)];
    const BatchControl& c = ctl[m.job];
    if (!(c.flags & kJobActive)) return;
    const PackedJob& J = jobs[m.job];
    precondition_node<S>(static_cast<const S*>(J.bsr), Mixed ? J.bsr64 : nullptr, J.diag_position, J.in.node_count,
                         c.damping, c.scaling, static_cast<S*>(J.damp), Mixed ? J.damp64 : nullptr,
                         static_cast<S*>(J.minv), J.barrier, static_cast<int64_t>(m.block) * 128 + threadIdx.x);
}

template <int K, typename S, bool Mixed>
__global__ void __launch_bounds__(kPcgThreads)
pcg_packed(const PackedJob* __restrict__ jobs, const BatchControl* __restrict__ ctl,
           const PackedBlock* __restrict__ tickets, unsigned* __restrict__ ticket) {
    __shared__ double lds[kPcgLdsDoubles];
    __shared__ float4 matrix_lds[kMatrixLdsBytes / 16];
    __shared__ int32_t job_lds;
    if (threadIdx.x == 0) {
        const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const PackedBlock m = tickets[t];
        job_lds = m.job;
        packed_group_set[0] = static_cast<unsigned>(m.block);
        packed_group_set[1] = static_cast<unsigned>(jobs[m.job].pcg_groups);
    }
    __syncthreads();
    const int32_t job = job_lds;
    if (!(ctl[job].flags & kJobActive)) return;
    // The descriptor read through the constant address space: its loads are
    // invariant, so the compiler reloads a pointer where it is used instead
    // of holding every one in registers for the whole solve.
    ConstJob* J = constant_job(jobs, job);
    const S* bsr = static_cast<const S*>(J->bsr);
    const double* bsr64 = Mixed ? J->bsr64 : nullptr;
    const double* damp64 = Mixed ? J->damp64 : nullptr;
    // One form of the body for every job (a job of one group meets itself at
    // its barriers): two inlined forms would not fit the registers.
    pcg_body<K, S, true>(J->row_offsets, J->columns, bsr, static_cast<const S*>(J->damp),
                         static_cast<const S*>(J->minv), bsr64, damp64, J->rhs, J->in.node_count, J->tolerance,
                         J->inner_tolerance, J->max_iterations, static_cast<S*>(J->vbuf), J->xbuf, J->slots,
                         J->barrier, J->delta, J->outcome, PackedGroups{}, lds, matrix_lds);
}

// Node slots per lane group K: at most 2 (register budget: every CG vector
// and the preconditioner rows stay in registers). Systems that fit one workgroup run without grid barriers; larger
// ones spread one slot per lane group over as many groups as needed (at most
// the co-resident count and kPcgThreads, which the fold of the partial sums
// assumes), growing K beyond that.
struct PcgShape {
    int k = 1, groups = 1;
    bool grid = false;
};

template <typename S>
PcgShape pcg_shape(int64_t nn) {
    constexpr int kmax = 2;
    const int64_t passes = (nn + kNodesPerPass - 1) / kNodesPerPass;
    PcgShape sh;
    if (passes <= kmax) {
        while (sh.k < passes) sh.k *= 2;
        return sh;
    }
    sh.grid = true;
    const int limit = std::min(
        persistent_group_budget(reinterpret_cast<const void*>(pcg_kernel<2, S, true>), kPcgThreads), kPcgThreads);
    int64_t need = passes;
    // A grid barrier costs more the more groups arrive at it; past
    // kBarrierGroups groups two node slots per lane group pay off.
    while ((need > limit || need > kBarrierGroups) && sh.k < kmax) {
        sh.k *= 2;
        need = (passes + sh.k - 1) / sh.k;
    }
    if (need > limit) throw InvalidArgument("linear system too large for one persistent launch");
    sh.groups = static_cast<int>(need);
    return sh;
}

// The solve on caller-provided workspace: damp (6N, S), damp64 (6N, only
// with bsr64), minv (36N, S), vbuf (12N, S), xbuf (6N), slots (6 x groups),
// barrier (2), outcome (4). Nothing is allocated here.
template <typename S>
void launch_pcg_raw(const int32_t* row_off, const int32_t* cols, const int32_t* diag_pos, int64_t nn, const S* bsr,
                    const double* bsr64, const double* rhs, double lambda, bool scaling, double tol, double inner_tol,
                    int max_iter, double* solution, const PcgShape& sh, S* damp, double* damp64, S* minv, S* vbuf,
                    double* xbuf, double* slots, unsigned* bar, double* out, hipEvent_t start, hipEvent_t stop,
                    hipStream_t stream) {
    // The preconditioner (and the barrier's reset), then the solve. In a
    // grid launch it is enqueued once the launch is admitted, so the two
    // reach the device back to back.
    struct Prelude {
        const S* bsr;
        const double* bsr64;
        const int32_t* diag_pos;
        int64_t nn;
        double lambda;
        bool scaling;
        S* damp;
        double* damp64;
        S* minv;
        unsigned* bar;
        hipEvent_t start;
        hipStream_t stream;
        static void run(void* p) {
            const Prelude& a = *static_cast<const Prelude*>(p);
            precondition_kernel<S><<<blocks_for(a.nn, 128), 128, 0, a.stream>>>(
                a.bsr, a.bsr64, a.diag_pos, a.nn, a.lambda, a.scaling ? 1 : 0, a.damp, a.damp64, a.minv, a.bar);
            NRICP_LAUNCH_CHECK("precondition_kernel");
            if (a.start) NRICP_CHECK(hipEventRecord(a.start, a.stream));
        }
    } prelude{bsr, bsr64, diag_pos, nn, lambda, scaling, damp, damp64, minv, bar, start, stream};
    const void* kernel = nullptr;
    if (sh.grid) {
        kernel = sh.k == 1 ? reinterpret_cast<const void*>(pcg_kernel<1, S, true>)
                           : reinterpret_cast<const void*>(pcg_kernel<2, S, true>);
    } else {
        kernel = sh.k == 1 ? reinterpret_cast<const void*>(pcg_kernel<1, S, false>)
                           : reinterpret_cast<const void*>(pcg_kernel<2, S, false>);
    }
    int64_t n = nn;
    double* d64 = damp64;
    void* args[] = {const_cast<int32_t**>(&row_off), const_cast<int32_t**>(&cols), const_cast<S**>(&bsr), &damp, &minv,
                    const_cast<double**>(&bsr64), &d64, const_cast<double**>(&rhs), &n, &tol, &inner_tol, &max_iter,
                    &vbuf, &xbuf, &slots, &bar, &solution, &out};
    if (sh.grid) {
        launch_persistent(kernel, sh.groups, kPcgThreads, args, stream, &Prelude::run, &prelude);
    } else {
        Prelude::run(&prelude);
        NRICP_CHECK(hipLaunchKernel(kernel, dim3(1), dim3(kPcgThreads), args, 0, stream));
    }
    if (stop) NRICP_CHECK(hipEventRecord(stop, stream));
}

template <typename S>
void launch_pcg(const LevelLayout& L, const S* bsr, const double* bsr64, const double* rhs, int64_t nn, double lambda,
                bool scaling, double tol, double inner_tol, int max_iter, double* solution, PcgWorkspace& ws, S* damp,
                double* damp64, S* minv, DeviceBuffer<S>& vbuf, hipStream_t stream) {
    if (nn <= 0) return;
    const PcgShape sh = pcg_shape<S>(nn);
    ws.barrier.resize(2);
    ws.reduce.resize(6 * static_cast<int64_t>(sh.groups));
    ws.outcome.resize(4);
    ws.xr.resize(6 * nn);
    vbuf.resize(12 * nn);
    launch_pcg_raw<S>(L.row_offsets.get(), L.columns.get(), L.diag_position.get(), nn, bsr, bsr64, rhs, lambda,
                      scaling, tol, inner_tol, max_iter, solution, sh, damp, damp64, minv, vbuf.get(), ws.xr.get(),
                      ws.reduce.get(), ws.barrier.get(), ws.outcome.get(), ws.kernel_start, ws.kernel_stop, stream);
    ws.last_groups = sh.groups;
}

// ------------------------------------------------------- sweeps (large) ---
// Levels above the persistent launch's node limit: the same block-Jacobi PCG
// (src/cuda/pcg_solver.cu:63-241: x0 = 0, stop when |r|/|b| <= tol on the
// recurrence residual, ok <=> converged) as a sequence of ordinary launches,
// three per iteration, with FP64 vectors and the matrix in its own precision.
// Each launch folds the previous launch's per-block partial sums in one fixed
// order (every block obtains the same bits), so a solve is reproducible and
// needs no grid barrier and no atomics. A solve that has converged or broken
// down turns the launches that remain into no-ops; the host reads the done
// flag every kSweepCheck iterations and stops enqueueing.
constexpr int kSweepThreads = 192;  // 32 nodes: a node's six rows never straddle two blocks
constexpr int kSweepNodes = kSweepThreads / 6;
constexpr int kSweepCheck = 8;
// The state after the partial sums: r.z by iteration parity, |b|, done, rel.
constexpr int kSweepRz = 0, kSweepBnorm = 2, kSweepDone = 3, kSweepRel = 4, kSweepStateDoubles = 8;

template <int C>
__device__ inline void fold_blocks(const double* __restrict__ part, int64_t blocks, double* lds, double (&out)[C]) {
    double v[C];
#pragma unroll
    for (int k = 0; k < C; ++k) v[k] = 0.0;
    for (int64_t b = threadIdx.x; b < blocks; b += kSweepThreads)
#pragma unroll
        for (int k = 0; k < C; ++k) v[k] += part[C * b + k];
    block_sum<C, kSweepThreads>(v, lds, out);
}

__device__ inline void sweep_finish(double* state, double* outcome, int it, double rel, bool converged,
                                    bool breakdown) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        outcome[0] = it;
        outcome[1] = rel;
        outcome[2] = converged ? 1.0 : 0.0;
        outcome[3] = breakdown ? 1.0 : 0.0;
        state[kSweepDone] = 1.0;
    }
}

// r = b, x = 0, z = M^-1 b; partials (r.z, r.r) per block.
template <typename S>
__global__ void __launch_bounds__(kSweepThreads)
sweep_init(const S* __restrict__ minv, const double* __restrict__ rhs, int64_t n6, double* __restrict__ x,
           double* __restrict__ r, double* __restrict__ z, double* __restrict__ part_a, double* __restrict__ state) {
    __shared__ double lds[2 * (kSweepThreads / kWave)];
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kSweepThreads + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (t < n6) {
        const int64_t node = t / 6;
        const int row = static_cast<int>(t % 6);
        const S* m = minv + 36 * node + 6 * row;
        double zt = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) zt += static_cast<double>(m[c]) * rhs[6 * node + c];
        const double b = rhs[t];
        x[t] = 0.0;
        r[t] = b;
        z[t] = zt;
        v[0] = b * zt;
        v[1] = b * b;
    }
    double s[2];
    block_sum<2, kSweepThreads>(v, lds, s);
    if (threadIdx.x == 0) {
        part_a[2 * blockIdx.x] = s[0];
        part_a[2 * blockIdx.x + 1] = s[1];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) state[kSweepDone] = 0.0;
}

// Iteration `it` begins: (r.z, r.r) of the last update, the stopping tests,
// then the search direction p = z + beta p (p = z at it = 0). With `last`
// (it = the iteration cap) only the tests run.
__global__ void __launch_bounds__(kSweepThreads)
sweep_direction(int it, int last, const double* __restrict__ part_a, int64_t blocks, double tol, int64_t n6,
                const double* __restrict__ z, double* __restrict__ p, double* __restrict__ state,
                double* __restrict__ outcome) {
    __shared__ double lds[2 * (kSweepThreads / kWave)];
    if (state[kSweepDone] != 0.0) return;
    double s[2];
    fold_blocks<2>(part_a, blocks, lds, s);
    const double rz = s[0], rr = s[1];
    const double bnorm = it == 0 ? sqrt(rr) : state[kSweepBnorm];
    const double rnorm = sqrt(rr);
    const double rel = bnorm > 0.0 ? rnorm / bnorm : 0.0;
    if (rnorm <= tol * bnorm) {
        sweep_finish(state, outcome, it, rel, true, false);
        return;
    }
    if (last) {
        sweep_finish(state, outcome, it, rel, false, false);
        return;
    }
    if (!(rz > 0.0) || !isfinite(rz)) {
        sweep_finish(state, outcome, it, rel, false, true);
        return;
    }
    const double beta = it == 0 ? 0.0 : rz / state[kSweepRz + ((it - 1) & 1)];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        state[kSweepRz + (it & 1)] = rz;
        state[kSweepRel] = rel;
        if (it == 0) state[kSweepBnorm] = bnorm;
    }
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kSweepThreads + threadIdx.x;
    if (t < n6) p[t] = it == 0 ? z[t] : z[t] + beta * p[t];
}

// q = (A + D) p over the block rows; partials p.q per block.
template <typename S>
__global__ void __launch_bounds__(kSweepThreads)
sweep_product(const int32_t* __restrict__ row_off, const int32_t* __restrict__ cols, const S* __restrict__ bsr,
              const S* __restrict__ damp, int64_t n6, const double* __restrict__ p, double* __restrict__ q,
              double* __restrict__ part_b, const double* __restrict__ state) {
    __shared__ double lds[kSweepThreads / kWave];
    if (state[kSweepDone] != 0.0) return;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kSweepThreads + threadIdx.x;
    double v[1] = {0.0};
    if (t < n6) {
        const int64_t node = t / 6;
        const int row = static_cast<int>(t % 6);
        double acc = static_cast<double>(damp[t]) * p[t];
        for (int32_t b = row_off[node]; b < row_off[node + 1]; ++b) {
            const S* blk = bsr + 36 * static_cast<int64_t>(b) + 6 * row;
            const double* pc = p + 6 * static_cast<int64_t>(cols[b]);
#pragma unroll
            for (int c = 0; c < 6; ++c) acc += static_cast<double>(blk[c]) * pc[c];
        }
        q[t] = acc;
        v[0] = p[t] * acc;
    }
    double s[1];
    block_sum<1, kSweepThreads>(v, lds, s);
    if (threadIdx.x == 0) part_b[blockIdx.x] = s[0];
}

// alpha = r.z / p.q; x += alpha p, r -= alpha q, z = M^-1 r; partials
// (r.z, r.r) per block.
template <typename S>
__global__ void __launch_bounds__(kSweepThreads)
sweep_update(int it, const double* __restrict__ part_b, int64_t blocks, const S* __restrict__ minv, int64_t n6,
             double* __restrict__ x, double* __restrict__ r, double* __restrict__ z, const double* __restrict__ p,
             const double* __restrict__ q, double* __restrict__ part_a, double* __restrict__ state,
             double* __restrict__ outcome) {
    __shared__ double lds[2 * (kSweepThreads / kWave)];
    __shared__ double rows[kSweepThreads];
    if (state[kSweepDone] != 0.0) return;
    double s1[1];
    fold_blocks<1>(part_b, blocks, lds, s1);
    const double pq = s1[0];
    if (!(pq > 0.0) || !isfinite(pq)) {
        sweep_finish(state, outcome, it, state[kSweepRel], false, true);
        return;
    }
    const double alpha = state[kSweepRz + (it & 1)] / pq;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kSweepThreads + threadIdx.x;
    double rt = 0.0;
    if (t < n6) {
        x[t] += alpha * p[t];
        rt = r[t] - alpha * q[t];
        r[t] = rt;
    }
    rows[threadIdx.x] = rt;
    __syncthreads();
    double v[2] = {0.0, 0.0};
    if (t < n6) {
        const int64_t node = t / 6;
        const int row = static_cast<int>(t % 6);
        const S* m = minv + 36 * node + 6 * row;
        const double* rn = rows + 6 * (threadIdx.x / 6);
        double zt = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) zt += static_cast<double>(m[c]) * rn[c];
        z[t] = zt;
        v[0] = rt * zt;
        v[1] = rt * rt;
    }
    double s[2];
    block_sum<2, kSweepThreads>(v, lds, s);
    if (threadIdx.x == 0) {
        part_a[2 * blockIdx.x] = s[0];
        part_a[2 * blockIdx.x + 1] = s[1];
    }
}

// The multi-launch solve on the session's workspace: damp (6N, S) and minv
// (36N, S) come in; ws.r, ws.u, ws.p, ws.w (6N each) hold r, z, p and q.
template <typename S>
void pcg_sweeps(const LevelLayout& L, const S* bsr, const double* rhs, int64_t nn, double lambda, bool scaling,
                double tol, int max_iter, double* solution, PcgWorkspace& ws, S* damp, S* minv, hipStream_t stream) {
    const int64_t n6 = 6 * nn;
    const int64_t blocks = (nn + kSweepNodes - 1) / kSweepNodes;
    if (blocks > static_cast<int64_t>(INT32_MAX)) throw InvalidArgument("linear system too large");
    ws.barrier.resize(2);
    ws.outcome.resize(4);
    ws.reduce.resize(3 * blocks + kSweepStateDoubles);
    ws.r.resize(n6);
    ws.u.resize(n6);
    ws.p.resize(n6);
    ws.w.resize(n6);
    double* part_a = ws.reduce.get();
    double* part_b = part_a + 2 * blocks;
    double* state = part_b + blocks;
    const unsigned grid = static_cast<unsigned>(blocks);
    precondition_kernel<S><<<blocks_for(nn, 128), 128, 0, stream>>>(bsr, nullptr, L.diag_position.get(), nn, lambda,
                                                                     scaling ? 1 : 0, damp, nullptr, minv,
                                                                     ws.barrier.get());
    NRICP_LAUNCH_CHECK("precondition_kernel");
    if (ws.kernel_start) NRICP_CHECK(hipEventRecord(ws.kernel_start, stream));
    sweep_init<S><<<grid, kSweepThreads, 0, stream>>>(minv, rhs, n6, solution, ws.r.get(), ws.u.get(), part_a, state);
    NRICP_LAUNCH_CHECK("sweep_init");
    const int cap = std::max(0, max_iter);
    for (int it = 0;; ++it) {
        const int last = it == cap ? 1 : 0;
        sweep_direction<<<grid, kSweepThreads, 0, stream>>>(it, last, part_a, blocks, tol, n6, ws.u.get(),
                                                            ws.p.get(), state, ws.outcome.get());
        NRICP_LAUNCH_CHECK("sweep_direction");
        if (last) break;
        sweep_product<S><<<grid, kSweepThreads, 0, stream>>>(L.row_offsets.get(), L.columns.get(), bsr, damp, n6,
                                                             ws.p.get(), ws.w.get(), part_b, state);
        NRICP_LAUNCH_CHECK("sweep_product");
        sweep_update<S><<<grid, kSweepThreads, 0, stream>>>(it, part_b, blocks, minv, n6, solution, ws.r.get(),
                                                            ws.u.get(), ws.p.get(), ws.w.get(), part_a, state,
                                                            ws.outcome.get());
        NRICP_LAUNCH_CHECK("sweep_update");
        if ((it + 1) % kSweepCheck == 0) {
            double done = 0.0;
            NRICP_CHECK(hipMemcpyAsync(&done, state + kSweepDone, sizeof(double), hipMemcpyDeviceToHost, stream));
            NRICP_CHECK(hipStreamSynchronize(stream));
            if (done != 0.0) break;
        }
    }
    if (ws.kernel_stop) NRICP_CHECK(hipEventRecord(ws.kernel_stop, stream));
    ws.last_groups = 0;
}

}  // namespace

void pcg_solve(const LevelLayout& L, const double* bsr, const double* rhs, int64_t nn, double lambda, bool scaling,
               double tol, int max_iter, double* solution, PcgWorkspace& ws, hipStream_t stream) {
    ws.damp.resize(6 * nn);
    ws.minv.resize(36 * nn);
    if (nn > pcg_max_nodes()) {
        pcg_sweeps<double>(L, bsr, rhs, nn, lambda, scaling, tol, max_iter, solution, ws, ws.damp.get(),
                           ws.minv.get(), stream);
        return;
    }
    launch_pcg<double>(L, bsr, nullptr, rhs, nn, lambda, scaling, tol, tol, max_iter, solution, ws, ws.damp.get(),
                       nullptr, ws.minv.get(), ws.u, stream);
}

int64_t packed_pcg_groups(int64_t nodes, int k) {
    const int64_t per_group = static_cast<int64_t>(k) * kNodesPerPass;
    return std::max<int64_t>(1, (nodes + per_group - 1) / per_group);
}

void packed_precondition(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, int precision,
                         unsigned* ticket, KernelSink& sink) {
    if (maps.node_blocks <= 0) return;
    const unsigned blocks = static_cast<unsigned>(maps.node_blocks);
    if (precision == NRICP_HIP_PRECISION_DOUBLE)
        sink_launch(sink, precondition_packed<double, false>, blocks, 128, jobs, ctl, maps.node, ticket);
    else if (precision == NRICP_HIP_PRECISION_MIXED)
        sink_launch(sink, precondition_packed<float, true>, blocks, 128, jobs, ctl, maps.node, ticket);
    else
        sink_launch(sink, precondition_packed<float, false>, blocks, 128, jobs, ctl, maps.node, ticket);
}

void packed_pcg(const PackedJob* jobs, const BatchControl* ctl, const PackedMaps& maps, int precision, int k,
                unsigned* ticket, KernelSink& sink) {
    if (maps.pcg_groups <= 0) return;
    const unsigned groups = static_cast<unsigned>(maps.pcg_groups);
    auto go = [&](auto kernel) { sink_launch(sink, kernel, groups, kPcgThreads, jobs, ctl, maps.pcg, ticket); };
    if (precision == NRICP_HIP_PRECISION_DOUBLE) {
        if (k == 1) go(pcg_packed<1, double, false>); else go(pcg_packed<2, double, false>);
    } else if (precision == NRICP_HIP_PRECISION_MIXED) {
        if (k == 1) go(pcg_packed<1, float, true>); else go(pcg_packed<2, float, true>);
    } else {
        if (k == 1) go(pcg_packed<1, float, false>); else go(pcg_packed<2, float, false>);
    }
}

// The most groups one job may hold in a packed launch (the groups the device
// keeps resident at once; packed_pcg's tickets rely on it).
int64_t packed_pcg_group_limit() {
    return std::min(persistent_group_budget(reinterpret_cast<const void*>(pcg_packed<2, double, false>), kPcgThreads),
                    kPcgThreads);
}

int64_t pcg_max_nodes() {
    // Both kernel precisions share the launch bound: two node slots per lane
    // group, at most the co-resident group count and kPcgThreads groups.
    const int limit = std::min(
        persistent_group_budget(reinterpret_cast<const void*>(pcg_kernel<2, float, true>), kPcgThreads),
        persistent_group_budget(reinterpret_cast<const void*>(pcg_kernel<2, double, true>), kPcgThreads));
    return static_cast<int64_t>(2) * std::min(limit, kPcgThreads) * kNodesPerPass;
}

int64_t pcg_workspace_groups(int64_t nn, bool fp32) {
    if (nn <= 0) return 0;
    return fp32 ? pcg_shape<float>(nn).groups : pcg_shape<double>(nn).groups;
}

void pcg_solve_on(const int32_t* row_off, const int32_t* cols, const int32_t* diag_pos, int64_t nn, const double* bsr,
                  const float* bsr32, const double* rhs, double lambda, bool scaling, double tol, int max_iter,
                  double* solution, const PcgWorkspaceView& w, hipStream_t stream) {
    if (nn <= 0) return;
    if (bsr32) {
        const PcgShape sh = pcg_shape<float>(nn);
        launch_pcg_raw<float>(row_off, cols, diag_pos, nn, bsr32, bsr, rhs, lambda, scaling, tol, std::max(tol, 1e-4),
                              max_iter, solution, sh, static_cast<float*>(w.damp), w.damp64,
                              static_cast<float*>(w.minv), static_cast<float*>(w.vectors), w.xbuf, w.slots, w.barrier,
                              w.outcome, nullptr, nullptr, stream);
    } else {
        const PcgShape sh = pcg_shape<double>(nn);
        launch_pcg_raw<double>(row_off, cols, diag_pos, nn, bsr, nullptr, rhs, lambda, scaling, tol, tol, max_iter,
                               solution, sh, static_cast<double*>(w.damp), nullptr, static_cast<double*>(w.minv),
                               static_cast<double*>(w.vectors), w.xbuf, w.slots, w.barrier, w.outcome, nullptr,
                               nullptr, stream);
    }
}

void pcg_solve_mixed(const LevelLayout& L, const float* bsr, const double* bsr64, const double* rhs, int64_t nn,
                     double lambda, bool scaling, double tol, int max_iter, double* solution, PcgWorkspace& ws,
                     hipStream_t stream) {
    if (nn > pcg_max_nodes()) {
        // Above one persistent launch: the multi-launch solve in FP64
        // arithmetic, on the FP64 blocks when the form has them (the mixed
        // form's target system), else on the FP32 ones.
        if (bsr64) {
            ws.damp.resize(6 * nn);
            ws.minv.resize(36 * nn);
            pcg_sweeps<double>(L, bsr64, rhs, nn, lambda, scaling, tol, max_iter, solution, ws, ws.damp.get(),
                               ws.minv.get(), stream);
        } else {
            ws.w32.resize(6 * nn);
            ws.minv32.resize(36 * nn);
            pcg_sweeps<float>(L, bsr, rhs, nn, lambda, scaling, tol, max_iter, solution, ws, ws.w32.get(),
                              ws.minv32.get(), stream);
        }
        return;
    }
    ws.w32.resize(6 * nn);  // damping vector (FP32)
    ws.minv32.resize(36 * nn);
    if (bsr64) ws.damp.resize(6 * nn);  // FP64 damping of the refinement residual
    // FP32 cycles stop at 1e-4 of the cycle's starting residual: the gap
    // between the pipelined recurrence residual and the true one grows with
    // the iterations of a cycle and reaches 1e-3 within a few dozen FP32
    // iterations on ill-conditioned systems, so the cycles restart (residual
    // replacement) well before it matters; FP64 refinement does the rest.
    launch_pcg<float>(L, bsr, bsr64, rhs, nn, lambda, scaling, tol, std::max(tol, 1e-4), max_iter, solution, ws,
                      ws.w32.get(), bsr64 ? ws.damp.get() : nullptr, ws.minv32.get(), ws.u32, stream);
}

}  // namespace nricp_hip
