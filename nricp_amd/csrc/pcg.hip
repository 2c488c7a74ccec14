// Block-Jacobi preconditioned conjugate gradients on the damped block system,
// as one persistent cooperative launch per solve.
//
// Reference behaviour: replaces src/cuda/pcg_solver.cu:15-326 (block-Jacobi
// PCG from x0 = 0, stop when |r|/|b| <= tol, ok <=> converged). Differences,
// chosen for gfx950:
//   * Chronopoulos-Gear recurrence: both inner products of an iteration are
//     reduced together, so an iteration costs two grid barriers, not four;
//   * the matrix is block-sparse rows of the full symmetric system, so the
//     product is a gather with no atomics (one lane per block row of a node,
//     eight lanes per node, six active);
//   * each lane keeps its rows of x, r, p, s in registers for the whole solve;
//     only the preconditioned residual u crosses lanes (through memory, for
//     the neighbours' products);
//   * damping is the CPU route's (src/solver/normal_equations.cpp:81-96):
//     lambda * max(A_ii, 1e-12) with diagonal scaling, lambda otherwise;
//   * inner products and the refinement residual are accumulated in FP64 and
//     reduced in a fixed order, so a solve is reproducible bit for bit;
//   * mixed mode runs the iteration in FP32 on an FP32 matrix and refines in
//     FP64 inside the same launch: x += e, r = b - A x evaluated in FP64
//     arithmetic, restart, until the FP64 relative residual reaches the
//     tolerance.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace nricp_hip {
namespace {

constexpr int kPcgThreads = 512;
constexpr int kNodesPerPass = kPcgThreads / 8;  // 8 lanes per node, 6 active

// Damped diagonal blocks inverted once per solve (Gauss-Jordan with partial
// pivoting; a singular block falls back to the identity as the reference's
// factor kernel does).
template <typename S>
__global__ void precondition_kernel(const S* __restrict__ bsr, const int32_t* __restrict__ diag_pos, int64_t nn,
                                    double lambda, int scaling, S* __restrict__ damp, S* __restrict__ minv) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const S* blk = bsr + 36 * static_cast<int64_t>(diag_pos[i]);
    double a[6][12];
    for (int r = 0; r < 6; ++r) {
        for (int c = 0; c < 6; ++c) {
            a[r][c] = static_cast<double>(blk[6 * r + c]);
            a[r][6 + c] = r == c ? 1.0 : 0.0;
        }
        double d = 0.0;
        if (lambda > 0.0) d = lambda * (scaling ? fmax(a[r][r], 1e-12) : 1.0);
        damp[6 * i + r] = static_cast<S>(d);
        a[r][r] += d;
    }
    bool ok = true;
    for (int c = 0; c < 6 && ok; ++c) {
        int piv = c;
        for (int r = c + 1; r < 6; ++r)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (!(fabs(a[piv][c]) > 0.0)) {
            ok = false;
            break;
        }
        if (piv != c)
            for (int k = 0; k < 12; ++k) {
                const double t = a[c][k];
                a[c][k] = a[piv][k];
                a[piv][k] = t;
            }
        const double inv = 1.0 / a[c][c];
        for (int k = 0; k < 12; ++k) a[c][k] *= inv;
        for (int r = 0; r < 6; ++r) {
            if (r == c) continue;
            const double f = a[r][c];
            if (f == 0.0) continue;
            for (int k = 0; k < 12; ++k) a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c)
            minv[36 * i + 6 * r + c] = static_cast<S>(ok ? a[r][6 + c] : (r == c ? 1.0 : 0.0));
}

template <bool Grid>
__device__ inline void sync_all(unsigned* barrier) {
    if (Grid)
        grid_barrier(barrier, gridDim.x);
    else
        __syncthreads();
}

// Sum of 3 FP64 partials over the block, then over all blocks in index order.
template <bool Grid>
__device__ inline void reduce3(double a, double b, double c, double* slots, double* lds, unsigned* barrier,
                               double (&out)[3]) {
    double v[3] = {a, b, c}, s[3];
    block_sum<3, kPcgThreads>(v, lds, s);
    if (!Grid) {
        out[0] = s[0];
        out[1] = s[1];
        out[2] = s[2];
        return;
    }
    if (threadIdx.x == 0) {
        slots[3 * blockIdx.x] = s[0];
        slots[3 * blockIdx.x + 1] = s[1];
        slots[3 * blockIdx.x + 2] = s[2];
    }
    grid_barrier(barrier, gridDim.x);
    out[0] = out[1] = out[2] = 0.0;
    for (unsigned g = 0; g < gridDim.x; ++g) {
        out[0] += slots[3 * g];
        out[1] += slots[3 * g + 1];
        out[2] += slots[3 * g + 2];
    }
}

// One row of (A + diag(damp)) u for a node: the node's blocks are
// contiguous, and the six lanes of the node read one 36-entry block together.
// Acc is the accumulation precision (the matrix precision for the iteration,
// FP64 for the refinement residual).
template <typename S, typename V, typename Acc>
__device__ inline Acc row_product(const int32_t* __restrict__ row_off, const int32_t* __restrict__ cols,
                                  const S* __restrict__ bsr, const S* __restrict__ damp, const V* __restrict__ u,
                                  int64_t node, int row) {
    Acc acc = Acc(0);
    for (int32_t b = row_off[node]; b < row_off[node + 1]; ++b) {
        const S* blk = bsr + 36 * static_cast<int64_t>(b) + 6 * row;
        const V* x = u + 6 * static_cast<int64_t>(cols[b]);
        Acc t = Acc(0);
#pragma unroll
        for (int c = 0; c < 6; ++c) t += static_cast<Acc>(blk[c]) * static_cast<Acc>(x[c]);
        acc += t;
    }
    return acc + static_cast<Acc>(damp[6 * node + row]) * static_cast<Acc>(u[6 * node + row]);
}

// S: matrix / iteration precision. Refine: FP64 outer refinement (S = float).
template <int K, typename S, bool Refine, bool Grid>
__global__ void __launch_bounds__(kPcgThreads)
pcg_kernel(const int32_t* __restrict__ row_off, const int32_t* __restrict__ cols, const S* __restrict__ bsr,
           const S* __restrict__ damp, const S* __restrict__ minv, const double* __restrict__ rhs, int64_t nn,
           double tol, double inner_tol, int max_iter, S* __restrict__ ubuf, double* __restrict__ xbuf,
           double* __restrict__ slots, unsigned* __restrict__ barrier, double* __restrict__ solution,
           double* __restrict__ outcome) {
    __shared__ double lds[3 * (kPcgThreads / kWave)];
    const int lane = threadIdx.x & 7;
    const int lane_base = (threadIdx.x & 63) & ~7;
    const bool row_lane = lane < 6;
    int64_t node[K];
    bool own[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        node[k] = (static_cast<int64_t>(blockIdx.x) * K + k) * kNodesPerPass + (threadIdx.x >> 3);
        own[k] = row_lane && node[k] < nn;
    }
    double b64[K], x64[K];
    S x[K], r[K], p[K], s[K], u[K], w[K];
    double bb_local = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        b64[k] = own[k] ? rhs[6 * node[k] + lane] : 0.0;
        x64[k] = 0.0;
        x[k] = p[k] = s[k] = u[k] = w[k] = S(0);
        r[k] = static_cast<S>(b64[k]);
        bb_local += b64[k] * b64[k];
    }
    double red[3];
    double* slot_a = slots;
    double* slot_b = slots + 3 * gridDim.x;
    reduce3<Grid>(bb_local, 0.0, 0.0, slot_a, lds, barrier, red);
    const double bnorm = sqrt(red[0]);
    double target = tol * bnorm;
    double inner_target = Refine ? inner_tol * bnorm : target;
    double gamma_prev = 1.0, alpha_prev = 1.0;
    int it = 0, cycle_it = 0;
    bool converged = bnorm == 0.0, breakdown = false;
    double rel = bnorm == 0.0 ? 0.0 : 1.0;
    int parity = 1;
    while (!converged && it < max_iter) {
        // u = M^-1 r (rows of the node exchanged through lane shuffles).
        double gamma_l = 0.0, rho_l = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            // Preconditioner row read per iteration (it stays in cache); the
            // node's residual rows come from the neighbouring lanes.
            const S* mrow = minv + 36 * (own[k] ? node[k] : 0) + 6 * (row_lane ? lane : 0);
            double uu = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const double rc = static_cast<double>(__shfl(r[k], lane_base + c, 64));
                uu += static_cast<double>(mrow[c]) * rc;
            }
            u[k] = own[k] ? static_cast<S>(uu) : S(0);
            if (own[k]) ubuf[6 * node[k] + lane] = u[k];
            gamma_l += static_cast<double>(r[k]) * static_cast<double>(u[k]);
            rho_l += static_cast<double>(r[k]) * static_cast<double>(r[k]);
        }
        sync_all<Grid>(barrier);
        double delta_l = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (own[k]) {
                w[k] = row_product<S, S, S>(row_off, cols, bsr, damp, ubuf, node[k], lane);
                delta_l += static_cast<double>(w[k]) * static_cast<double>(u[k]);
            }
        }
        double* sl = parity ? slot_b : slot_a;
        parity ^= 1;
        reduce3<Grid>(gamma_l, rho_l, delta_l, sl, lds, barrier, red);
        const double gamma = red[0], rho = red[1], delta = red[2];
        const double rnorm = sqrt(rho);
        if (!Refine) rel = bnorm > 0.0 ? rnorm / bnorm : 0.0;
        if (rnorm <= inner_target) {
            if (!Refine) {
                converged = true;
                break;
            }
            // Refinement: x64 += x, r64 = b - A x64 in FP64 arithmetic.
#pragma unroll
            for (int k = 0; k < K; ++k) {
                x64[k] += static_cast<double>(x[k]);
                if (own[k]) xbuf[6 * node[k] + lane] = x64[k];
            }
            sync_all<Grid>(barrier);
            double rr_l = 0.0;
            double r64[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                r64[k] = own[k] ? b64[k] - row_product<S, double, double>(row_off, cols, bsr, damp, xbuf, node[k], lane) : 0.0;
                rr_l += r64[k] * r64[k];
            }
            double* sl2 = parity ? slot_b : slot_a;
            parity ^= 1;
            reduce3<Grid>(rr_l, 0.0, 0.0, sl2, lds, barrier, red);
            const double r64norm = sqrt(red[0]);
            rel = bnorm > 0.0 ? r64norm / bnorm : 0.0;
            if (r64norm <= target) {
                converged = true;
                break;
            }
            // Restart the inner iteration on the FP64 residual.
            inner_target = inner_tol * r64norm;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                r[k] = static_cast<S>(r64[k]);
                x[k] = p[k] = s[k] = S(0);
            }
            cycle_it = 0;
            ++it;
            continue;
        }
        double beta, alpha;
        if (cycle_it == 0) {
            beta = 0.0;
            alpha = delta != 0.0 ? gamma / delta : 0.0;
        } else {
            beta = gamma / gamma_prev;
            alpha = gamma / (delta - beta * gamma / alpha_prev);
        }
        if (!(alpha > 0.0) || !isfinite(alpha)) {
            breakdown = true;
            break;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            p[k] = static_cast<S>(static_cast<double>(u[k]) + beta * static_cast<double>(p[k]));
            s[k] = static_cast<S>(static_cast<double>(w[k]) + beta * static_cast<double>(s[k]));
            x[k] = static_cast<S>(static_cast<double>(x[k]) + alpha * static_cast<double>(p[k]));
            r[k] = static_cast<S>(static_cast<double>(r[k]) - alpha * static_cast<double>(s[k]));
        }
        gamma_prev = gamma;
        alpha_prev = alpha;
        ++it;
        ++cycle_it;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (!own[k]) continue;
        solution[6 * node[k] + lane] = Refine ? x64[k] + (converged ? 0.0 : static_cast<double>(x[k]))
                                              : static_cast<double>(x[k]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        outcome[0] = it;
        outcome[1] = rel;
        outcome[2] = converged ? 1.0 : 0.0;
        outcome[3] = breakdown ? 1.0 : 0.0;
    }
}

struct Launch {
    int k;
    int groups;
};

int resident_groups(const void* kernel) {
    int device = 0;
    NRICP_CHECK(hipGetDevice(&device));
    int cus = 0;
    NRICP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    int per_cu = 0;
    NRICP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kPcgThreads, 0));
    return cus * std::max(per_cu, 1);
}

template <typename S, bool Refine>
void launch_pcg(const LevelLayout& L, const S* bsr, const double* rhs, int64_t nn, double lambda, bool scaling,
                double tol, int max_iter, double* solution, PcgWorkspace& ws, S* damp, S* minv, S* ubuf,
                hipStream_t stream) {
    if (nn <= 0) return;
    precondition_kernel<S><<<blocks_for(nn, 128), 128, 0, stream>>>(bsr, L.diag_position.get(), nn, lambda,
                                                                     scaling ? 1 : 0, damp, minv);
    NRICP_LAUNCH_CHECK("precondition_kernel");
    // Node slots per group: 64 per K. Small systems run in one workgroup
    // (no grid barrier); larger ones spread one K=1 slot per lane group
    // over as many groups as are co-resident, growing K beyond that.
    const int64_t passes = (nn + kNodesPerPass - 1) / kNodesPerPass;
    int k = 1;
    bool grid = true;
    if (passes <= 4) {
        k = passes <= 1 ? 1 : (passes <= 2 ? 2 : 4);
        grid = false;
    }
    const void* kernel = nullptr;
#define NRICP_PCG_PICK(KK, GG) reinterpret_cast<const void*>(pcg_kernel<KK, S, Refine, GG>)
    int groups = 1;
    if (grid) {
        const int resident = resident_groups(NRICP_PCG_PICK(1, true));
        int64_t need = passes;
        while (need > resident && k < 4) {
            k *= 2;
            need = (passes + k - 1) / k;
        }
        if (need > resident) throw InvalidArgument("linear system too large for one persistent launch");
        groups = static_cast<int>(need);
        if (groups == 1) grid = false;
    }
    switch (k * 2 + (grid ? 1 : 0)) {
        case 2: kernel = NRICP_PCG_PICK(1, false); break;
        case 3: kernel = NRICP_PCG_PICK(1, true); break;
        case 4: kernel = NRICP_PCG_PICK(2, false); break;
        case 5: kernel = NRICP_PCG_PICK(2, true); break;
        case 8: kernel = NRICP_PCG_PICK(4, false); break;
        default: kernel = NRICP_PCG_PICK(4, true); break;
    }
#undef NRICP_PCG_PICK
    ws.reduce.resize(6 * static_cast<int64_t>(groups) + 6);
    ws.barrier.resize(2);
    ws.barrier.zero(stream);
    ws.outcome.resize(4);
    ws.xr.resize(6 * nn);
    const int32_t* row_off = L.row_offsets.get();
    const int32_t* cols = L.columns.get();
    double inner_tol = Refine ? std::max(tol, 1e-5) : tol;
    int64_t n = nn;
    double* xbuf = ws.xr.get();
    double* slots = ws.reduce.get();
    unsigned* bar = ws.barrier.get();
    double* out = ws.outcome.get();
    void* args[] = {&row_off, &cols, const_cast<S**>(&bsr), &damp, &minv, const_cast<double**>(&rhs), &n, &tol,
                    &inner_tol, &max_iter, &ubuf, &xbuf, &slots, &bar, &solution, &out};
    if (ws.kernel_start) NRICP_CHECK(hipEventRecord(ws.kernel_start, stream));
    if (grid) {
        NRICP_CHECK(hipLaunchCooperativeKernel(kernel, dim3(groups), dim3(kPcgThreads), args, 0, stream));
    } else {
        NRICP_CHECK(hipLaunchKernel(kernel, dim3(1), dim3(kPcgThreads), args, 0, stream));
    }
    if (ws.kernel_stop) NRICP_CHECK(hipEventRecord(ws.kernel_stop, stream));
    ws.last_groups = groups;
}

}  // namespace

void pcg_solve(const LevelLayout& L, const double* bsr, const double* rhs, int64_t nn, double lambda, bool scaling,
               double tol, int max_iter, double* solution, PcgWorkspace& ws, hipStream_t stream) {
    ws.damp.resize(6 * nn);
    ws.minv.resize(36 * nn);
    ws.u.resize(6 * nn);
    launch_pcg<double, false>(L, bsr, rhs, nn, lambda, scaling, tol, max_iter, solution, ws, ws.damp.get(),
                              ws.minv.get(), ws.u.get(), stream);
}

void pcg_solve_mixed(const LevelLayout& L, const float* bsr, const double* rhs, int64_t nn, double lambda,
                     bool scaling, double tol, int max_iter, double* solution, PcgWorkspace& ws, hipStream_t stream) {
    ws.u32.resize(6 * nn);
    ws.w32.resize(6 * nn);  // damping vector (FP32)
    ws.minv32.resize(36 * nn);
    launch_pcg<float, true>(L, bsr, rhs, nn, lambda, scaling, tol, max_iter, solution, ws, ws.w32.get(),
                            ws.minv32.get(), ws.u32.get(), stream);
}

}  // namespace nricp_hip
