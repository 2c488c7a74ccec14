// The local/global route's device steps: the per-node rotation fit (local
// step) and the restriction of the assembled system to its translation corner
// (global step; its right-hand side comes from translation_rhs).
//
// Reference behaviour: src/cuda/local_global.cu -- fit_rotations_kernel
// (:188-302), horn_quaternion (:53-129), rotation_from_quaternion (:134-157),
// the restriction kernels (:453-485) and the 0.2 under-relaxation (:18). The
// arithmetic here is FP64 and the correlation is reduced in a fixed order, so
// the FP64 CPU restatement checks it to rounding.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "so3.hpp"

namespace nricp_hip {
namespace {

constexpr int kThreads = 256;

// S += w x y^T, row major.
__device__ inline void add_correlation(double* S, double w, d3 x, d3 y) {
    const double xv[3] = {x.x, x.y, x.z}, yv[3] = {y.x, y.y, y.z};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) S[3 * r + c] += w * xv[r] * yv[c];
}

// Eigenvector of the largest eigenvalue of Horn's symmetric 4x4, by cyclic
// Jacobi rotations (at most 12 sweeps, fixed order); the zero correlation
// gives the identity quaternion.
__device__ inline void horn_quaternion(const double* S, double q[4]) {
    const double sxx = S[0], sxy = S[1], sxz = S[2];
    const double syx = S[3], syy = S[4], syz = S[5];
    const double szx = S[6], szy = S[7], szz = S[8];
    double a[4][4];
    a[0][0] = sxx + syy + szz;
    a[0][1] = syz - szy;
    a[0][2] = szx - sxz;
    a[0][3] = sxy - syx;
    a[1][1] = sxx - syy - szz;
    a[1][2] = sxy + syx;
    a[1][3] = szx + sxz;
    a[2][2] = -sxx + syy - szz;
    a[2][3] = syz + szy;
    a[3][3] = -sxx - syy + szz;
#pragma unroll
    for (int r = 1; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < r; ++c) a[r][c] = a[c][r];
    double v[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) v[r][c] = r == c ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 12; ++sweep) {
        double off = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = r + 1; c < 4; ++c) off += a[r][c] * a[r][c];
        if (off <= 0.0) break;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
#pragma unroll
            for (int qq = p + 1; qq < 4; ++qq) {
                if (a[p][qq] == 0.0) continue;
                const double theta = (a[qq][qq] - a[p][p]) / (2.0 * a[p][qq]);
                const double sign = theta >= 0.0 ? 1.0 : -1.0;
                const double t = sign / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double cs = 1.0 / sqrt(t * t + 1.0);
                const double sn = t * cs;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double f = a[k][p], s = a[k][qq];
                    a[k][p] = cs * f - sn * s;
                    a[k][qq] = sn * f + cs * s;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double f = a[p][k], s = a[qq][k];
                    a[p][k] = cs * f - sn * s;
                    a[qq][k] = sn * f + cs * s;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const double f = v[k][p], s = v[k][qq];
                    v[k][p] = cs * f - sn * s;
                    v[k][qq] = sn * f + cs * s;
                }
            }
        }
    }
    int best = 0;
    double top = a[0][0];
#pragma unroll
    for (int i = 1; i < 4; ++i)
        if (a[i][i] > top) {
            top = a[i][i];
            best = i;
        }
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = best == 0 ? v[k][0] : (best == 1 ? v[k][1] : (best == 2 ? v[k][2] : v[k][3]));
}

// Row-major rotation of the normalised quaternion (identity when it is zero).
__device__ inline m3 rotation_from_quaternion(const double q[4]) {
    const double len = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (!(len > 0.0)) return m3{{1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0}};
    const double w = q[0] / len, x = q[1] / len, y = q[2] / len, z = q[3] / len;
    return m3{{w * w + x * x - y * y - z * z, 2.0 * (x * y - w * z), 2.0 * (x * z + w * y),
               2.0 * (x * y + w * z), w * w - x * x + y * y - z * z, 2.0 * (y * z - w * x),
               2.0 * (x * z - w * y), 2.0 * (y * z + w * x), w * w - x * x - y * y + z * z}};
}

// One wave per node: the lanes split the node's bound (vertex, slot) rows and
// its edges, the nine correlation sums are reduced with wave butterflies, and
// lane 0 fits and writes the rotation.
__global__ void __launch_bounds__(kThreads)
fit_rotations_kernel(AssemblyInputs in, const int32_t* __restrict__ goff, const int32_t* __restrict__ gent,
                     const int32_t* __restrict__ eoff, const int32_t* __restrict__ eent, double data_weight,
                     double relaxation, double* __restrict__ rotations) {
    const int lane = threadIdx.x % kWave;
    const int64_t node = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
    if (node >= in.node_count) return;
    double S[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) S[k] = 0.0;
    const d3 g = ld3(in.rest, node), t = ld3(in.translations, node);
    for (int32_t k = goff[node] + lane; k < goff[node + 1]; k += kWave) {
        const int32_t ent = gent[k];
        const int64_t v = ent >> 2;
        const double own = in.bind_weights[4 * v + (ent & 3)];
        if (own == 0.0) continue;
        const d3 source = ld3(in.points, v) - g;
        const bool data = in.active ? in.active[v] != 0 : in.rejection[v] == 0;
        if (data) add_correlation(S, data_weight * in.robust[v] * own, source, (ld3(in.corr_points, v) - g) - t);
        if (in.w.landmark > 0.0 && in.landmark_flag[v] != 0)
            add_correlation(S, in.w.landmark * own, source, (ld3(in.landmark_target, v) - g) - t);
    }
    for (int32_t k = eoff[node] + lane; k < eoff[node + 1]; k += kWave) {
        const int32_t ent = eent[k];
        if (ent & 1) continue;  // only edges this node starts
        const int64_t e = ent >> 1;
        const int64_t b = in.edges[2 * e + 1];
        if (b < 0 || b >= in.node_count) continue;
        const d3 gb = ld3(in.rest, b);
        add_correlation(S, in.w.smooth, gb - g, ((gb + ld3(in.translations, b)) - g) - t);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) S[k] = wave_sum(S[k]);
    if (lane != 0) return;
    double q[4];
    horn_quaternion(S, q);
    const m3 fitted = rotation_from_quaternion(q);
    double* R = rotations + 9 * node;
    m3 cur;
#pragma unroll
    for (int k = 0; k < 9; ++k) cur.m[k] = R[k];
    if (relaxation >= 1.0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) R[k] = fitted.m[k];
        return;
    }
    m3 tr;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) tr.m[3 * r + c] = cur.m[3 * c + r];
    const m3 rel = mat_mul(fitted, tr);
    const d3 w = log_so3(rel.m) * relaxation;
    const m3 out = mat_mul(exp_so3(w), cur);
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = out.m[k];
}

template <typename Scalar>
__global__ void restrict_kernel(const int32_t* __restrict__ row_off, const int32_t* __restrict__ cols, int64_t nn,
                                Scalar* __restrict__ bsr) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    for (int32_t b = row_off[i]; b < row_off[i + 1]; ++b) {
        Scalar* blk = bsr + 36 * static_cast<int64_t>(b);
        const bool diag = cols[b] == i;
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                if (r >= 3 && c >= 3) continue;
                blk[6 * r + c] = (diag && r == c) ? Scalar(1) : Scalar(0);
            }
    }
}

}  // namespace

template <typename Scalar>
void restrict_to_translations(const LevelLayout& L, Scalar* bsr, hipStream_t stream) {
    if (L.node_count <= 0) return;
    restrict_kernel<Scalar><<<blocks_for(L.node_count, kThreads), kThreads, 0, stream>>>(
        L.row_offsets.get(), L.columns.get(), L.node_count, bsr);
    NRICP_LAUNCH_CHECK("restrict_kernel");
}
template void restrict_to_translations<double>(const LevelLayout&, double*, hipStream_t);
template void restrict_to_translations<float>(const LevelLayout&, float*, hipStream_t);

void fit_rotations(const AssemblyInputs& in, const LevelLayout& L, double relaxation, double* rotations,
                   hipStream_t stream) {
    if (in.node_count <= 0) return;
    // Both data terms share one fit weight: at a closest point they ask the
    // vertex for the same direction (src/cuda/local_global.cu:541-546).
    const double data_weight = in.w.p2p + in.w.plane;
    fit_rotations_kernel<<<blocks_for(in.node_count, kThreads / kWave), kThreads, 0, stream>>>(
        in, L.gather_offsets.get(), L.gather_entries.get(), L.node_edge_offsets.get(), L.node_edge_entries.get(),
        data_weight, relaxation, rotations);
    NRICP_LAUNCH_CHECK("fit_rotations_kernel");
}

}  // namespace nricp_hip
