// SO(3) helpers shared by the device kernels, FP64, with the operation order
// of src/core/rotation.h (exp: Rodrigues with the second-order series below
// theta^2 = 1e-24, :28-46; log: through the w >= 0 unit quaternion, :50-62).
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace nricp_hip {

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

// log through the (w >= 0) unit quaternion, as src/core/rotation.h:50-62.
__device__ inline d3 log_so3(const double* r) {
    double w, x, y, z;
    const double tr = r[0] + r[4] + r[8];
    if (tr > 0.0) {
        const double s = sqrt(tr + 1.0) * 2.0;
        w = 0.25 * s; x = (r[7] - r[5]) / s; y = (r[2] - r[6]) / s; z = (r[3] - r[1]) / s;
    } else if (r[0] > r[4] && r[0] > r[8]) {
        const double s = sqrt(1.0 + r[0] - r[4] - r[8]) * 2.0;
        w = (r[7] - r[5]) / s; x = 0.25 * s; y = (r[1] + r[3]) / s; z = (r[2] + r[6]) / s;
    } else if (r[4] > r[8]) {
        const double s = sqrt(1.0 + r[4] - r[0] - r[8]) * 2.0;
        w = (r[2] - r[6]) / s; x = (r[1] + r[3]) / s; y = 0.25 * s; z = (r[5] + r[7]) / s;
    } else {
        const double s = sqrt(1.0 + r[8] - r[0] - r[4]) * 2.0;
        w = (r[3] - r[1]) / s; x = (r[2] + r[6]) / s; y = (r[5] + r[7]) / s; z = 0.25 * s;
    }
    const double l = sqrt(w * w + x * x + y * y + z * z);
    if (l > 0.0) {
        w /= l; x /= l; y /= l; z /= l;
    } else {
        w = 1.0; x = y = z = 0.0;
    }
    if (w < 0.0) {
        w = -w; x = -x; y = -y; z = -z;
    }
    const double len = sqrt(x * x + y * y + z * z);
    if (len < 1e-12) return d3{0.0, 0.0, 0.0};
    const double angle = 2.0 * atan2(len, w);
    return d3{x, y, z} * (angle / len);
}

}  // namespace nricp_hip
