// Mapping of a launch's workgroups onto MI355X's eight XCDs.
//
// The dispatcher deals the workgroups of one launch to the XCDs in turn:
// workgroup b runs on XCD b % 8. Each XCD has its own L2, so a gather kernel
// whose consecutive workgroups touch neighbouring data should give each XCD a
// contiguous share of the work, not every eighth item. The functions here
// turn a workgroup index into such a logical index. Plain C++ (no HIP types),
// so the host tests compile the same code.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define NRICP_HD __host__ __device__
#else
#define NRICP_HD
#endif

namespace nricp_hip {

constexpr int64_t kXcds = 8;

NRICP_HD inline int64_t xcd_min(int64_t a, int64_t b) { return a < b ? a : b; }

// A bijection of [0, total): the workgroups on XCD x (b = x, x + 8, ...)
// take the logical indices [x q + min(x, r), ...) in order, where
// total = 8 q + r, so XCD x holds one contiguous eighth.
NRICP_HD inline int64_t xcd_contiguous(int64_t b, int64_t total) {
    const int64_t x = b % kXcds, i = b / kXcds;
    const int64_t q = total / kXcds, r = total % kXcds;
    return x * q + xcd_min(x, r) + i;
}

// A bijection of [0, total) that gives each XCD runs of `chunk` consecutive
// logical indices, the runs dealt to the XCDs in turn: workgroup b (XCD
// b % 8, its (b / 8)-th there) takes run (b / 8) / chunk of that XCD. The
// indices past the last whole round of 8 runs map to themselves.
NRICP_HD inline int64_t xcd_chunked(int64_t b, int64_t total, int64_t chunk) {
    const int64_t round = kXcds * chunk;
    if (chunk <= 1 || b >= total / round * round) return b;
    const int64_t x = b % kXcds, i = b / kXcds;
    return (i / chunk) * round + x * chunk + i % chunk;
}

// A launch of `total` workgroups that does two kinds of work: the first
// `first` logical items of kind 0, the remaining total - first of kind 1.
// Each XCD takes a contiguous eighth of kind 0 (as xcd_contiguous over
// `first`) and then a contiguous share of kind 1, so both kinds of one region
// of the data meet in the same L2. Returns the kind in *kind and the index
// within that kind.
NRICP_HD inline int64_t xcd_split(int64_t b, int64_t total, int64_t first, int* kind) {
    const int64_t x = b % kXcds, i = b / kXcds;
    const int64_t qt = total / kXcds, rt = total % kXcds;
    const int64_t qn = first / kXcds, rn = first % kXcds;
    const int64_t nx = qn + (x < rn ? 1 : 0);  // kind-0 workgroups on XCD x
    if (i < nx) {
        *kind = 0;
        return x * qn + xcd_min(x, rn) + i;
    }
    // XCD x holds cnt_x = qt + (x < rt) workgroups; those past its kind-0
    // share are kind 1. Their start is the sum over the XCDs before it of
    // cnt_y - n_y.
    *kind = 1;
    return x * (qt - qn) + xcd_min(x, rt) - xcd_min(x, rn) + (i - nx);
}

}  // namespace nricp_hip
