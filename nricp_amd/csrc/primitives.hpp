// Device primitives the layout and hierarchy builders rely on: a stable LSD
// radix sort of (u64 key, u32 value) pairs, an exclusive scan, CSR offsets
// from sorted keys, and first-of-run flags. All are deterministic: equal keys
// keep their input order, and sums are over fixed orders.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "buffers.hpp"

namespace nricp_hip {

// Digit passes of one sort (64-bit keys, 8-bit digits) and the words of one
// sort's table: the digit counts of every pass, then one tile ticket per pass.
constexpr int kSortMaxPasses = 8;
constexpr int kSortTableWords = kSortMaxPasses * 256 + kSortMaxPasses;

// Working memory of the sort and the scans. One scratch serves one stream at
// a time: sorts on it run in that stream's order.
struct SortScratch {
    DeviceBuffer<uint64_t> keys_alt;
    DeviceBuffer<uint32_t> vals_alt;
    DeviceBuffer<uint32_t> tables;   // two sort tables, used in turn
    DeviceBuffer<uint64_t> state;    // look-back words, 256 per tile
    int64_t state_words = 0;         // words of `state` cleared since it was allocated
    uint32_t epoch = 0;              // the last digit pass's tag
    int parity = 0;                  // the table the next sort uses
    DeviceBuffer<int64_t> scan_partials;  // tile sums of the scans
};

// Sorts keys[0..n) (and vals alongside, when non-null) by the low `key_bits`
// bits, stably (n < 2^31). The result is left in keys/vals.
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, int64_t n, int key_bits, SortScratch& scratch,
                      hipStream_t stream);

// out[i] = sum of in[0..i), out[n] = total. `out` has n + 1 entries and may
// alias `in`. Multi-workgroup, fixed summation order.
void exclusive_scan(const int64_t* in, int64_t* out, int64_t n, SortScratch& scratch, hipStream_t stream);
void exclusive_scan32(const int32_t* in, int32_t* out, int64_t n, SortScratch& scratch, hipStream_t stream);

// offsets[r] = first index i with (keys[i] >> shift) >= r, r in [0, rows];
// keys sorted ascending.
void offsets_from_sorted(const uint64_t* keys, int64_t n, int shift, uint64_t mask, int64_t rows,
                         int32_t* offsets, hipStream_t stream);

// flags[i] = 1 when keys[i] differs from keys[i-1] (flags[0] = 1 when n > 0).
void run_starts(const uint64_t* keys, int64_t n, int32_t* flags, hipStream_t stream);

int bits_for(uint64_t max_value);

}  // namespace nricp_hip
