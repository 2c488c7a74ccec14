// Device radix sort (nricp_hip::radix_sort_pairs) against std::stable_sort on
// the host: random keys with heavy duplication, sizes on and around the
// 2,048-key tile boundary, 1 to 8 digit passes, and many sorts in a row on one
// scratch (its two count tables, look-back epochs and growing and shrinking
// sizes). Exit status 0 when every case matches bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>

#include "primitives.hpp"

using namespace nricp_hip;

namespace {

bool run_case(SortScratch& scratch, hipStream_t stream, std::mt19937_64& rng, int64_t n, int key_bits,
              int distinct_bits, bool with_values) {
    const uint64_t mask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1);
    const uint64_t dmask = distinct_bits >= 64 ? ~0ull : ((1ull << distinct_bits) - 1);
    std::vector<uint64_t> keys(n);
    for (auto& k : keys) {
        // Bits above key_bits are set too: the sort must order by the low
        // key_bits only and carry the rest along.
        const uint64_t low = (rng() & dmask) * (mask / std::max<uint64_t>(dmask, 1));
        k = (rng() & ~mask) | (low & mask);
    }
    std::vector<uint32_t> vals(n);
    std::iota(vals.begin(), vals.end(), 0u);
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return (keys[a] & mask) < (keys[b] & mask); });

    DeviceBuffer<uint64_t> dk;
    DeviceBuffer<uint32_t> dv;
    dk.upload(keys, stream);
    if (with_values) dv.upload(vals, stream);
    radix_sort_pairs(dk.get(), with_values ? dv.get() : nullptr, n, key_bits, scratch, stream);
    const std::vector<uint64_t> got_k = dk.to_host(stream);
    std::vector<uint32_t> got_v;
    if (with_values) got_v = dv.to_host(stream);
    for (int64_t i = 0; i < n; ++i) {
        // Keys with equal low bits may differ above them; the stable order
        // fixes which key lands where, so compare whole keys.
        const bool key_ok = got_k[i] == keys[order[i]];
        const bool val_ok = !with_values || got_v[i] == order[i];
        if (!key_ok || !val_ok) {
            std::printf("MISMATCH n=%lld bits=%d distinct=%d values=%d at %lld: key %llx want %llx, val %u want %u\n",
                        static_cast<long long>(n), key_bits, distinct_bits, with_values ? 1 : 0,
                        static_cast<long long>(i), static_cast<unsigned long long>(got_k[i]),
                        static_cast<unsigned long long>(keys[order[i]]), with_values ? got_v[i] : 0u, order[i]);
            return false;
        }
    }
    std::printf("ok n=%lld bits=%d distinct=%d values=%d\n", static_cast<long long>(n), key_bits, distinct_bits,
                with_values ? 1 : 0);
    std::fflush(stdout);
    return true;
}

}  // namespace

int main() {
    hipStream_t stream;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
        std::printf("no HIP stream\n");
        return 2;
    }
    std::mt19937_64 rng(20240611);
    bool ok = true;
    {
        StreamScope scope(stream);
        SortScratch scratch;
        const int64_t sizes[] = {2, 3, 255, 2047, 2048, 2049, 4096, 6143, 100003, 1300000, 40000, 2049, 700001};
        const int bits[] = {1, 7, 8, 9, 13, 26, 33, 50, 64};
        int c = 0;
        for (int64_t n : sizes) {
            for (int b : bits) {
                // Few distinct keys (long runs of equal keys across tiles)
                // and many, in turn; values carried every other case.
                const int distinct = (c % 3 == 0) ? std::min(b, 3) : b;
                ok = run_case(scratch, stream, rng, n, b, distinct, c % 2 == 0) && ok;
                ++c;
                if (!ok) break;
            }
            if (!ok) break;
        }
        // A fresh scratch for a large sort first (its words are allocated at
        // the largest size straight away).
        if (ok) {
            SortScratch other;
            ok = run_case(other, stream, rng, 2500000, 50, 50, true) && run_case(other, stream, rng, 5000, 18, 2, true);
        }
    }
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
    std::printf(ok ? "ALL OK\n" : "FAILED\n");
    return ok ? 0 : 1;
}
