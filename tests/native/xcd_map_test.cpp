// Host test of the XCD work mapping (nricp_amd/csrc/xcd_map.hpp): the
// mappings are bijections; xcd_contiguous and xcd_split give each XCD one
// contiguous run (per kind), in XCD order; xcd_chunked gives it runs of
// `chunk` dealt in turn. Exits non-zero on failure.
#include <cstdio>
#include <vector>

#include "xcd_map.hpp"

using namespace nricp_hip;

static int failures = 0;
#define EXPECT(cond, ...)                                                  \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::fprintf(stderr, "%s:%d: expected %s: ", __FILE__, __LINE__, #cond); \
            std::fprintf(stderr, __VA_ARGS__);                             \
            std::fprintf(stderr, "\n");                                    \
            if (++failures > 20) return 1;                                 \
        }                                                                  \
    } while (0)

int main() {
    // xcd_chunked: a bijection; within each whole round of 8 runs, XCD x
    // (workgroups b = x + 8 i) takes the run [round start + x chunk, + chunk)
    // in order; past the last whole round every index maps to itself.
    for (int64_t chunk : {1, 2, 3, 16}) {
        for (int64_t total = 0; total <= 700; total += 7) {
            std::vector<int> seen(total, 0);
            const int64_t round = kXcds * chunk, whole = total / round * round;
            for (int64_t b = 0; b < total; ++b) {
                const int64_t l = xcd_chunked(b, total, chunk);
                EXPECT(l >= 0 && l < total, "chunk %lld total %lld b %lld -> %lld", (long long)chunk,
                       (long long)total, (long long)b, (long long)l);
                if (l < 0 || l >= total) continue;
                ++seen[l];
                const int64_t x = b % kXcds, i = b / kXcds;
                const int64_t want = b >= whole || chunk <= 1 ? b : (i / chunk) * round + x * chunk + i % chunk;
                EXPECT(l == want, "chunk %lld total %lld b %lld -> %lld, not %lld", (long long)chunk,
                       (long long)total, (long long)b, (long long)l, (long long)want);
                if (b < whole && chunk > 1)
                    EXPECT(l / chunk % kXcds == x, "chunk %lld total %lld b %lld: run of another XCD",
                           (long long)chunk, (long long)total, (long long)b);
            }
            for (int64_t l = 0; l < total; ++l)
                EXPECT(seen[l] == 1, "chunk %lld total %lld index %lld seen %d", (long long)chunk, (long long)total,
                       (long long)l, seen[l]);
        }
    }
    for (int64_t total = 0; total <= 300; ++total) {
        std::vector<int> seen(total, 0);
        std::vector<int64_t> lo(kXcds, total), hi(kXcds, -1), cnt(kXcds, 0);
        for (int64_t b = 0; b < total; ++b) {
            const int64_t l = xcd_contiguous(b, total);
            EXPECT(l >= 0 && l < total, "total %lld b %lld -> %lld", (long long)total, (long long)b, (long long)l);
            if (l < 0 || l >= total) continue;
            ++seen[l];
            const int64_t x = b % kXcds;
            lo[x] = l < lo[x] ? l : lo[x];
            hi[x] = l > hi[x] ? l : hi[x];
            ++cnt[x];
        }
        for (int64_t l = 0; l < total; ++l) EXPECT(seen[l] == 1, "total %lld index %lld seen %d", (long long)total, (long long)l, seen[l]);
        int64_t next = 0;
        for (int64_t x = 0; x < kXcds; ++x) {
            if (cnt[x] == 0) continue;
            EXPECT(hi[x] - lo[x] + 1 == cnt[x], "total %lld xcd %lld not contiguous", (long long)total, (long long)x);
            EXPECT(lo[x] == next, "total %lld xcd %lld starts at %lld, not %lld", (long long)total, (long long)x,
                   (long long)lo[x], (long long)next);
            next = hi[x] + 1;
        }
        for (int64_t first = 0; first <= total; ++first) {
            std::vector<int> seen0(first, 0), seen1(total - first, 0);
            std::vector<int64_t> lo0(kXcds, total), hi0(kXcds, -1), c0(kXcds, 0);
            std::vector<int64_t> lo1(kXcds, total), hi1(kXcds, -1), c1(kXcds, 0);
            for (int64_t b = 0; b < total; ++b) {
                int kind = -1;
                const int64_t l = xcd_split(b, total, first, &kind);
                const int64_t x = b % kXcds;
                if (kind == 0) {
                    EXPECT(l >= 0 && l < first, "split %lld/%lld b %lld kind 0 -> %lld", (long long)first,
                           (long long)total, (long long)b, (long long)l);
                    if (l < 0 || l >= first) continue;
                    ++seen0[l];
                    lo0[x] = l < lo0[x] ? l : lo0[x];
                    hi0[x] = l > hi0[x] ? l : hi0[x];
                    ++c0[x];
                } else {
                    EXPECT(kind == 1, "kind %d", kind);
                    EXPECT(l >= 0 && l < total - first, "split %lld/%lld b %lld kind 1 -> %lld", (long long)first,
                           (long long)total, (long long)b, (long long)l);
                    if (l < 0 || l >= total - first) continue;
                    ++seen1[l];
                    lo1[x] = l < lo1[x] ? l : lo1[x];
                    hi1[x] = l > hi1[x] ? l : hi1[x];
                    ++c1[x];
                }
            }
            for (int64_t l = 0; l < first; ++l) EXPECT(seen0[l] == 1, "split %lld/%lld kind 0 index %lld seen %d", (long long)first, (long long)total, (long long)l, seen0[l]);
            for (int64_t l = 0; l < total - first; ++l) EXPECT(seen1[l] == 1, "split %lld/%lld kind 1 index %lld seen %d", (long long)first, (long long)total, (long long)l, seen1[l]);
            int64_t n0 = 0, n1 = 0;
            for (int64_t x = 0; x < kXcds; ++x) {
                if (c0[x] > 0) {
                    EXPECT(hi0[x] - lo0[x] + 1 == c0[x] && lo0[x] == n0, "split %lld/%lld xcd %lld kind 0 run",
                           (long long)first, (long long)total, (long long)x);
                    n0 = hi0[x] + 1;
                }
                if (c1[x] > 0) {
                    EXPECT(hi1[x] - lo1[x] + 1 == c1[x] && lo1[x] == n1, "split %lld/%lld xcd %lld kind 1 run",
                           (long long)first, (long long)total, (long long)x);
                    n1 = hi1[x] + 1;
                }
            }
        }
    }
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
