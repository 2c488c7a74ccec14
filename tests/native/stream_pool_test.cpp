// Host test of the device-memory cache's ordering policy
// (nricp_amd/csrc/stream_pool.hpp), with integers for streams and blocks.
// Exits non-zero with a message on the first broken rule.
#include <cstdio>
#include <cstdlib>

#include "stream_pool.hpp"

using Pool = nricp_hip::StreamOrderedPool<long, long>;

static int failures = 0;
#define EXPECT(cond)                                                        \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                     \
        }                                                                   \
    } while (0)

int main() {
    // Size classes: powers of two to 1 MiB, whole MiB above.
    EXPECT(Pool::size_class(1) == 256);
    EXPECT(Pool::size_class(257) == 512);
    EXPECT(Pool::size_class(1 << 20) == (1 << 20));
    EXPECT(Pool::size_class((1 << 20) + 1) == (2 << 20));

    const int dev = 0;
    const long A = 11, B = 22;
    const int64_t cls = 4096;
    long got = 0;
    {
        Pool p;
        p.stream_created(A);
        p.stream_created(B);
        p.release(dev, A, cls, 100);
        // Released on A while A lives: B must not get it (A's queued work may
        // still use it), a request on no stream neither; A gets it back.
        EXPECT(!p.take(dev, B, cls, &got));
        EXPECT(!p.take(dev, 0, cls, &got));
        EXPECT(p.take(dev, A, cls, &got) && got == 100);
        EXPECT(p.cached() == 0);
    }
    {
        Pool p;
        p.stream_created(A);
        p.stream_created(B);
        p.release(dev, A, cls, 100);
        // A retires (synchronised, destroyed): its blocks are idle for anyone.
        p.stream_retired(A);
        EXPECT(p.take(dev, B, cls, &got) && got == 100);
    }
    {
        Pool p;
        p.stream_created(B);
        // Released on a stream the pool never saw live (already retired):
        // idle at once.
        p.release(dev, A, cls, 200);
        EXPECT(p.take(dev, B, cls, &got) && got == 200);
        // Released on no stream: idle.
        p.release(dev, 0, cls, 300);
        EXPECT(p.take(dev, B, cls, &got) && got == 300);
    }
    {
        Pool p;
        p.stream_created(A);
        p.stream_created(B);
        p.release(dev, A, cls, 100);
        p.release(dev, 0, cls, 300);
        // A request takes its own stream's block before an idle one; other
        // devices and classes never mix.
        EXPECT(p.take(dev, A, cls, &got) && got == 100);
        EXPECT(!p.take(1, A, cls, &got));
        EXPECT(!p.take(dev, A, 2 * cls, &got));
        EXPECT(p.take(dev, A, cls, &got) && got == 300);
    }
    {
        Pool p;
        p.stream_created(A);
        p.release(0, A, cls, 1);
        p.release(1, A, cls, 2);
        p.release(1, 0, cls, 3);
        long given = 0;
        int count = 0;
        p.drain(1, [&](int device, long b) {
            EXPECT(device == 1);
            given += b;
            ++count;
        });
        EXPECT(count == 2 && given == 5 && p.cached() == 1);
        p.drain(-1, [&](int, long) { ++count; });
        EXPECT(count == 3 && p.cached() == 0);
    }
    if (failures) return 1;
    std::puts("ok");
    return 0;
}
