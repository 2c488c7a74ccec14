// The ordering policy of the device-memory cache, independent of the HIP
// runtime (memory.hip instantiates it with hipStream_t and device pointers; a
// host test instantiates it with plain integers).
//
// A released block goes back to the free list of the stream it was ordered
// on, and only a request ordered on that same stream takes it again: the work
// still queued on the stream that uses the old contents then runs before the
// new owner's first kernel. A block released on no stream, or on a stream
// that has since retired (synchronised and destroyed: no work left), is idle
// and any stream may take it. A request takes from its own stream's list
// first, then from the idle list.
#pragma once

#include <cstdint>
#include <map>
#include <set>
#include <tuple>
#include <utility>
#include <vector>

namespace nricp_hip {

template <typename Stream, typename Block>
class StreamOrderedPool {
public:
    // Powers of two up to 1 MiB, then whole MiB.
    static int64_t size_class(int64_t bytes) {
        if (bytes <= 256) return 256;
        constexpr int64_t kMiB = int64_t{1} << 20;
        if (bytes <= kMiB) {
            int64_t c = 256;
            while (c < bytes) c <<= 1;
            return c;
        }
        return (bytes + kMiB - 1) / kMiB * kMiB;
    }

    // A cached block of class `cls` for a request ordered on `stream` (a null
    // stream takes only idle blocks); false when none is cached.
    bool take(int device, Stream stream, int64_t cls, Block* out) {
        auto it = stream != Stream{} ? free_.find(Key{device, stream, cls}) : free_.end();
        if (it == free_.end()) it = free_.find(Key{device, Stream{}, cls});
        if (it == free_.end()) return false;
        *out = it->second;
        free_.erase(it);
        return true;
    }

    // A block released by its owner, ordered on `stream` (null: no stream,
    // the caller has drained the device).
    void release(int device, Stream stream, int64_t cls, Block block) {
        const Stream owner = stream != Stream{} && live_.count(stream) ? stream : Stream{};
        free_.emplace(Key{device, owner, cls}, block);
    }

    void stream_created(Stream s) { live_.insert(s); }

    // The stream has no work left: its blocks become idle.
    void stream_retired(Stream s) {
        live_.erase(s);
        std::vector<std::pair<Key, Block>> moved;
        for (auto it = free_.begin(); it != free_.end();) {
            if (std::get<1>(it->first) == s) {
                moved.emplace_back(Key{std::get<0>(it->first), Stream{}, std::get<2>(it->first)}, it->second);
                it = free_.erase(it);
            } else {
                ++it;
            }
        }
        for (auto& m : moved) free_.emplace(m.first, m.second);
    }

    // Hands every cached block of `device` (all devices when negative) to
    // `give_back`, emptying those lists; the caller has drained the device.
    template <typename F>
    void drain(int device, F&& give_back) {
        for (auto it = free_.begin(); it != free_.end();) {
            if (device < 0 || std::get<0>(it->first) == device) {
                give_back(std::get<0>(it->first), it->second);
                it = free_.erase(it);
            } else {
                ++it;
            }
        }
    }

    size_t cached() const { return free_.size(); }

private:
    using Key = std::tuple<int, Stream, int64_t>;  // (device, owning stream or idle, class)
    std::multimap<Key, Block> free_;
    std::set<Stream> live_;  // streams created and not yet retired
};

}  // namespace nricp_hip
