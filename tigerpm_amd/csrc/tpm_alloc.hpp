// tpm_alloc.hpp -- first-fit range allocator of the batch arenas (host
// code only, no HIP dependency).
#pragma once

#include <cstdint>
#include <iterator>
#include <map>

// First-fit allocator of offsets in [0, cap) (host side; the arenas are
// allocated once, boxes take and return ranges of them).
struct tpm_range_alloc {
    int64_t cap = 0;
    std::map<int64_t, int64_t> free_;  // offset -> size, coalesced
    void reset(int64_t c) {
        cap = c;
        free_.clear();
        if (c > 0) free_[0] = c;
    }
    int64_t alloc(int64_t size, int64_t align) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            const int64_t off = (it->first + align - 1) / align * align;
            const int64_t end = it->first + it->second;
            if (off + size > end) continue;
            const int64_t b0 = it->first;
            free_.erase(it);
            if (off > b0) free_[b0] = off - b0;
            if (end > off + size) free_[off + size] = end - (off + size);
            return off;
        }
        return -1;
    }
    void release(int64_t off, int64_t size) {
        auto it = free_.emplace(off, size).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    // One past the highest offset in use.
    int64_t high_water() const {
        if (free_.empty()) return cap;
        auto last = std::prev(free_.end());
        return last->first + last->second == cap ? last->first : cap;
    }
};

