// dyc_arena.hpp -- best-fit allocator over one device allocation (host bookkeeping only).
//
// The serving loop keeps every table's subtables in one hipMalloc'd region so that "free HBM"
// is an exact, host-known number and no hipMalloc/hipFree (which synchronise the device) runs
// while ticks are in flight. Offsets are 256-byte aligned. A range released here may be handed
// out again at once: every user of the region issues its work on one stream, so a kernel that
// writes a reused range is ordered after every kernel that read it before the release.
#pragma once

#include <algorithm>
#include <cstdint>
#include <iterator>
#include <map>

namespace dyc {

class Arena {
  public:
    static constexpr uint64_t kAlign = 256;
    static constexpr uint64_t kNone = ~0ull;

    void reset(uint64_t bytes)
    {
        free_.clear();
        live_.clear();
        size_ = bytes & ~(kAlign - 1);
        used_ = peak_ = 0;
        if (size_) free_[0] = size_;
    }

    // Offset of a new range of at least `bytes` (smallest free block that fits), or kNone.
    uint64_t alloc(uint64_t bytes)
    {
        bytes = std::max<uint64_t>(kAlign, (bytes + kAlign - 1) & ~(kAlign - 1));
        auto best = free_.end();
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes && (best == free_.end() || it->second < best->second)) best = it;
        if (best == free_.end()) return kNone;
        const uint64_t off = best->first, sz = best->second;
        free_.erase(best);
        if (sz > bytes) free_[off + bytes] = sz - bytes;
        live_[off] = bytes;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return off;
    }

    // Return a range handed out by alloc; false if `off` is not the start of a live range.
    bool release(uint64_t off)
    {
        auto a = live_.find(off);
        if (a == live_.end()) return false;
        const uint64_t sz = a->second;
        live_.erase(a);
        used_ -= sz;
        auto it = free_.emplace(off, sz).first;
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
        return true;
    }

    uint64_t largest_free() const
    {
        uint64_t m = 0;
        for (const auto &f : free_) m = std::max(m, f.second);
        return m;
    }
    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }
    uint64_t peak() const { return peak_; }
    uint64_t free_bytes() const { return size_ - used_; }
    size_t free_blocks() const { return free_.size(); }
    size_t live_blocks() const { return live_.size(); }

  private:
    std::map<uint64_t, uint64_t> free_;  // offset -> bytes, no two adjacent
    std::map<uint64_t, uint64_t> live_;  // offset -> bytes
    uint64_t size_ = 0, used_ = 0, peak_ = 0;
};

}  // namespace dyc
