// arena.h — first-fit range allocator over one device allocation (the serving loop's HBM budget).
//
// Host-only bookkeeping (no HIP calls): offsets into an arena of `size` bytes, aligned to `align`, with free-range
// coalescing. Not thread-safe; the owner serialises calls.
#pragma once
#include <cstdint>
#include <iterator>
#include <map>

namespace bz {

class Arena {
public:
    static constexpr uint64_t kFail = ~0ull;

    explicit Arena(uint64_t size = 0, uint64_t align = 256) { reset(size, align); }

    void reset(uint64_t size, uint64_t align = 256)
    {
        size_ = size - size % align;
        align_ = align;
        used_ = peak_ = 0;
        free_.clear();
        if (size_) free_[0] = size_;
    }

    static uint64_t round_up(uint64_t n, uint64_t a) { return (n + a - 1) / a * a; }

    // Offset of a block of at least `n` bytes (first fit), or kFail.
    uint64_t alloc(uint64_t n)
    {
        if (n == 0) n = 1;
        n = round_up(n, align_);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const uint64_t off = it->first, rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            used_ += n;
            if (used_ > peak_) peak_ = used_;
            return off;
        }
        return kFail;
    }

    // Returns a block; `n` must be the size given to alloc.
    void release(uint64_t off, uint64_t n)
    {
        if (n == 0) n = 1;
        n = round_up(n, align_);
        used_ -= n;
        auto next = free_.lower_bound(off);
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {  // merge with the preceding free range
                off = prev->first;
                n += prev->second;
                free_.erase(prev);
            }
        }
        next = free_.lower_bound(off);
        if (next != free_.end() && off + n == next->first) {  // and the following one
            n += next->second;
            free_.erase(next);
        }
        free_[off] = n;
    }

    // Whether a block of n bytes could ever fit (the whole arena free).
    bool fits_empty(uint64_t n) const { return round_up(n ? n : 1, align_) <= size_; }
    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }
    uint64_t peak() const { return peak_; }
    size_t free_ranges() const { return free_.size(); }
    uint64_t largest_free() const
    {
        uint64_t m = 0;
        for (const auto& kv : free_)
            if (kv.second > m) m = kv.second;
        return m;
    }

private:
    uint64_t size_ = 0, align_ = 256, used_ = 0, peak_ = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> length
};

}  // namespace bz
