// sw2d_pool.hpp -- device memory pool of a batch.
//
// One hipMalloc'd region, carved on the host with first-fit over an
// offset-ordered free list; freed ranges are merged with their neighbours.
// Scenarios of a batch allocate every device array from it, so admission is
// a question of whether a scenario's arrays fit (sw2d_batch_footprint), no
// hipMalloc/hipFree (which synchronise the device) runs while the batch is
// stepping, and freed HBM is reused by the next scenario.
#pragma once
#include <stddef.h>

#include <map>
#include <unordered_map>

namespace sw2d {

constexpr size_t kPoolAlign = 256;

class Pool {
public:
    void init(char *base, size_t cap)
    {
        base_ = base;
        cap_ = cap;
        free_.clear();
        used_.clear();
        in_use_ = peak_ = 0;
        if (cap) free_[0] = cap;
    }
    static size_t rounded(size_t bytes)
    {
        if (bytes == 0) bytes = 1;
        return (bytes + kPoolAlign - 1) / kPoolAlign * kPoolAlign;
    }
    // NULL when no free range is large enough
    void *alloc(size_t bytes)
    {
        const size_t n = rounded(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[off + n] = len - n;
            used_[off] = n;
            in_use_ += n;
            if (in_use_ > peak_) peak_ = in_use_;
            return base_ + off;
        }
        return nullptr;
    }
    // false when p was not allocated here
    bool release(void *p)
    {
        if (!p || (char *)p < base_ || (char *)p >= base_ + cap_) return false;
        const size_t off = (size_t)((char *)p - base_);
        auto u = used_.find(off);
        if (u == used_.end()) return false;
        size_t start = off, len = u->second;
        used_.erase(u);
        in_use_ -= len;
        auto next = free_.lower_bound(start);
        if (next != free_.end() && start + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
        return true;
    }
    size_t capacity() const { return cap_; }
    size_t in_use() const { return in_use_; }
    size_t peak() const { return peak_; }
    size_t largest_free() const
    {
        size_t m = 0;
        for (auto &f : free_) m = f.second > m ? f.second : m;
        return m;
    }
    char *base() const { return base_; }

private:
    char *base_ = nullptr;
    size_t cap_ = 0;
    std::map<size_t, size_t> free_;            // offset -> length
    std::unordered_map<size_t, size_t> used_;  // offset -> length
    size_t in_use_ = 0, peak_ = 0;
};

}  // namespace sw2d
