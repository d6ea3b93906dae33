// Offset allocator over one reserved range [0, size): best fit by size, and
// freed ranges coalesce with free neighbours.  Host-only; the device pool
// places its blocks at base + offset.  A serving worker reserves its memory
// budget once, so batches of changing sizes make no device allocations (a
// large hipMalloc that the runtime cannot serve from freed memory takes
// about a second).
#pragma once

#include <cstddef>
#include <map>

namespace tr {

class Arena {
  public:
    void init(size_t size) {
        by_off_.clear();
        by_size_.clear();
        size_ = size;
        free_bytes_ = 0;
        if (size) insert(0, size);
    }
    size_t size() const { return size_; }
    size_t free_bytes() const { return free_bytes_; }
    size_t free_ranges() const { return by_off_.size(); }

    // smallest free range that holds `bytes`; false if none does
    bool alloc(size_t bytes, size_t &off) {
        if (bytes == 0) return false;
        auto it = by_size_.lower_bound(bytes);
        if (it == by_size_.end()) return false;
        const size_t sz = it->first, o = it->second;
        by_size_.erase(it);
        by_off_.erase(o);
        free_bytes_ -= sz;
        if (sz > bytes) insert(o + bytes, sz - bytes);
        off = o;
        return true;
    }

    // returns [off, off + bytes), merging it with adjacent free ranges
    void free(size_t off, size_t bytes) {
        size_t o = off, sz = bytes;
        auto next = by_off_.lower_bound(o);
        if (next != by_off_.end() && next->first == o + sz) {
            sz += next->second;
            erase(next);
        }
        auto prev = by_off_.lower_bound(o);
        if (prev != by_off_.begin()) {
            --prev;
            if (prev->first + prev->second == o) {
                o = prev->first;
                sz += prev->second;
                erase(prev);
            }
        }
        insert(o, sz);
    }

  private:
    void insert(size_t o, size_t sz) {
        by_off_[o] = sz;
        by_size_.emplace(sz, o);
        free_bytes_ += sz;
    }
    void erase(std::map<size_t, size_t>::iterator it) {
        auto r = by_size_.equal_range(it->second);
        for (auto j = r.first; j != r.second; ++j)
            if (j->second == it->first) {
                by_size_.erase(j);
                break;
            }
        free_bytes_ -= it->second;
        by_off_.erase(it);
    }

    std::map<size_t, size_t> by_off_;         // free ranges: offset -> size
    std::multimap<size_t, size_t> by_size_;  // free ranges: size -> offset
    size_t size_ = 0, free_bytes_ = 0;
};

}  // namespace tr
