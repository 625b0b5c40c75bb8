// First-fit block allocator over one pre-allocated range (device HBM or pinned
// host memory) for the request server: admission reserves a request's buffers
// here, so serving never calls the runtime allocator per request.  Host-only
// bookkeeping, not thread-safe (the server serialises calls).
#pragma once
#include <cstddef>
#include <iterator>
#include <map>

namespace fpcc {

class Arena {
public:
    static constexpr size_t kAlign = 256;

    explicit Arena(size_t size = 0) { reset(size); }

    void reset(size_t size)
    {
        size_ = size / kAlign * kAlign;
        free_.clear();
        used_ = 0;
        if (size_) free_[0] = size_;
    }

    static size_t round(size_t bytes) { return (bytes + kAlign - 1) / kAlign * kAlign; }

    // Reserves round(bytes) (at least one block) at the lowest offset that fits.
    bool alloc(size_t bytes, size_t &off)
    {
        const size_t need = round(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            off = it->first;
            const size_t rest = it->second - need;
            free_.erase(it);
            if (rest) free_[off + need] = rest;
            used_ += need;
            return true;
        }
        return false;
    }

    // Returns a block from alloc(bytes) and merges it with free neighbours.
    void release(size_t off, size_t bytes)
    {
        size_t len = round(bytes ? bytes : 1);
        used_ -= len;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += len;
                return;
            }
        }
        free_[off] = len;
    }

    // Whether blocks of these sizes could all be reserved right now (first fit on a
    // copy of the free list).
    template <size_t K>
    bool fits(const size_t (&bytes)[K]) const
    {
        Arena t(*this);
        size_t off;
        for (size_t k = 0; k < K; ++k)
            if (bytes[k] && !t.alloc(bytes[k], off)) return false;
        return true;
    }

    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t free_bytes() const { return size_ - used_; }
    size_t largest_free() const
    {
        size_t m = 0;
        for (const auto &f : free_) m = f.second > m ? f.second : m;
        return m;
    }
    size_t fragments() const { return free_.size(); }

private:
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length, disjoint and non-adjacent
};

}  // namespace fpcc
