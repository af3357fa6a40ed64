// mkf_arena.hpp - first-fit offset allocator over one reservation, freed
// blocks merging with their neighbours. Host code only, no device calls: the
// serving loop owns the device memory and asks this class for offsets.
#pragma once

#include <stddef.h>

#include <iterator>
#include <map>

namespace mkf {

class ArenaMap {
 public:
  static constexpr size_t kAlign = 256;
  static constexpr size_t kNone = ~(size_t)0;

  static size_t alignUp(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

  // Starts over with one free block of `bytes` (rounded up to kAlign).
  void reset(size_t bytes) {
    cap_ = alignUp(bytes);
    used_ = 0;
    free_.clear();
    if (cap_) free_[0] = cap_;
  }
  // Offset of a block of n bytes (rounded up to kAlign), the lowest that
  // fits, or kNone.
  size_t alloc(size_t n) {
    n = alignUp(n);
    if (n == 0) n = kAlign;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      const size_t off = it->first, rest = it->second - n;
      free_.erase(it);
      if (rest) free_[off + n] = rest;
      used_ += n;
      return off;
    }
    return kNone;
  }
  // Returns the block at off of n bytes (as passed to alloc).
  void release(size_t off, size_t n) {
    n = alignUp(n);
    if (n == 0) n = kAlign;
    used_ -= n;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + n == next->first) {
      n += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += n;
        return;
      }
    }
    free_[off] = n;
  }
  size_t capacity() const { return cap_; }
  size_t used() const { return used_; }
  // Free blocks (offset -> size), for checks.
  const std::map<size_t, size_t>& freeBlocks() const { return free_; }

 private:
  size_t cap_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;
};

}  // namespace mkf
