// arena_test.cpp - the serving loop's first-fit allocator (mkf_arena.hpp)
// against a brute-force model: a map of 256-byte granules. Random
// allocations and releases; after each step every live block lies inside the
// reservation, aligned, disjoint from the others, at the offset first fit
// gives (the start of the lowest maximal free run long enough), the free
// list is exactly the model's maximal free runs, and used() is the live
// total. tests/test_sanitize.py builds it with g++ under ASan/UBSan.
#include <stdio.h>
#include <stdlib.h>

#include <utility>
#include <vector>

#include "mkf_arena.hpp"

using mkf::ArenaMap;

static int fails = 0;
#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      if (++fails > 20) exit(1);                                          \
    }                                                                     \
  } while (0)

int main() {
  const size_t G = ArenaMap::kAlign;
  unsigned long long rng = 12345;
  auto rnd = [&]() {
    rng = rng * 6364136223846793005ull + 1442695040888963407ull;
    return (size_t)(rng >> 33);
  };
  for (int trial = 0; trial < 40; ++trial) {
    const size_t cap = G * (1 + rnd() % 400) - (trial % 3) * 7;  // not always a multiple of the granule
    ArenaMap a;
    a.reset(cap);
    const size_t granules = a.capacity() / G;
    CHECK(a.capacity() == ArenaMap::alignUp(cap) && a.capacity() % G == 0);
    std::vector<char> busy(granules, 0);
    std::vector<std::pair<size_t, size_t>> live;  // offset, requested bytes
    for (int step = 0; step < 2000; ++step) {
      if (live.empty() || rnd() % 100 < 55) {
        const size_t n = 1 + rnd() % (G * (1 + granules / 4));
        const size_t need = (ArenaMap::alignUp(n)) / G;
        // model: start of the lowest maximal free run of at least `need` granules
        size_t want = ArenaMap::kNone;
        for (size_t i = 0; i < granules;) {
          if (busy[i]) { ++i; continue; }
          size_t j = i;
          while (j < granules && !busy[j]) ++j;
          if (j - i >= need) { want = i * G; break; }
          i = j;
        }
        const size_t off = a.alloc(n);
        CHECK(off == want);
        if (off != ArenaMap::kNone) {
          CHECK(off % G == 0 && off + need * G <= a.capacity());
          for (size_t g = off / G; g < off / G + need; ++g) {
            CHECK(!busy[g]);
            busy[g] = 1;
          }
          live.push_back({off, n});
        }
      } else {
        const size_t k = rnd() % live.size();
        const size_t off = live[k].first, n = live[k].second;
        a.release(off, n);
        for (size_t g = off / G; g < off / G + ArenaMap::alignUp(n) / G; ++g) busy[g] = 0;
        live[k] = live.back();
        live.pop_back();
      }
      size_t used = 0;
      for (char b : busy) used += b ? G : 0;
      CHECK(a.used() == used);
      // the free list is exactly the maximal free runs of the model
      std::vector<std::pair<size_t, size_t>> runs;
      for (size_t i = 0; i < granules;) {
        if (busy[i]) { ++i; continue; }
        size_t j = i;
        while (j < granules && !busy[j]) ++j;
        runs.push_back({i * G, (j - i) * G});
        i = j;
      }
      CHECK(runs.size() == a.freeBlocks().size());
      size_t r = 0;
      for (const auto& fb : a.freeBlocks()) {
        if (r < runs.size()) CHECK(fb.first == runs[r].first && fb.second == runs[r].second);
        ++r;
      }
    }
    for (const auto& b : live) a.release(b.first, b.second);
    CHECK(a.used() == 0 && a.freeBlocks().size() == (granules ? 1u : 0u));
  }
  if (fails) return 1;
  printf("arena: ok\n");
  return 0;
}
