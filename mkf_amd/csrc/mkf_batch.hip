// mkf_batch.hip - batched multi-level histograms: many fields of mixed dims
// and element types, each evaluated at up to kMaxLevels threshold levels, in
// one persistent launch over a device work queue, captured in a HIP graph.
//
// Every field is read from HBM once for all of its levels: a wave loads one
// stack of rows per z-step and derives the pair bits of every level from the
// same registers. Work items are (field, x-tile, y-block, slab) tiles sorted
// by field; persistent workgroups take items from an atomic counter and add
// their LDS histograms to the field's outputs when the next item belongs to
// another field (or at the end).
//
// A workgroup of kBatchWaves waves holds 8 / segs row stacks side by side
// (wave w: segment w % segs of the x-tile, stack w / segs), so narrow fields
// still use every wave. Neighbouring segments of one stack exchange the
// voxel right of each row through LDS; the last segment of an x-tile reads
// it from memory.
//
// Per level and row the wave forms the row's relation mask with one or two
// ballots and packs the 2x2 pair nibbles of its 4 row pairs into the 4 bytes
// of one word H. The cell codes of plane pair (z-1, z) are the bytes of
// Hprev | H << 4, with the bytes of cells that do not exist cleared; all-one
// cells are counted per lane by a byte-wise test, mixed cells go to the
// workgroup's LDS histogram (one SDWA compare and one atomic per byte), and
// all-zero cells are what remains of the cell count.
//
// Fields whose levels all share one relation family (">=" / "<", or ">" /
// "<="; batchItemGray) take a faster form: the levels sorted, the pair bits
// and packing done once per Gray plane of a voxel's level count instead of
// once per level, and each level's code word recovered by a bitwise formula.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "mkf_device.hpp"
#include "mkf_internal.h"

namespace mkf {

constexpr int kBatchWaves = 8;
constexpr int kBatchTY = 4;  // row pairs per stack, per-level form
constexpr int kGrayTY = 4;   // row pairs per stack, Gray-plane form
constexpr int kMaxLevels = MKF_BATCH_MAX_LEVELS;
// Slab depth limit: one item adds at most 8 x 64 x 8 x 4096 cells, so the
// 32-bit LDS counters, flushed once they pass 2^31, cannot wrap.
constexpr int kMaxPlanePairs = 4096;

// Level families of the Gray-plane form (batchItemGray): the bit of a level
// is !(f < t) (kFamGE: relations ">=", and "<" as its complement) or f > t
// (kFamGT: ">", and "<=" as its complement). NaN takes the EQ bit: set for
// ">=" and "<=", clear for ">" and "<", so the complements hold for it too.
constexpr int kFamGE = 1, kFamGT = 2;

// Device descriptor of one field of a batch (or of one part of its levels).
struct BatchField {
  const void* p;
  unsigned long long* out;  // outLevels x 256 u64, added to
  long long sy, sz;
  int enc, X, Y, Z;
  int nLevels;           // levels tallied by this descriptor
  int outLevels;         // histograms at out (every part of a field shares them)
  int segs, stacks, TZ;  // segments per x-tile, row stacks per workgroup, plane pairs per slab
  int nXT, nYB, nZR;     // x-tiles, y-blocks, slabs
  int gray, fam;         // Gray-plane form: 4 or 8 padded levels and the family; 0: per-level form
  uint32_t comp;         // bit l: level l's histogram is stored complemented (bin c -> 255 - c)
  uint8_t perm[kMaxLevels];  // output histogram of level l
  double t[kMaxLevels];  // f32 fields: the threshold rounded to float (Gray form: sorted ascending)
  uint32_t m[kMaxLevels];
};

struct BatchItem {
  int field, xt, yb, zr;
};

// Shared state of a batch workgroup. hpk holds, per wave, level and lane,
// the previous plane's H word (four row-pair nibbles, one per byte); ones,
// per level and lane, the all-one cells counted so far (every wave adds its
// lane l's count to slot l).
struct BatchShared {
  uint32_t hist[kMaxLevels][256];
  uint32_t hpk[kBatchWaves][kMaxLevels][kWave];
  uint32_t ones[kMaxLevels][kWave];
  double xv[2][kBatchTY + 1][kBatchWaves];  // voxel at lane 0 of each row, per wave
  uint32_t msum[kMaxLevels];
  uint32_t cells;
  int item;
};

// Adds the workgroup's histograms of field fd to its outputs and clears them.
// Bin 0 of a level is the workgroup's cells minus its all-one and mixed cells.
__device__ void batchFlush(BatchShared& sh, const BatchField& fd) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x >> 6;
  const int nL = fd.nLevels;
  __syncthreads();
  // Wave w folds the all-one counts of levels w, w + 8, ... into bin 255.
  for (int l = w; l < nL; l += kBatchWaves) {
    const uint32_t n1 = waveSum(sh.ones[l][lane]);
    sh.ones[l][lane] = 0;
    if (lane == 0) sh.hist[l][255] += n1;
  }
  __syncthreads();
  for (int l = 0; l < nL; ++l) {
    uint32_t part = 0;
    for (int i = threadIdx.x; i < 256; i += blockDim.x) part += (i >= 1 && i <= 254) ? sh.hist[l][i] : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    if (lane == 0 && part) atomicAdd(&sh.msum[l], part);
  }
  __syncthreads();
  for (int l = 0; l < nL; ++l) {
    unsigned long long* o = fd.out + (size_t)fd.perm[l] * 256;
    const int flip = ((fd.comp >> l) & 1u) ? 255 : 0;
    for (int i = threadIdx.x; i < 256; i += blockDim.x) {
      const uint32_t v = (i == 0) ? sh.cells - sh.msum[l] - sh.hist[l][255] : sh.hist[l][i];
      if (v) atomicAdd(o + (i ^ flip), (unsigned long long)v);
    }
  }
  __syncthreads();
  for (int l = 0; l < nL; ++l)
    for (int i = threadIdx.x; i < 256; i += blockDim.x) sh.hist[l][i] = 0;
  if (threadIdx.x < kMaxLevels) sh.msum[threadIdx.x] = 0;
  if (threadIdx.x == 0) sh.cells = 0;
  __syncthreads();
}

// Pair bits of one row at relation M: bit 0 from the lane's voxel, bit 1
// from its right neighbour. The lane's bit enters as the carry of an add (or
// the borrow of a subtract, for relations that are the complement of a
// ballot: ">=", "<=", "=", all), so a row costs its compares, one v_cndmask
// and one v_addc / v_subb. NaN compares false both ways and so takes the EQ
// bit.
template <int M, typename T>
__device__ __forceinline__ uint32_t pairBits(T f, T fr, T t) {
  constexpr bool inv = (M == 2 || M == 3 || M == 6 || M == 7);
  uint64_t a = 0, b = 0;
  if constexpr (M == 1 || M == 6) {
    a = ballot(f < t);
    b = ballot(fr < t);
  } else if constexpr (M == 3 || M == 4) {
    a = ballot(f > t);
    b = ballot(fr > t);
  } else if constexpr (M == 2 || M == 5) {
    a = ballot(f < t) | ballot(f > t);
    b = ballot(fr < t) | ballot(fr > t);
  }
  if constexpr (inv) {
    return subLaneBit(laneBit(b) ? 1u : 3u, a);
  } else {
    return addLaneBit(laneBit(b) ? 2u : 0u, a);
  }
}

// One level of one z-step: the pair nibbles of the wave's 5 rows at level
// relation M, then (from the second plane on) the tally of its row pairs'
// cells: byte r of C is the code of the lane's cell in row pair r, cleared
// by kb where that cell does not exist (so it counts as neither all-one nor
// mixed). rgt[r] is the voxel right of each lane's voxel (lane 63: the halo
// voxel). All-one cells are counted per lane by a byte-wise test (added to
// the lane's slot of onesL), mixed cells go to histL with one SDWA compare
// and one LDS atomic per byte, and all-zero cells are what remains of the
// cell count.
template <int M, typename T>
__device__ __forceinline__ uint32_t levelStep(const T (&cur)[kBatchTY + 1], const T (&rgt)[kBatchTY + 1], T thr,
                                              uint32_t kb, bool count, uint32_t hp, uint32_t* histL,
                                              uint32_t* onesL) {
  uint32_t P[kBatchTY + 1];
#pragma unroll
  for (int r = 0; r <= kBatchTY; ++r) P[r] = pairBits<M>(cur[r], rgt[r], thr);
  const uint32_t lo = P[0] | (P[1] << 8) | (P[2] << 16) | (P[3] << 24);
  const uint32_t hi = (lo >> 8) | (P[4] << 24);
  const uint32_t H = lo | (hi << 2);
  if (count) {
    const uint32_t C = (hp | (H << 4)) & kb;
    const uint32_t X = mixedBits(C);
    __hip_atomic_fetch_add(onesL, (uint32_t)__builtin_popcount(allOneBytesX(C, X)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    Tally t;
    t.addMixed<0>(C, X, histL);
    t.addMixed<1>(C, X, histL);
    t.addMixed<2>(C, X, histL);
    t.addMixed<3>(C, X, histL);
  }
  return H;
}

template <typename T>
__device__ __forceinline__ uint32_t levelDispatch(uint32_t m, const T (&cur)[kBatchTY + 1],
                                                  const T (&rgt)[kBatchTY + 1], T thr, uint32_t kb, bool count,
                                                  uint32_t hp, uint32_t* histL, uint32_t* onesL) {
  switch (m) {
    case 1: return levelStep<1>(cur, rgt, thr, kb, count, hp, histL, onesL);
    case 2: return levelStep<2>(cur, rgt, thr, kb, count, hp, histL, onesL);
    case 3: return levelStep<3>(cur, rgt, thr, kb, count, hp, histL, onesL);
    case 4: return levelStep<4>(cur, rgt, thr, kb, count, hp, histL, onesL);
    case 5: return levelStep<5>(cur, rgt, thr, kb, count, hp, histL, onesL);
    case 6: return levelStep<6>(cur, rgt, thr, kb, count, hp, histL, onesL);
    case 7: return levelStep<7>(cur, rgt, thr, kb, count, hp, histL, onesL);
    default: return levelStep<0>(cur, rgt, thr, kb, count, hp, histL, onesL);
  }
}

// One work item: the wave's stack of rows over the item's slab, all levels.
template <typename T>
__device__ void batchItem(BatchShared& sh, const BatchField& fd, const BatchItem it) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int X = fd.X, Y = fd.Y;
  const int nL = fd.nLevels;
  const int segs = fd.segs;
  const int seg = w % segs;
  const int stack = w / segs;
  const bool active = stack < fd.stacks;
  const int x0 = (it.xt * segs + seg) * kWave;
  const int y0 = (it.yb * fd.stacks + (active ? stack : 0)) * kBatchTY;
  const int nY = active ? max(0, min(kBatchTY, Y - 1 - y0)) : 0;  // 0: the stack counts nothing
  const int yc = min(y0, Y - 1);
  const int z0 = it.zr * fd.TZ;
  const int nZ = min(fd.TZ, fd.Z - 1 - z0);
  HaloFrom halo = kHaloNone;
  if (nY > 0 && x0 + kWave < X) halo = (seg + 1 < segs) ? kHaloLds : kHaloMem;

  const int xb = min(x0, X - 1);
  const T* fw = (const T*)fd.p + xb;
  const long long sy = fd.sy, sz = fd.sz;
  const uint32_t lx = (uint32_t)(min(x0 + lane, X - 1) - xb);
  const int hxo = min(x0 + kWave, X - 1) - xb;
  const uint64_t valid = ballot(x0 + lane < X - 1);
  if (lane == 0 && nY > 0) atomicAdd(&sh.cells, (uint32_t)__builtin_popcountll(valid) * (uint32_t)(nY * nZ));
  // Bytes of the lane's cells that exist: row pairs 0 .. nY-1, valid lanes.
  const uint32_t kb = laneBit(valid) ? (nY >= 4 ? 0xFFFFFFFFu : (1u << (8 * nY)) - 1u) : 0u;

  // Lane l holds level l's threshold and relation; the level loop reads them
  // with readlane.
  const int ll = min(lane, nL - 1);
  const T tl = (T)fd.t[ll];
  const uint32_t ml = fd.m[ll];
  const bool last = lane == kWave - 1;

  const int nb = min(w + 1, kBatchWaves - 1);
  long long yo[kBatchTY + 1];
#pragma unroll
  for (int r = 0; r <= kBatchTY; ++r) yo[r] = (long long)(yc + min(r, nY)) * sy;
  T cur[kBatchTY + 1], nxt[kBatchTY + 1];
  T hc[kBatchTY + 1], hn[kBatchTY + 1];
#pragma unroll
  for (int r = 0; r <= kBatchTY; ++r) {
    cur[r] = globalLoad(fw, yo[r] + (long long)z0 * sz + lx);
    hc[r] = (halo == kHaloMem) ? uniformLoad(fw + (yo[r] + (long long)z0 * sz + hxo)) : T(0);
  }
  uint32_t* hpk = &sh.hpk[w][0][lane];
  for (int z = z0; z <= z0 + nZ; ++z) {
    const long long zon = (long long)min(z + 1, z0 + nZ) * sz;
#pragma unroll
    for (int r = 0; r <= kBatchTY; ++r) {
      nxt[r] = globalLoad(fw, yo[r] + zon + lx);
      hn[r] = (halo == kHaloMem) ? uniformLoad(fw + (yo[r] + zon + hxo)) : T(0);
    }
    const int par = z & 1;
    if (segs > 1) {
      if (lane == 0) {
#pragma unroll
        for (int r = 0; r <= kBatchTY; ++r) sh.xv[par][r][w] = (double)cur[r];
      }
      __syncthreads();
    }
    // Right neighbours: lane l+1's voxel, and the halo voxel for lane 63.
    T rgt[kBatchTY + 1];
#pragma unroll
    for (int r = 0; r <= kBatchTY; ++r) {
      const T hv = (halo == kHaloLds) ? (T)sh.xv[par][r][nb] : hc[r];
      const T n1 = laneNext(cur[r]);
      rgt[r] = last ? hv : n1;
    }
    const bool count = z > z0;
    uint32_t hpNext = hpk[0];
    for (int l = 0; l < nL; ++l) {
      const uint32_t hp = hpNext;
      if (l + 1 < nL) hpNext = hpk[(l + 1) * kWave];
      const T thr = readLane(tl, l);
      const uint32_t m = readLane(ml, l);
      hpk[l * kWave] = levelDispatch<T>(m, cur, rgt, thr, kb, count, hp, sh.hist[l], &sh.ones[l][lane]);
    }
#pragma unroll
    for (int r = 0; r <= kBatchTY; ++r) {
      cur[r] = nxt[r];
      hc[r] = hn[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Gray-plane form. With one family's levels sorted by threshold, a voxel's
// bits over levels 0..NL-1 form a thermometer: bit L is set iff L < r, r the
// number of set bits (0..NL). The wave ballots the NL level masks A_L of a
// row and keeps only the Gray code of r, in NP = log2(NL) + 1 plane masks:
//   NL = 8: g0 = A0^A2^A4^A6, g1 = A1^A5, g2 = A3, g3 = A7
//   NL = 4: g0 = A0^A2,       g1 = A1,    g2 = A3
// (scalar XORs of the ballots: A_L & ~A_(L+1) are disjoint since the A_L
// are nested). Pair bits, row packing and the previous plane's word are
// formed once per plane instead of once per level; the code word of level L
// (byte k: cell code of row pair k) is then a bitwise function of the plane
// words, applied to all 4 row pairs and all 8 corners at once (grayLevel).
// A segment covers 63 cells: lane 63 holds the voxel right of lane 62, so
// every right neighbour is bit l + 1 of the same ballot and the waves of a
// workgroup exchange nothing.
// ---------------------------------------------------------------------------
template <int NL>
constexpr int kGrayPlanes = NL == 8 ? 4 : 3;

template <int FAM, typename T>
__device__ __forceinline__ uint64_t famMask(T f, T t) {
  if constexpr (FAM == kFamGE)
    return ~ballot(f < t);
  else
    return ballot(f > t);
}

// Code word of level L (bit i of byte k set iff corner i of cell k has
// r > L) from the plane words c[j] (bit i of byte k: Gray bit j of corner
// i's r). Every form maps all-zero planes (cells that do not exist) to 0.
template <int NL, int L>
__device__ __forceinline__ uint32_t grayLevel(const uint32_t (&c)[kGrayPlanes<NL>]) {
  if constexpr (NL == 8) {
    if constexpr (L == 0) return c[0] | c[1] | c[2];
    if constexpr (L == 1) return c[1] | c[2];
    if constexpr (L == 2) return c[2] | (c[1] & ~c[0]);
    if constexpr (L == 3) return c[2];
    if constexpr (L == 4) return c[2] & ~(c[1] & ~c[0]);
    if constexpr (L == 5) return c[2] & ~c[1];
    if constexpr (L == 6) return c[2] & ~(c[1] | c[0]);
    return c[3];
  } else {
    if constexpr (L == 0) return c[0] | c[1];
    if constexpr (L == 1) return c[1];
    if constexpr (L == 2) return c[1] & ~c[0];
    return c[2];
  }
}

// Tally of one level's code word: all-one cells counted per lane, mixed
// cells to the level's LDS histogram.
__device__ __forceinline__ void grayTally(uint32_t W, uint32_t* hist, uint32_t& ones) {
  const uint32_t X = mixedBits(W);
  ones += (uint32_t)__builtin_popcount(allOneBytesX(W, X));
  Tally t;
  t.addMixed<0>(W, X, hist);
  t.addMixed<1>(W, X, hist);
  t.addMixed<2>(W, X, hist);
  t.addMixed<3>(W, X, hist);
}

// Levels L.. of one z-step; c[k][j]: plane j of row-pair word k.
template <int NL, int NW, int L>
__device__ __forceinline__ void grayLevels(const uint32_t (&c)[NW][kGrayPlanes<NL>], int nL, BatchShared& sh,
                                           uint32_t (&ones)[NL]) {
  if constexpr (L < NL) {
    if (L < nL) {
#pragma unroll
      for (int k = 0; k < NW; ++k) grayTally(grayLevel<NL, L>(c[k]), sh.hist[L], ones[L]);
      grayLevels<NL, NW, L + 1>(c, nL, sh, ones);
    }
  }
}

// One work item in the Gray-plane form: the wave's stack of TY + 1 rows (TY
// row pairs, TY / 4 code words per plane) over the item's slab, all levels.
template <typename T, int FAM, int NL, int TY>
__device__ void batchItemGray(BatchShared& sh, const BatchField& fd, const BatchItem it) {
  constexpr int NP = kGrayPlanes<NL>;
  constexpr int NW = TY / 4;         // row-pair words per plane
  constexpr int kCells = kWave - 1;  // cells per segment
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int X = fd.X, Y = fd.Y;
  const int nL = fd.nLevels;
  const int segs = fd.segs;
  const int seg = w % segs;
  const int stack = w / segs;
  const bool active = stack < fd.stacks;
  const int x0 = (it.xt * segs + seg) * kCells;
  const int y0 = (it.yb * fd.stacks + (active ? stack : 0)) * TY;
  const int nY = active ? max(0, min(TY, Y - 1 - y0)) : 0;  // 0: the stack counts nothing
  const int yc = min(y0, Y - 1);
  const int z0 = it.zr * fd.TZ;
  const int nZ = min(fd.TZ, fd.Z - 1 - z0);

  const int xb = min(x0, X - 1);
  const T* fw = (const T*)fd.p + xb;
  const long long sy = fd.sy, sz = fd.sz;
  const uint32_t lx = (uint32_t)(min(x0 + lane, X - 1) - xb);
  const uint64_t valid = ballot(lane < kCells && x0 + lane < X - 1);
  if (lane == 0 && nY > 0) atomicAdd(&sh.cells, (uint32_t)__builtin_popcountll(valid) * (uint32_t)(nY * nZ));
  // Bytes of cells that exist, per row-pair word.
  uint32_t kb[NW];
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const int n = min(max(nY - 4 * k, 0), 4);
    kb[k] = laneBit(valid) ? (n >= 4 ? 0xFFFFFFFFu : (1u << (8 * n)) - 1u) : 0u;
  }

  T thr[NL];
#pragma unroll
  for (int l = 0; l < NL; ++l) thr[l] = (T)fd.t[l];

  long long yo[TY + 1];
#pragma unroll
  for (int r = 0; r <= TY; ++r) yo[r] = (long long)(yc + min(r, nY)) * sy;
  T cur[TY + 1], nxt[TY + 1];
#pragma unroll
  for (int r = 0; r <= TY; ++r) cur[r] = globalLoad(fw, yo[r] + (long long)z0 * sz + lx);
  uint32_t hp[NW][NP] = {};
  uint32_t ones[NL];
#pragma unroll
  for (int l = 0; l < NL; ++l) ones[l] = 0;
  for (int z = z0; z <= z0 + nZ; ++z) {
    const long long zon = (long long)min(z + 1, z0 + nZ) * sz;
#pragma unroll
    for (int r = 0; r <= TY; ++r) nxt[r] = globalLoad(fw, yo[r] + zon + lx);
    // Pair bits of every plane and row: bit 0 the lane's voxel, bit 1 its
    // right neighbour (lane l + 1).
    uint32_t P[NP][TY + 1];
#pragma unroll
    for (int r = 0; r <= TY; ++r) {
      uint64_t A[NL];
#pragma unroll
      for (int l = 0; l < NL; ++l) A[l] = famMask<FAM>(cur[r], thr[l]);
      uint64_t g[NP];
      if constexpr (NL == 8) {
        g[0] = A[0] ^ A[2] ^ A[4] ^ A[6];
        g[1] = A[1] ^ A[5];
        g[2] = A[3];
        g[3] = A[7];
      } else {
        g[0] = A[0] ^ A[2];
        g[1] = A[1];
        g[2] = A[3];
      }
#pragma unroll
      for (int j = 0; j < NP; ++j) P[j][r] = addLaneBit(laneBit(g[j] >> 1) ? 2u : 0u, g[j]);
    }
    uint32_t c[NW][NP];
#pragma unroll
    for (int k = 0; k < NW; ++k) {
#pragma unroll
      for (int j = 0; j < NP; ++j) {
        const uint32_t* q = &P[j][4 * k];
        const uint32_t lo = q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24);
        const uint32_t hi = (lo >> 8) | (q[4] << 24);
        const uint32_t H = lo | (hi << 2);
        c[k][j] = (hp[k][j] | (H << 4)) & kb[k];
        hp[k][j] = H;
      }
    }
    if (z > z0) grayLevels<NL, NW, 0>(c, nL, sh, ones);
#pragma unroll
    for (int r = 0; r <= TY; ++r) cur[r] = nxt[r];
  }
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    if (l < nL) __hip_atomic_fetch_add(&sh.ones[l][lane], ones[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

__global__ void kBatchReset(unsigned int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

// Copy of n 16-byte words by the compute units (either side may be mapped
// pinned host memory), so small transfers on a compute stream never queue
// behind bulk DMA transfers on other streams.
__global__ void kCopyWords(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void kZeroList(unsigned long long* const* outs, const int* counts, int n) {
  for (int f = blockIdx.x; f < n; f += gridDim.x)
    for (int i = threadIdx.x; i < counts[f]; i += blockDim.x) outs[f][i] = 0ull;
}

__global__ __launch_bounds__(kBatchWaves* kWave) void kBatch(const BatchField* __restrict__ fields,
                                                              const BatchItem* __restrict__ items, int nItems,
                                                              unsigned int* __restrict__ counter) {
  __shared__ BatchShared sh;
  for (int l = 0; l < kMaxLevels; ++l) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) sh.hist[l][i] = 0;
    if (threadIdx.x < kWave) sh.ones[l][threadIdx.x] = 0;
  }
  if (threadIdx.x < kMaxLevels) sh.msum[threadIdx.x] = 0;
  if (threadIdx.x == 0) sh.cells = 0;
  int curField = -1;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) sh.item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(sh.item);
    const int f = i < nItems ? items[i].field : -1;
    const bool full = sh.cells > 0x80000000u;
    if ((f != curField || full) && curField >= 0) batchFlush(sh, fields[curField]);
    if (i >= nItems) break;
    curField = f;
    const BatchField& fd = fields[f];
    const BatchItem it = items[i];
    const bool f32 = fd.enc == ENC_F32;
    if (fd.gray == 8) {
      if (fd.fam == kFamGE) {
        if (f32) batchItemGray<float, kFamGE, 8, kGrayTY>(sh, fd, it);
        else batchItemGray<double, kFamGE, 8, kGrayTY>(sh, fd, it);
      } else {
        if (f32) batchItemGray<float, kFamGT, 8, kGrayTY>(sh, fd, it);
        else batchItemGray<double, kFamGT, 8, kGrayTY>(sh, fd, it);
      }
    } else if (fd.gray == 4) {
      if (fd.fam == kFamGE) {
        if (f32) batchItemGray<float, kFamGE, 4, kGrayTY>(sh, fd, it);
        else batchItemGray<double, kFamGE, 4, kGrayTY>(sh, fd, it);
      } else {
        if (f32) batchItemGray<float, kFamGT, 4, kGrayTY>(sh, fd, it);
        else batchItemGray<double, kFamGT, 4, kGrayTY>(sh, fd, it);
      }
    } else if (f32) {
      batchItem<float>(sh, fd, it);
    } else {
      batchItem<double>(sh, fd, it);
    }
  }
}

}  // namespace mkf

using namespace mkf;

// ---------------------------------------------------------------------------
// Host side: field descriptions, work items, the reusable launcher and the
// graph-captured plan
// ---------------------------------------------------------------------------
namespace mkf {
namespace {

int bfail(int code, const char* msg) { return setError(code, "batch: %s", msg); }

long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

bool shapeOk(const MkfBatchShape& s) {
  return s.planePairs >= 1 && s.planePairs <= kMaxPlanePairs &&
         (s.segments == 1 || s.segments == 2 || s.segments == 4 || s.segments == 8);
}

MkfBatchShape normalise(const MkfBatchShape* s) {
  MkfBatchShape r{32, 8};
  if (s) {
    r.planePairs = s->planePairs == 0 ? 32 : s->planePairs;
    r.segments = s->segments == 0 ? 8 : s->segments;
  }
  return r;
}

// Family of the Gray-plane form for a field's levels, 0 if they do not
// qualify: every relation ">=" or "<" (kFamGE), or every one ">" or "<="
// (kFamGT), "none" (0) and "all" (7) joining either; no threshold NaN.
int grayFamily(const MkfBatchField& src) {
  int fam = 0;
  for (int l = 0; l < src.nLevels; ++l) {
    const uint32_t m = src.pM[l].m & 7u;
    const double t = src.pM[l].t[0];
    if (t != t || m == 2 || m == 5) return 0;
    const int f = (m == 6 || m == 1) ? kFamGE : (m == 4 || m == 3) ? kFamGT : 0;
    if (f != 0) {
      if (fam != 0 && fam != f) return 0;
      fam = f;
    }
  }
  return fam != 0 ? fam : kFamGE;
}

// Validates field `index` and appends its device descriptors (without
// shape). A field with 3 or more levels of one Gray family gets the Gray-plane
// form: its levels sorted by threshold, each as the family's base relation
// with a complement flag ("<" = not ">=", "<=" = not ">"; "all" and "none"
// as levels at an infinite threshold), in parts of at most 8 levels that
// read the field once each. Other fields keep their levels in order (the
// per-level form).
int describeField(const MkfBatchField& src, int index, std::vector<BatchField>& out) {
  FieldDef d[3];
  const int nd = copyNonZeroDef(d, src.def, 3);
  if (!src.pF || !src.pBPFD || (src.enc != ENC_F32 && src.enc != ENC_F64) || src.nLevels < 1 ||
      src.nLevels > kMaxLevels || !src.pM || nd <= 0 || src.def[0] < 0 || src.def[1] < 0 || src.def[2] < 0)
    return setError(MKF_ERR_ARG, "batch: invalid description of field %d", index);
  BatchField b;
  memset(&b, 0, sizeof(b));
  b.p = src.pF;
  b.out = (unsigned long long*)src.pBPFD;
  b.enc = src.enc;
  b.X = d[0];
  b.Y = d[1];
  b.Z = d[2];
  b.sy = d[0];
  b.sz = (long long)d[0] * d[1];
  b.outLevels = src.nLevels;
  // f32 fields compare against the threshold rounded to float.
  auto thr = [&](int l) { return (src.enc == ENC_F32) ? (double)(float)src.pM[l].t[0] : src.pM[l].t[0]; };
  const int fam = src.nLevels >= 3 ? grayFamily(src) : 0;
  if (fam == 0) {
    b.nLevels = src.nLevels;
    for (int l = 0; l < src.nLevels; ++l) {
      b.t[l] = thr(l);
      b.m[l] = src.pM[l].m & 7u;
      b.perm[l] = (uint8_t)l;
    }
    out.push_back(b);
    return MKF_OK;
  }
  struct Lv {
    double t;
    int idx;
    bool comp;
  };
  std::vector<Lv> lv;
  const double inf = __builtin_inf();
  for (int l = 0; l < src.nLevels; ++l) {
    const uint32_t m = src.pM[l].m & 7u;
    if (fam == kFamGE) {
      // bit = !(f < t); "all" is ">= -inf", "none" its complement
      if (m == 6 || m == 1) lv.push_back({thr(l), l, m == 1});
      else lv.push_back({-inf, l, m == 0});
    } else {
      // bit = f > t; "none" is "> +inf", "all" its complement
      if (m == 4 || m == 3) lv.push_back({thr(l), l, m == 3});
      else lv.push_back({inf, l, m == 7});
    }
  }
  std::stable_sort(lv.begin(), lv.end(), [](const Lv& a, const Lv& c) { return a.t < c.t; });
  const int n = src.nLevels;
  const int parts = (n + 7) / 8;
  for (int p = 0, first = 0; p < parts; ++p) {
    const int cnt = n / parts + (p < n % parts ? 1 : 0);
    BatchField q = b;
    q.nLevels = cnt;
    q.gray = cnt <= 4 ? 4 : 8;
    q.fam = fam;
    for (int l = 0; l < q.gray; ++l) {
      if (l < cnt) {
        const Lv& e = lv[first + l];
        q.t[l] = e.t;
        q.perm[l] = (uint8_t)e.idx;
        if (e.comp) q.comp |= 1u << l;
      } else {
        q.t[l] = inf;  // above every level: keeps the bits a thermometer
      }
    }
    out.push_back(q);
    first += cnt;
  }
  return MKF_OK;
}

// Cells per 64-lane segment: the Gray-plane form gives lane 63 to the right
// neighbour of lane 62.
int segCells(const BatchField& b) { return b.gray ? kWave - 1 : kWave; }

// Tile shape of one field: the largest power-of-two x-tile up to the shape's
// limit that leaves the fewest empty segments at the ends of the rows.
void shapeField(BatchField& b, const MkfBatchShape& s) {
  const int nSeg = (int)std::max<long long>(1, cdiv(b.X - 1, segCells(b)));
  int best = 1;
  long long bestWaste = 1LL << 40;
  for (int g = s.segments; g >= 1; g >>= 1) {
    const long long waste = cdiv(nSeg, g) * g - nSeg;
    if (waste < bestWaste) {
      bestWaste = waste;
      best = g;
    }
  }
  b.segs = best;
  b.stacks = kBatchWaves / best;
  b.TZ = s.planePairs;
  b.nXT = (int)cdiv(nSeg, best);
  b.nYB = (int)cdiv(cdiv(std::max(b.Y - 1, 1), b.gray ? kGrayTY : kBatchTY), b.stacks);
  b.nZR = (int)cdiv(std::max(b.Z - 1, 1), b.TZ);
}

// Items sorted by field, so a workgroup flushes when it moves to the next
// field. Fields without cells (a dim of 1) have no items.
void buildItems(const std::vector<BatchField>& fs, std::vector<BatchItem>& items) {
  items.clear();
  for (int f = 0; f < (int)fs.size(); ++f) {
    const BatchField& b = fs[f];
    if (b.X < 2 || b.Y < 2 || b.Z < 2) continue;
    for (int zr = 0; zr < b.nZR; ++zr)
      for (int yb = 0; yb < b.nYB; ++yb)
        for (int xt = 0; xt < b.nXT; ++xt) items.push_back(BatchItem{f, xt, yb, zr});
  }
}

// Persistent grid: every CU holds as many workgroups as fit.
int persistentGrid() {
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  int perCU = 1;
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kBatch, kBatchWaves * kWave, 0);
  return std::max(1, cus * std::max(1, perCU));
}

// Queue reset (a kernel, so a captured graph has no memset node), optional
// output clearing, then the persistent kernel.
void launchAll(const BatchField* dFields, const BatchItem* dItems, int nItems, unsigned int* dCounter,
               unsigned long long* const* dOuts, const int* dCounts, int nFields, bool zero, int grid,
               hipStream_t s) {
  hipLaunchKernelGGL(kBatchReset, dim3(1), dim3(64), 0, s, dCounter);
  if (zero)
    hipLaunchKernelGGL(kZeroList, dim3(std::min<int>(1024, nFields)), dim3(256), 0, s, dOuts, dCounts, nFields);
  if (nItems > 0)
    hipLaunchKernelGGL(kBatch, dim3(grid), dim3(kBatchWaves * kWave), 0, s, dFields, dItems, nItems, dCounter);
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

hipError_t copyWords(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes % 16 != 0) return hipErrorInvalidValue;
  const size_t n = bytes / 16;
  if (n == 0) return hipSuccess;
  const int blocks = (int)std::min<size_t>(256, (n + 255) / 256);
  hipLaunchKernelGGL(kCopyWords, dim3(blocks), dim3(256), 0, s, (uint4*)dst, (const uint4*)src, n);
  return hipGetLastError();
}

// Launcher for a set of fields that changes from launch to launch (the
// serving loop): descriptors and items go through pinned staging and one
// copy per launch; buffers grow as needed and are kept.
struct BatchLauncher {
  char* hStage = nullptr;
  char* hStageDev = nullptr;  // device address of hStage (mapped pinned memory)
  char* dBuf = nullptr;
  size_t cap = 0;
  unsigned int* dCounter = nullptr;
  hipEvent_t staged = nullptr;
  int grid = 0;
  std::vector<BatchField> fs;
  std::vector<BatchItem> items;
};

BatchLauncher* batchLauncherCreate() {
  BatchLauncher* L = new BatchLauncher();
  if (hipMalloc((void**)&L->dCounter, sizeof(unsigned int)) != hipSuccess ||
      hipEventCreateWithFlags(&L->staged, hipEventDisableTiming) != hipSuccess) {
    bfail(MKF_ERR_MEM, "launcher allocation");
    batchLauncherDestroy(L);
    return nullptr;
  }
  L->grid = persistentGrid();
  return L;
}

void batchLauncherDestroy(BatchLauncher* L) {
  if (!L) return;
  if (L->staged) {
    (void)hipEventSynchronize(L->staged);
    (void)hipEventDestroy(L->staged);
  }
  if (L->hStage) (void)hipHostFree(L->hStage);
  if (L->dBuf) (void)hipFree(L->dBuf);
  if (L->dCounter) (void)hipFree(L->dCounter);
  delete L;
}

int batchLauncherRun(BatchLauncher* L, const MkfBatchField* fields, int n, const MkfBatchShape* shape, int flags,
                     hipStream_t s) {
  if (!L || !fields || n <= 0) return bfail(MKF_ERR_ARG, "no fields");
  const MkfBatchShape sh = normalise(shape);
  if (!shapeOk(sh)) return bfail(MKF_ERR_ARG, "shape: planePairs in 1..4096 and segments in {1, 2, 4, 8}");
  L->fs.clear();
  for (int f = 0; f < n; ++f) {
    const int rc = describeField(fields[f], f, L->fs);
    if (rc != MKF_OK) return rc;
  }
  for (auto& b : L->fs) shapeField(b, sh);
  n = (int)L->fs.size();  // descriptors from here on (a field may have several)
  buildItems(L->fs, L->items);
  if (L->items.size() > (size_t)0x7fffffff - 1) return bfail(MKF_ERR_UNSUPPORTED, "too many work items");
  const size_t oF = 0, oI = align256(sizeof(BatchField) * n);
  const size_t oO = oI + align256(sizeof(BatchItem) * L->items.size());
  const size_t oC = oO + align256(sizeof(void*) * n);
  const size_t total = oC + align256(sizeof(int) * n);
  // The previous launch's staging copy must be done before the buffer is
  // rewritten (or replaced).
  if (hipEventSynchronize(L->staged) != hipSuccess) return bfail(MKF_ERR_HIP, "staging wait");
  if (total > L->cap) {
    const size_t cap = std::max(total, L->cap + L->cap / 2);
    if (L->hStage) (void)hipHostFree(L->hStage);
    if (L->dBuf) (void)hipFree(L->dBuf);
    L->hStage = nullptr;
    L->dBuf = nullptr;
    L->cap = 0;
    if (hipHostMalloc((void**)&L->hStage, cap, hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void**)&L->hStageDev, L->hStage, 0) != hipSuccess ||
        hipMalloc((void**)&L->dBuf, cap) != hipSuccess)
      return bfail(MKF_ERR_MEM, "launcher buffers");
    L->cap = cap;
  }
  memcpy(L->hStage + oF, L->fs.data(), sizeof(BatchField) * n);
  if (!L->items.empty()) memcpy(L->hStage + oI, L->items.data(), sizeof(BatchItem) * L->items.size());
  unsigned long long** outs = (unsigned long long**)(L->hStage + oO);
  int* counts = (int*)(L->hStage + oC);
  for (int f = 0; f < n; ++f) {
    outs[f] = L->fs[f].out;
    counts[f] = 256 * L->fs[f].outLevels;
  }
  if (copyWords(L->dBuf, L->hStageDev, total, s) != hipSuccess || hipEventRecord(L->staged, s) != hipSuccess)
    return bfail(MKF_ERR_HIP, "descriptor upload");
  launchAll((const BatchField*)(L->dBuf + oF), (const BatchItem*)(L->dBuf + oI), (int)L->items.size(), L->dCounter,
            (unsigned long long* const*)(L->dBuf + oO), (const int*)(L->dBuf + oC), n,
            (flags & MKF_BATCH_ZERO_OUTPUT) != 0, L->grid, s);
  return hipGetLastError() == hipSuccess ? MKF_OK : bfail(MKF_ERR_HIP, "batch launch");
}

}  // namespace mkf

// Graph-captured plan of a fixed set of fields.
struct MkfBatch {
  std::vector<BatchField> host;
  BatchField* dFields = nullptr;
  BatchItem* dItems = nullptr;
  size_t itemCap = 0;
  unsigned int* dCounter = nullptr;
  unsigned long long** dOuts = nullptr;
  int* dCounts = nullptr;
  int nItems = 0;
  int grid = 0;
  MkfBatchShape shape{32, 8};
  int flags = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  // Set when a re-plan failed and the previous plan could not be restored:
  // the device queue and the graph may then disagree, so runs are refused.
  bool broken = false;
};

namespace {

int upload(MkfBatch* B, const MkfBatchShape& s, hipStream_t st) {
  if (faultDue("batch.upload")) return bfail(MKF_ERR_MEM, "work queue allocation (injected fault)");
  for (auto& b : B->host) shapeField(b, s);
  std::vector<BatchItem> items;
  buildItems(B->host, items);
  if (items.size() > (size_t)0x7fffffff - 1) return bfail(MKF_ERR_UNSUPPORTED, "too many work items");
  if (items.size() > B->itemCap) {
    if (B->dItems) (void)hipFree(B->dItems);
    B->dItems = nullptr;
    B->itemCap = 0;
    if (hipMalloc((void**)&B->dItems, sizeof(BatchItem) * items.size()) != hipSuccess)
      return bfail(MKF_ERR_MEM, "work queue allocation");
    B->itemCap = items.size();
  }
  B->nItems = (int)items.size();
  B->shape = s;
  if (hipMemcpyAsync(B->dFields, B->host.data(), sizeof(BatchField) * B->host.size(), hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      (!items.empty() && hipMemcpyAsync(B->dItems, items.data(), sizeof(BatchItem) * items.size(),
                                        hipMemcpyHostToDevice, st) != hipSuccess))
    return bfail(MKF_ERR_HIP, "upload of the batch description");
  return hipStreamSynchronize(st) == hipSuccess ? MKF_OK : bfail(MKF_ERR_HIP, "upload sync");
}

hipError_t record(MkfBatch* B, hipStream_t s) {
  launchAll(B->dFields, B->dItems, B->nItems, B->dCounter, B->dOuts, B->dCounts, (int)B->host.size(),
            (B->flags & MKF_BATCH_ZERO_OUTPUT) != 0, B->grid, s);
  return hipGetLastError();
}

int capture(MkfBatch* B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
  if (faultDue("batch.capture")) return bfail(MKF_ERR_HIP, "stream capture (injected fault)");
  if (hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return bfail(MKF_ERR_HIP, "stream capture begin");
  const hipError_t e = record(B, B->cap);
  if (hipStreamEndCapture(B->cap, &B->graph) != hipSuccess || e != hipSuccess)
    return bfail(MKF_ERR_HIP, "stream capture");
  if (hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0) != hipSuccess)
    return bfail(MKF_ERR_HIP, "graph instantiation");
  return MKF_OK;
}

}  // namespace

extern "C" {

MkfBatch* mkfBatchCreate(const MkfBatchField* fields, int nFields, const MkfBatchShape* shape, int flags,
                         void* stream) {
  if (!fields || nFields <= 0) {
    bfail(MKF_ERR_ARG, "no fields");
    return nullptr;
  }
  const MkfBatchShape s = normalise(shape);
  if (!shapeOk(s)) {
    bfail(MKF_ERR_ARG, "shape: planePairs in 1..4096 and segments in {1, 2, 4, 8}");
    return nullptr;
  }
  std::vector<BatchField> host;
  for (int f = 0; f < nFields; ++f)
    if (describeField(fields[f], f, host) != MKF_OK) return nullptr;
  const int nDesc = (int)host.size();  // a field may have several descriptors
  std::vector<unsigned long long*> outs;
  std::vector<int> counts;
  for (const auto& b : host) {
    outs.push_back(b.out);
    counts.push_back(256 * b.outLevels);
  }
  MkfBatch* B = new MkfBatch();
  B->flags = flags;
  B->host.swap(host);
  hipStream_t st = (hipStream_t)stream;
  B->grid = persistentGrid();
  if (hipMalloc((void**)&B->dFields, sizeof(BatchField) * nDesc) != hipSuccess ||
      hipMalloc((void**)&B->dCounter, sizeof(unsigned int)) != hipSuccess ||
      hipMalloc((void**)&B->dOuts, sizeof(void*) * nDesc) != hipSuccess ||
      hipMalloc((void**)&B->dCounts, sizeof(int) * nDesc) != hipSuccess ||
      hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking) != hipSuccess) {
    bfail(MKF_ERR_MEM, "batch allocation");
    mkfBatchDestroy(B);
    return nullptr;
  }
  int rc = MKF_OK;
  if (hipMemcpyAsync(B->dOuts, outs.data(), sizeof(void*) * nDesc, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(B->dCounts, counts.data(), sizeof(int) * nDesc, hipMemcpyHostToDevice, st) != hipSuccess)
    rc = bfail(MKF_ERR_HIP, "upload of the output table");
  if (rc == MKF_OK) rc = upload(B, s, st);
  if (rc == MKF_OK) rc = capture(B);
  if (rc != MKF_OK) {
    mkfBatchDestroy(B);
    return nullptr;
  }
  return B;
}

int mkfBatchRun(MkfBatch* B, void* stream) {
  if (B && B->broken) return bfail(MKF_ERR_ARG, "unusable after a failed re-plan; destroy it");
  if (!B || !B->exec) return bfail(MKF_ERR_ARG, "no batch");
  const hipError_t e = hipGraphLaunch(B->exec, (hipStream_t)stream);
  return e == hipSuccess ? MKF_OK : bfail(MKF_ERR_HIP, "graph launch");
}

int mkfBatchRunDirect(MkfBatch* B, void* stream) {
  if (!B) return bfail(MKF_ERR_ARG, "no batch");
  if (B->broken) return bfail(MKF_ERR_ARG, "unusable after a failed re-plan; destroy it");
  return record(B, (hipStream_t)stream) == hipSuccess ? MKF_OK : bfail(MKF_ERR_HIP, "batch launch");
}

int mkfBatchAutotune(MkfBatch* B, const MkfBatchShape* candidates, int nCandidates, int reps,
                     MkfBatchShape* chosen, void* stream) {
  if (!B || !candidates || nCandidates <= 0) return bfail(MKF_ERR_ARG, "no candidates");
  if (B->broken) return bfail(MKF_ERR_ARG, "unusable after a failed re-plan; destroy it");
  hipStream_t st = (hipStream_t)stream;
  const MkfBatchShape prev = B->shape;
  // Point every field at a scratch histogram while tuning (the parts of a
  // field, which are consecutive and share their output, share it too).
  size_t total = 0;
  for (auto& b : B->host) total += (size_t)b.outLevels * 256;
  unsigned long long* scratch = nullptr;
  if (hipMalloc((void**)&scratch, total * 8) != hipSuccess) return bfail(MKF_ERR_MEM, "tuning scratch");
  std::vector<unsigned long long*> saved;
  size_t off = 0;
  for (size_t f = 0; f < B->host.size(); ++f) {
    BatchField& b = B->host[f];
    saved.push_back(b.out);
    if (f > 0 && b.out == saved[f - 1]) {
      b.out = B->host[f - 1].out;
      continue;
    }
    b.out = scratch + off;
    off += (size_t)b.outLevels * 256;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  MkfBatchShape best{0, 0};
  float bestMs = 1e30f;
  const int saveFlags = B->flags;
  B->flags &= ~MKF_BATCH_ZERO_OUTPUT;
  int rc = MKF_OK;
  for (int c = 0; c < nCandidates && rc == MKF_OK; ++c) {
    const MkfBatchShape s = normalise(&candidates[c]);
    if (!shapeOk(s)) continue;
    if ((rc = upload(B, s, st)) != MKF_OK) break;
    (void)record(B, st);  // warm-up
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < std::max(1, reps); ++r) (void)record(B, st);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess) {
      rc = bfail(MKF_ERR_HIP, "tuning run");
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < bestMs) {
      bestMs = ms;
      best = s;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (size_t f = 0; f < B->host.size(); ++f) B->host[f].out = saved[f];
  B->flags = saveFlags;
  (void)hipFree(scratch);
  if (rc == MKF_OK && best.planePairs <= 0) rc = bfail(MKF_ERR_ARG, "no usable candidate");
  // Plan the fastest shape, or, after an error, the previous one again: the
  // tuning runs replaced the device queue (and may have reallocated it), so
  // the queue and the graph are rebuilt either way. If that fails too, the
  // batch refuses further runs.
  int rp = upload(B, rc == MKF_OK ? best : prev, st);
  if (rp == MKF_OK) rp = capture(B);
  if (rp != MKF_OK) {
    B->broken = true;
    return rc != MKF_OK ? rc : rp;
  }
  if (rc != MKF_OK) return rc;
  if (chosen) *chosen = best;
  return MKF_OK;
}

int mkfBatchInfo(const MkfBatch* B, int* nItems, MkfBatchShape* shape, int* grid) {
  if (!B) return bfail(MKF_ERR_ARG, "no batch");
  if (nItems) *nItems = B->nItems;
  if (shape) *shape = B->shape;
  if (grid) *grid = B->grid;
  return MKF_OK;
}

void mkfBatchDestroy(MkfBatch* B) {
  if (!B) return;
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  if (B->cap) (void)hipStreamDestroy(B->cap);
  if (B->dFields) (void)hipFree(B->dFields);
  if (B->dItems) (void)hipFree(B->dItems);
  if (B->dCounter) (void)hipFree(B->dCounter);
  if (B->dOuts) (void)hipFree(B->dOuts);
  if (B->dCounts) (void)hipFree(B->dCounts);
  delete B;
}

}  // extern "C"
