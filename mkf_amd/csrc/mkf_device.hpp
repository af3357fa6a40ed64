// mkf_device.hpp - wave64 device building blocks for gfx950 (CDNA4).
//
// Layout conventions (shared by every kernel):
//  * A wave owns a 64-wide x-segment: lane l <-> x = x0 + l, x0 = 64 * seg.
//  * A row (y, z) of a segment is thresholded by ONE v_cmp; its ballot is the
//    64-bit PBM mask R of the segment, bit l = voxel x0 + l, which is exactly
//    the packed-map bit order (bit x & 31 of 32-bit word x >> 5, words little
//    endian), so R is stored as two consecutive 32-bit PBM words.
//  * E = R >> 1 with the first voxel of the next segment shifted into bit 63
//    holds bit x + 1 for lane x, so lane l's pair value (bit x, bit x+1) is
//    read straight out of the two SGPR masks by v_cndmask (inverse ballot).
//  * 2x2x2 cell code (src/bpfdCUDA.cuh:45-61, src/mkfACC.c:78-91):
//      bit i = P(x + (i & 1), y + ((i >> 1) & 1), z + (i >> 2)).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mkf {

constexpr int kWave = 64;

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Threshold relation of a whole wave as a 64-bit lane mask
// (src/binMapCUDA.cuh:141-145): with d = 1 - (f < t) + (f > t) the bit is
// (m >> d) & 1, so an unordered compare (NaN) selects the EQ bit. Every mask
// is built from the ballots of the two ORDERED compares f < t and f > t:
// hipcc lowers a ballot of (f < t || f > t) to v_cmp_neq, which is true for
// NaN, so composite predicates are never balloted directly. Lanes must all
// be active (inactive lanes would read as EQ).
template <int M, typename T>
__device__ __forceinline__ uint64_t relMask(T f, T t) {
  if constexpr (M == 0) {
    return 0ull;
  } else if constexpr (M == 1) {
    return ballot(f < t);
  } else if constexpr (M == 2) {
    return ~(ballot(f < t) | ballot(f > t));
  } else if constexpr (M == 3) {
    return ~ballot(f > t);
  } else if constexpr (M == 4) {
    return ballot(f > t);
  } else if constexpr (M == 5) {
    return ballot(f < t) | ballot(f > t);
  } else if constexpr (M == 6) {
    return ~ballot(f < t);
  } else {
    return ~0ull;
  }
}

// The same with the relation mask known only at run time.
template <typename T>
__device__ __forceinline__ uint64_t relMaskDyn(T f, T t, uint32_t m) {
  const uint64_t lt = ballot(f < t);
  const uint64_t gt = ballot(f > t);
  const uint64_t eq = ~(lt | gt);
  return ((m & 1u) ? lt : 0ull) | ((m & 2u) ? eq : 0ull) | ((m & 4u) ? gt : 0ull);
}

// Lane l reads bit l of a wave-uniform 64-bit mask (one v_cndmask).
__device__ __forceinline__ bool laneBit(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

__device__ __forceinline__ uint64_t lowBits(int n) {
  return n >= 64 ? ~0ull : (n <= 0 ? 0ull : ((1ull << n) - 1ull));
}

// Load of a wave-uniform address through the scalar unit (s_load): the
// address is made uniform with readfirstlane and read through the constant
// address space. Used only for read-only inputs.
template <typename T>
__device__ __forceinline__ T uniformLoad(const T* p) {
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  typedef const __attribute__((address_space(4))) T* ConstPtr;
  return *(ConstPtr)((((uint64_t)hi) << 32) | lo);
}

// Pair value for lane l from the segment masks: bit x and bit x + 1.
__device__ __forceinline__ uint32_t pairBits(uint64_t R, uint64_t E) {
  return (laneBit(R) ? 1u : 0u) | (laneBit(E) ? 2u : 0u);
}

// Rows of a scalar field thresholded on the fly (fused path).
template <typename T, int M>
struct FieldRows {
  const T* f;
  T thr;
  long long sy, sz;  // element strides of y and z (x stride is 1)
  int xc;            // this lane's x, clamped into the row
  int haloX;         // x of the next segment's first voxel, clamped into the row
  bool halo;         // whether that voxel exists
  uint64_t tail;     // voxels of this segment that exist

  struct Raw {
    T v, h;
  };

  // Every address is inside the row, so loads need no branch.
  __device__ __forceinline__ Raw load(int y, int z) const {
    const long long b = (long long)y * sy + (long long)z * sz;
    Raw r;
    r.v = f[b + xc];
    r.h = uniformLoad(f + b + haloX);
    return r;
  }

  // Returns the lane's pair value; R receives the PBM mask of the segment.
  // The halo value is wave-uniform, so any lane of its mask holds its bit.
  __device__ __forceinline__ uint32_t pair(const Raw& r, uint64_t& R) const {
    R = relMask<M>(r.v, thr) & tail;
    const uint64_t hb = halo ? ((relMask<M>(r.h, thr) & 1ull) << 63) : 0ull;
    return pairBits(R, (R >> 1) | hb);
  }
};

// Rows of an existing packed binary map (BPFD-only path).
struct PbmRows {
  const uint32_t* w;
  uint32_t rowWS, planeWS;
  uint32_t wi0;   // first word of the segment (2 * seg)
  uint32_t o1;    // offset of the segment's second word, 0 if outside the row
  uint32_t o2;    // offset of the next segment's first word, 0 if outside
  bool halo;      // next segment's first bit exists
  uint64_t tail;  // bits of this segment inside the row

  struct Raw {
    uint32_t w0, w1, w2;
  };

  // Offsets outside the row are replaced by 0 and masked off after the load.
  __device__ __forceinline__ Raw load(int y, int z) const {
    const uint32_t* p = w + (size_t)y * rowWS + (size_t)z * planeWS + wi0;
    Raw r;
    r.w0 = uniformLoad(p);
    r.w1 = uniformLoad(p + o1);
    r.w2 = uniformLoad(p + o2);
    return r;
  }

  __device__ __forceinline__ uint32_t pair(const Raw& r, uint64_t& R) const {
    R = ((((uint64_t)r.w1) << 32) | r.w0) & tail;
    const uint64_t hb = halo ? (((uint64_t)(r.w2 & 1u)) << 63) : 0ull;
    return pairBits(R, (R >> 1) | hb);
  }
};

// Per-wave tally: all-zero and all-one cells are counted from ballots in
// scalar registers; only mixed cells touch the wave's LDS sub-histogram.
struct Tally {
  uint32_t c0 = 0, c255 = 0;

  __device__ __forceinline__ void add(uint32_t code, bool laneCell, uint64_t vmask, uint32_t* hist) {
    c255 += __builtin_popcountll(ballot(code == 255u) & vmask);
    c0 += __builtin_popcountll(ballot(code == 0u) & vmask);
    if (laneCell && (code - 1u) < 254u) {
      __hip_atomic_fetch_add(hist + code, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
  }
};

// Walks row pairs y0 .. y0+nY-1 of plane pairs z0 .. z0+nT-1 for one
// x-segment: each step loads row y of planes z0..z0+nT (nT + 1 loads, the
// next step's loads are issued before this step's cells are counted), and
// counts the cells between rows y-1 and y.
template <int TP, class Src>
__device__ __forceinline__ void walkSegment(const Src& src, int y0, int nY, int z0, int nT, bool laneCell,
                                            uint64_t vmask, uint32_t* hist, Tally& tally) {
  // Planes beyond z0 + nT re-read the last plane (always in bounds); their
  // cells are never counted.
  int zt[TP + 1];
#pragma unroll
  for (int t = 0; t <= TP; ++t) zt[t] = z0 + min(t, nT);
  typename Src::Raw cur[TP + 1], nxt[TP + 1];
#pragma unroll
  for (int t = 0; t <= TP; ++t) cur[t] = src.load(y0, zt[t]);
  uint32_t qprev[TP + 1];
  for (int y = y0; y <= y0 + nY; ++y) {
    const int yn = min(y + 1, y0 + nY);
#pragma unroll
    for (int t = 0; t <= TP; ++t) nxt[t] = src.load(yn, zt[t]);
    uint32_t q[TP + 1];
#pragma unroll
    for (int t = 0; t <= TP; ++t) {
      uint64_t R;
      q[t] = src.pair(cur[t], R);
    }
    if (y > y0) {
      uint32_t h[TP + 1];
#pragma unroll
      for (int t = 0; t <= TP; ++t) h[t] = qprev[t] | (q[t] << 2);
#pragma unroll
      for (int t = 0; t < TP; ++t) {
        if (t < nT) tally.add(h[t] | (h[t + 1] << 4), laneCell, vmask, hist);
      }
    }
#pragma unroll
    for (int t = 0; t <= TP; ++t) {
      qprev[t] = q[t];
      cur[t] = nxt[t];
    }
  }
}

}  // namespace mkf
