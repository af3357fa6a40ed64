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
//    That one bit comes from the neighbouring wave through LDS.
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

// Rows of a scalar field thresholded on the fly (fused path). Addresses are
// a wave-uniform row pointer (scalar registers) plus a 32-bit lane offset,
// so loads take the global_load saddr form. The wave's row mask R comes from
// the compare ballots; the bit right of the segment (the next segment's
// first voxel) is read by the wave itself only when that segment belongs to
// another workgroup, else it is exchanged through LDS by the walker.
template <typename T, int M>
struct FieldRows {
  const T* fw;       // field + xb, xb = min(x0, X - 1) (uniform)
  T thr;
  long long sy, sz;  // element strides of y and z (x stride is 1)
  uint32_t lx;       // lane offset from xb, clamped into the row
  int hxo;           // offset from xb of the next segment's first voxel, clamped
  uint64_t tail;     // voxels of this segment that exist

  struct Raw {
    T v;
  };

  // Uniform row pointer + zero-extended 32-bit lane offset: the saddr form.
  __device__ __forceinline__ Raw load(long long off) const {
    const T* row = fw + off;
    return Raw{row[lx]};
  }
  __device__ __forceinline__ uint64_t mask(const Raw& r) const { return relMask<M>(r.v, thr) & tail; }
  // Bit of the voxel right of the segment, read from memory (wave-uniform).
  __device__ __forceinline__ uint32_t haloBit(long long off) const {
    return (uint32_t)(relMask<M>(uniformLoad(fw + (off + hxo)), thr) & 1ull);
  }
};

// Rows of an existing packed binary map (BPFD-only path): the segment's two
// words are read by the scalar unit. Offsets are in words (sy = rowWS,
// sz = planeWS).
struct PbmRows {
  const uint32_t* ws;  // PBM + first word of the segment, clamped into the row
  long long sy, sz;
  uint32_t o1;         // offset of the segment's second word, 0 if outside the row
  uint32_t hw;         // offset of the word holding the next segment's first bit
  uint64_t tail;       // bits of this segment inside the row

  struct Raw {
    uint32_t w0, w1;
  };

  __device__ __forceinline__ Raw load(long long off) const {
    const uint32_t* p = ws + off;
    return Raw{uniformLoad(p), uniformLoad(p + o1)};
  }
  __device__ __forceinline__ uint64_t mask(const Raw& r) const {
    return ((((uint64_t)r.w1) << 32) | r.w0) & tail;
  }
  __device__ __forceinline__ uint32_t haloBit(long long off) const { return uniformLoad(ws + (off + hw)) & 1u; }
};

// Per-wave tally: all-one cells and mixed cells are counted from ballots in
// scalar registers; only mixed cells touch the wave's LDS sub-histogram. The
// all-zero count is what remains of the wave's cell count.
struct Tally {
  uint32_t c255 = 0, cmix = 0;

  __device__ __forceinline__ void add(uint32_t code, uint64_t vmask, uint32_t* hist) {
    c255 += __builtin_popcountll(ballot(code == 255u) & vmask);
    const uint64_t mixed = ballot((code - 1u) < 254u) & vmask;
    cmix += __builtin_popcountll(mixed);
    if (laneBit(mixed)) __hip_atomic_fetch_add(hist + code, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  }
};

// How a wave obtains the bit right of its segment.
enum HaloFrom : int { kHaloNone = 0, kHaloLds = 1, kHaloMem = 2 };

// Row-stack walker. The workgroup's waves sit side by side along x; each
// holds rows y0 .. y0+nY (nY <= TY) and marches z from z0 to z0+nZ, so one
// z-step of the workgroup reads nY+1 consecutive rows of one plane. At every
// step each wave publishes the low word of its row masks in LDS (double
// buffered by z parity), one barrier later it reads its right neighbour's
// word for the bit x+1 of its last lane, forms lane l's pair value
// (bit x, bit x+1) and the cell codes between this plane and the previous:
//   hc[r] = q(y0+r) | q(y0+r+1) << 2;  code = hc_prev[r] | hc[r] << 4.
// Every wave of the workgroup runs the same number of steps (barriers).
template <int TY, class Src>
__device__ __forceinline__ void walkStack(const Src& src, uint32_t (*xch)[TY + 1][16], int w, HaloFrom halo,
                                          int y0, int nY, int z0, int nZ, uint64_t vmask, uint32_t* hist,
                                          Tally& tally) {
  const int lane = threadIdx.x & (kWave - 1);
  // The neighbour slot read when the halo comes from LDS (clamped so the
  // read stays inside the array for the workgroup's last wave).
  const int nb = min(w + 1, 15);
  const uint32_t ldsKeep = (halo == kHaloLds) ? 1u : 0u;
  long long yo[TY + 1];
#pragma unroll
  for (int r = 0; r <= TY; ++r) yo[r] = (long long)(y0 + min(r, nY)) * src.sy;
  typename Src::Raw cur[TY + 1], nxt[TY + 1];
  {
    const long long zo = (long long)z0 * src.sz;
#pragma unroll
    for (int r = 0; r <= TY; ++r) cur[r] = src.load(yo[r] + zo);
  }
  uint32_t hp[TY];
  for (int z = z0; z <= z0 + nZ; ++z) {
    const long long zon = (long long)min(z + 1, z0 + nZ) * src.sz;
#pragma unroll
    for (int r = 0; r <= TY; ++r) nxt[r] = src.load(yo[r] + zon);
    const int par = z & 1;
    uint64_t R[TY + 1];
#pragma unroll
    for (int r = 0; r <= TY; ++r) {
      R[r] = src.mask(cur[r]);
      if (lane == 0) xch[par][r][w] = (uint32_t)R[r];
    }
    __syncthreads();
    uint32_t hb[TY + 1];
    if (halo == kHaloMem) {
      const long long zo = (long long)z * src.sz;
#pragma unroll
      for (int r = 0; r <= TY; ++r) hb[r] = src.haloBit(yo[r] + zo);
    } else {
#pragma unroll
      for (int r = 0; r <= TY; ++r) hb[r] = __builtin_amdgcn_readfirstlane(xch[par][r][nb]) & ldsKeep;
    }
    uint32_t q[TY + 1];
#pragma unroll
    for (int r = 0; r <= TY; ++r) q[r] = pairBits(R[r], (R[r] >> 1) | ((uint64_t)hb[r] << 63));
    uint32_t hc[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) hc[r] = q[r] | (q[r + 1] << 2);
    if (z > z0) {
#pragma unroll
      for (int r = 0; r < TY; ++r) {
        if (r < nY) tally.add(hp[r] | (hc[r] << 4), vmask, hist);
      }
    }
#pragma unroll
    for (int r = 0; r < TY; ++r) hp[r] = hc[r];
#pragma unroll
    for (int r = 0; r <= TY; ++r) cur[r] = nxt[r];
  }
}

}  // namespace mkf
