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

#include <type_traits>
#include <utility>

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

// x + (bit l of mask) for lane l: one v_addc_co_u32 with the mask as its
// carry-in (register operands only; the carry-out is discarded).
__device__ __forceinline__ uint32_t addLaneBit(uint32_t x, uint64_t mask) {
  uint32_t r;
  uint64_t co;
  asm("v_addc_co_u32_e64 %0, %1, %2, 0, %3" : "=v"(r), "=s"(co) : "v"(x), "s"(mask));
  return r;
}
// x - (bit l of mask) for lane l: one v_subb_co_u32 with the mask as borrow.
__device__ __forceinline__ uint32_t subLaneBit(uint32_t x, uint64_t mask) {
  uint32_t r;
  uint64_t co;
  asm("v_subb_co_u32_e64 %0, %1, %2, 0, %3" : "=v"(r), "=s"(co) : "v"(x), "s"(mask));
  return r;
}

// 2x + (bit l of mask) for lane l: one v_addc_co_u32 (x + x + carry).
__device__ __forceinline__ uint32_t dblAddLaneBit(uint32_t x, uint64_t mask) {
  uint32_t r;
  uint64_t co;
  asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(co) : "v"(x), "s"(mask));
  return r;
}

// v | (the value of v in lane l ^ 1), (l ^ 2), lane l + 4 or l + 8 of the
// same 16-lane row (0 where that lane is outside the row): DPP quad_perm
// [1,0,3,2], quad_perm [2,3,0,1], row_shl:4, row_shl:8.
__device__ __forceinline__ uint32_t orQuadXor1(uint32_t v) {
  return v | (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t orQuadXor2(uint32_t v) {
  return v | (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t orRowShl4(uint32_t v) {
  return v | (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x104, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t orRowShl8(uint32_t v) {
  return v | (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x108, 0xF, 0xF, true);
}
// v | (the value of v in lane 7 - i of the same 8-lane half-row), and in
// lane 15 - i of the same 16-lane row: DPP row_half_mirror, row_mirror. After
// orQuadXor1 and orQuadXor2 every lane of a quad holds the quad's OR; one
// half-mirror step then gives every lane its 8-lane group's OR, and one more
// mirror step its 16-lane row's OR.
__device__ __forceinline__ uint32_t orHalfMirror(uint32_t v) {
  return v | (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t orRowMirror(uint32_t v) {
  return v | (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, true);
}

// Lane l receives lane l+1's value through DPP (wave_shl:1, one v_mov per
// dword); lane 63, which has no source, keeps its own value.
__device__ __forceinline__ uint32_t laneNext(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x130, 0xF, 0xF, false);
}
__device__ __forceinline__ float laneNext(float v) { return __uint_as_float(laneNext(__float_as_uint(v))); }
__device__ __forceinline__ double laneNext(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint64_t lo = laneNext((uint32_t)b), hi = laneNext((uint32_t)(b >> 32));
  return __longlong_as_double((long long)(lo | (hi << 32)));
}

// Wave-uniform read of lane `l` of v (v_readlane; l is uniform).
__device__ __forceinline__ uint32_t readLane(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ float readLane(float v, int l) { return __uint_as_float(readLane(__float_as_uint(v), l)); }
__device__ __forceinline__ double readLane(double v, int l) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint64_t lo = readLane((uint32_t)b, l), hi = readLane((uint32_t)(b >> 32), l);
  return __longlong_as_double((long long)(lo | (hi << 32)));
}

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

// Per-lane load through a pointer known to be in global memory (a pointer
// read from a descriptor is generic and would otherwise load with flat_load).
template <typename T>
__device__ __forceinline__ T globalLoad(const T* p, long long i) {
  typedef const __attribute__((address_space(1))) T* GlobalPtr;
  return ((GlobalPtr)p)[i];
}
// Load of a 32-bit word at a per-lane byte offset from a wave-uniform base,
// as a raw buffer load: the base goes into a buffer descriptor in scalar
// registers (readfirstlane makes its uniformity explicit) and the offset is
// a 32-bit vector operand, so no 64-bit address is formed per load. Offsets
// must be below 2^31; the descriptor's range check returns 0 beyond that.
__device__ __forceinline__ uint32_t bufferLoadWord(const uint32_t* base, uint32_t byteOff) {
  const uint64_t a = reinterpret_cast<uint64_t>(base);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc((void*)((((uint64_t)hi) << 32) | lo), 0, 0x7FFFFFFF, 0x00020000);
  return __builtin_amdgcn_raw_buffer_load_b32(r, (int)byteOff, 0, 0);
}

// Byte-packed pair words. A lane holds V consecutive voxels (V = 1, 2 or 4);
// byte k of its pair word P holds (bit of voxel k, bit of voxel k+1) in bits
// 8k and 8k+1. Then, for rows r, r+1 and planes z-1, z, all V cell codes of a
// lane come from two word operations:
//   Hc = P(r) | P(r+1) << 2          (one plane, two rows: bits 0..3 per byte)
//   C  = Hc(z-1) | Hc(z) << 4        (byte k = code of the lane's cell k)
// which is the reference's code bit order (src/bpfdCUDA.cuh:45-61).
template <int V>
__device__ __forceinline__ uint32_t packPairs(const uint32_t (&b)[V], uint32_t right) {
  uint32_t P = 0;
#pragma unroll
  for (int k = 0; k < V; ++k) {
    P |= b[k] << (8 * k);
    if (k > 0) P |= b[k] << (8 * (k - 1) + 1);
  }
  return P | (right << (8 * (V - 1) + 1));
}

// Bit position of the neighbour-supplied bit in a pair word.
template <int V>
constexpr int kRightBit = 8 * (V - 1) + 1;

template <typename T, int V>
struct VecLoad;
template <typename T>
struct VecLoad<T, 1> {
  typedef T type;
};
template <>
struct VecLoad<float, 2> {
  typedef float type __attribute__((ext_vector_type(2), aligned(4)));
};
template <>
struct VecLoad<float, 4> {
  typedef float type __attribute__((ext_vector_type(4), aligned(4)));
};
template <>
struct VecLoad<double, 2> {
  typedef double type __attribute__((ext_vector_type(2), aligned(8)));
};

// Kernel argument: up to 4 fields of identical dims and element strides,
// thresholded as their sum in table order.
constexpr int kMaxFieldsDev = 4;  // fields summed at most (BMCU_FIELD_MAX)

template <typename T>
struct FieldArgs {
  const T* p[4];
  int nF;
  long long sx, sy, sz;
};

// Rows of a scalar field thresholded on the fly (fused path). Lane l holds
// voxels x0 + V*l + k. Addresses are a wave-uniform row pointer plus a lane
// offset; a wave whose voxels all exist reads them with one vector load per
// row (4 floats or 2 doubles = 16 bytes), a partial wave clamps each voxel.
// The general form (G: several fields, or an x stride) gathers each voxel
// and sums the fields in table order (src/binMapCUDA.cuh:196-201); the
// plain form compiles to the single-field, unit-stride loads only. The bit
// right of the wave's last voxel comes from the next wave through LDS
// (walker), or from memory when that wave belongs to another workgroup.
template <typename T, int M, int V, bool G = false, bool PART = false>
struct FieldRows {
  static constexpr bool kSelfHalo = false;
  static constexpr int kM = M;
  const T* fw[4];    // fields + xb * sx, xb = min(x0, X - 1) (uniform)
  int nF;            // fields summed (1..4)
  long long sx;      // element stride of x
  T thr;
  long long sy, sz;  // element strides of y and z
  uint32_t lx;       // lane offset (in voxels) of voxel 0 from xb, clamped into the row
  int last;          // X - 1 - xb: clamp for partial waves
  bool vec;          // unit x stride and every voxel of every lane exists
  int hxo;           // offset (in voxels) from xb of the next wave's first voxel, clamped
  uint32_t keep;     // pair-word bytes of cells that exist (per lane)
  // PART (partial waves: unit stride, V > 1, some voxels past the row's
  // end): the lane's V voxels are read by one vector load from vs (offset
  // from xb, ending inside the row; X >= V is checked by the host) and
  // shifted back by shv (0..V-1 for lanes whose first voxel is in the row)
  // where they are used. The kernel compiles the walk once per form, so
  // whole waves carry no partial-wave code.
  int vs;
  uint32_t shv;

  struct Raw {
    T v[V];
  };

  template <bool NT = false>
  __device__ __forceinline__ Raw loadOne(const T* row) const {
    Raw r;
    if (!G && !PART && vec) {
      typedef typename VecLoad<T, V>::type VT;
      typedef const __attribute__((address_space(1))) VT* GVT;
      VT x;
      if constexpr (NT)
        x = __builtin_nontemporal_load((GVT)(row + lx));
      else
        x = *(const VT*)(row + lx);
      if constexpr (V == 1) {
        r.v[0] = x;
      } else {
#pragma unroll
        for (int k = 0; k < V; ++k) r.v[k] = x[k];
      }
    } else if constexpr (PART) {
      typedef typename VecLoad<T, V>::type VT;
      typedef const __attribute__((address_space(1))) VT* GVT;
      VT x;
      if constexpr (NT)
        x = __builtin_nontemporal_load((GVT)(row + vs));
      else
        x = *(const VT*)(row + vs);
#pragma unroll
      for (int k = 0; k < V; ++k) r.v[k] = x[k];
    } else {
#pragma unroll
      for (int k = 0; k < V; ++k) r.v[k] = row[G ? (long long)min((int)lx + k, last) * sx : min((int)lx + k, last)];
    }
    return r;
  }
  // The lane's voxels in place: partial waves shift the vector back by shv.
  __device__ __forceinline__ Raw placed(const Raw& r) const {
    if constexpr (PART) {
      {
        Raw o = r;
        const bool s1 = (shv & 1u) != 0;
        if constexpr (V == 4) {
          const bool s2 = (shv & 2u) != 0;
          const T t0 = s1 ? r.v[1] : r.v[0], t1 = s1 ? r.v[2] : r.v[1], t2 = s1 ? r.v[3] : r.v[2];
          o.v[0] = s2 ? t2 : t0;
          o.v[1] = s2 ? r.v[3] : t1;
          o.v[2] = t2;
        } else {
          o.v[0] = s1 ? r.v[1] : r.v[0];
        }
        return o;
      }
    }
    return r;
  }
  template <bool NT = false>
  __device__ __forceinline__ Raw load(long long off) const {
    Raw r = loadOne<NT>(fw[0] + off);
    if constexpr (G) {
      // unrolled with guards so the pointers stay in registers
#pragma unroll
      for (int f = 1; f < kMaxFieldsDev; ++f) {
        if (f < nF) {
          const Raw q = loadOne(fw[f] + off);
#pragma unroll
          for (int k = 0; k < V; ++k) r.v[k] += q.v[k];
        }
      }
    }
    return r;
  }
  // Pair word without the neighbour bit of lane 63; R0 receives the wave
  // mask of voxel 0 (the right neighbours of the lanes' last voxels). Bits of
  // voxels that do not exist (clamped loads) only reach cells that do not
  // exist, whose bytes `keep` clears.
  __device__ __forceinline__ uint32_t pairs(const Raw& r0, uint64_t& R0) const {
    const Raw r = placed(r0);
    uint32_t b[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const uint64_t mk = relMask<M>(r.v[k], thr);
      if (k == 0) R0 = mk;
      b[k] = laneBit(mk) ? 1u : 0u;
    }
    const uint32_t right = laneBit(R0 >> 1) ? 1u : 0u;
    return packPairs<V>(b, right) & keep;
  }
  // Value and bit of the voxel right of the wave, read from memory
  // (wave-uniform).
  __device__ __forceinline__ T haloVal(long long off) const {
    const long long o = off + (G ? (long long)hxo * sx : (long long)hxo);
    T h = uniformLoad(fw[0] + o);
    if constexpr (G) {
#pragma unroll
      for (int f = 1; f < kMaxFieldsDev; ++f)
        if (f < nF) h += uniformLoad(fw[f] + o);
    }
    return h;
  }
  __device__ __forceinline__ uint32_t haloBit(long long off) const {
    return (uint32_t)(relMask<M>(haloVal(off), thr) & 1ull);
  }
  // Phase moments of the lane's voxels selected by own (bit k: voxel k).
  template <class Acc>
  __device__ __forceinline__ void moments(const Raw& r0, uint32_t own, Acc& acc) const {
    const Raw r = placed(r0);
#pragma unroll
    for (int k = 0; k < V; ++k) acc.add(r.v[k], laneBit(relMask<M>(r.v[k], thr)), ((own >> k) & 1u) != 0);
  }
};

// Rows of an existing packed binary map (BPFD-only path, one bit per lane):
// the wave's two words are read by the scalar unit. Offsets are in words
// (sy = rowWS, sz = planeWS).
struct PbmRows {
  static constexpr bool kSelfHalo = false;
  const uint32_t* ws;  // PBM + first word of the segment, clamped into the row
  long long sy, sz;
  uint32_t o1;         // offset of the segment's second word, 0 if outside the row
  uint32_t hw;         // offset of the word holding the next segment's first bit
  uint32_t keep;       // pair-word bits of cells that exist (per lane)

  struct Raw {
    uint32_t w0, w1;
  };

  template <bool NT = false>
  __device__ __forceinline__ Raw load(long long off) const {
    const uint32_t* p = ws + off;
    return Raw{uniformLoad(p), uniformLoad(p + o1)};
  }
  // Bits beyond the row (tail bits of the last word, or a clamped word) only
  // reach cells that do not exist, whose bits `keep` clears.
  __device__ __forceinline__ uint32_t pairs(const Raw& r, uint64_t& R0) const {
    R0 = (((uint64_t)r.w1) << 32) | r.w0;
    const uint32_t b[1] = {laneBit(R0) ? 1u : 0u};
    return packPairs<1>(b, laneBit(R0 >> 1) ? 1u : 0u) & keep;
  }
  __device__ __forceinline__ uint32_t haloBit(long long off) const { return uniformLoad(ws + (off + hw)) & 1u; }
};

// Rows of an existing packed binary map, 4 voxels per lane: lane l holds
// voxels x0 + 4l .. 4l+3, nibble l & 7 of word x0/32 + (l >> 3). It reads
// that word and the next one, so the bit right of its last voxel -- lane 63's
// too -- is in its own registers and no halo is exchanged. The five bits
// b0..b4 become the pair word (byte k = b_k | b_(k+1) << 1) by one multiply:
// v * 0x204081 places copies of v at bits 0, 7, 14 and 21 without overlap, so
// byte k holds bits k and k+1 of v at its bits 0 and 1.
struct PbmRows4 {
  static constexpr bool kSelfHalo = true;
  const uint32_t* row;  // PBM (offsets below are from the row start)
  long long sy, sz;     // word strides of y and z
  uint32_t b0, b1;      // byte offsets of the lane's word and the next, clamped into the row
  uint32_t sh;          // 4 * (lane & 7)
  uint32_t keep;        // pair-word bytes of cells that exist (per lane)

  struct Raw {
    uint32_t lo, hi;
  };

  template <bool NT = false>
  __device__ __forceinline__ Raw load(long long off) const {
    const uint32_t* rp = row + off;  // wave-uniform
    return Raw{bufferLoadWord(rp, b0), bufferLoadWord(rp, b1)};
  }
  __device__ __forceinline__ uint32_t pairs(const Raw& r, uint64_t& R0) const {
    R0 = 0;
    const uint32_t v = __builtin_amdgcn_alignbit(r.hi, r.lo, sh) & 31u;
    return (v * 0x204081u) & 0x03030303u & keep;
  }
  __device__ __forceinline__ uint32_t haloBit(long long) const { return 0u; }
};

// PBM rows of a narrow map (rows of at most 4 words) read in row groups: the
// wave's lanes form groups of 8 GW lanes, each group reading GW words of its
// own stack of rows (a per-lane row offset folded into b0 and b1), so one
// wave walks several row stacks side by side. kp[r] holds the pair-word
// bytes of the lane's cells of row pair r that exist and that this group
// owns; the walk applies it to the row pair's word, so pairs() needs no mask.
template <int TY>
struct PbmRowsR : PbmRows4 {
  static constexpr bool kPairKeep = true;
  uint32_t kp[TY];

  __device__ __forceinline__ uint32_t pairs(const Raw& r, uint64_t& R0) const {
    R0 = 0;
    const uint32_t v = __builtin_amdgcn_alignbit(r.hi, r.lo, sh) & 31u;
    return (v * 0x204081u) & 0x03030303u;
  }
  __device__ __forceinline__ uint32_t pairKeep(int r) const { return kp[r]; }
};

// Whether a row source masks its cells per row pair (kPairKeep) rather than
// per row.
template <class S, class = void>
struct HasPairKeep : std::false_type {};
template <class S>
struct HasPairKeep<S, std::void_t<decltype(S::kPairKeep)>> : std::bool_constant<S::kPairKeep> {};

// Per-lane phase moments: sums in double of the values and their squares per
// phase (map bit 0 or 1) and the counts. Selects, not differences, split the
// phases, so a NaN stays in its own phase. The selects act on the value
// before its conversion to double (one 32-bit select each for a float).
struct MomAcc {
  double s0 = 0, q0 = 0, s1 = 0, q1 = 0;
  uint32_t n1 = 0, nAll = 0;

  template <typename T>
  __device__ __forceinline__ void add(T v, bool in, bool own) {
    const double fi = (double)((own && in) ? v : T(0));
    const double fo = (double)((own && !in) ? v : T(0));
    s1 += fi;
    q1 = __builtin_fma(fi, fi, q1);
    s0 += fo;
    q0 = __builtin_fma(fo, fo, q0);
    n1 += (own && in) ? 1u : 0u;
    nAll += own ? 1u : 0u;
  }
};

// v in the lanes of mask, +0 in the others: one v_cndmask on the float. (A
// select written in C++ lets the compiler convert first and select both
// halves of the double.)
__device__ __forceinline__ float selectLanes(float v, uint64_t mask) {
  float r;
  asm("v_cndmask_b32_e64 %0, 0, %1, %2" : "=v"(r) : "v"(v), "s"(mask));
  return r;
}

// +0 in the lanes of mask, v in the others.
__device__ __forceinline__ float selectOtherLanes(float v, uint64_t mask) {
  float r;
  asm("v_cndmask_b32_e64 %0, %1, 0, %2" : "=v"(r) : "v"(v), "s"(mask));
  return r;
}

// Phase moments with the phases and the owned voxels given as wave masks:
// the sums per lane in double, the counts per wave (wave-uniform, so in
// scalar registers).
struct MomWave {
  double s0 = 0, q0 = 0, s1 = 0, q1 = 0;
  uint64_t n1 = 0, nAll = 0;

  __device__ __forceinline__ void sums(double fi, double fo) {
    s1 += fi;
    q1 = __builtin_fma(fi, fi, q1);
    s0 += fo;
    q0 = __builtin_fma(fo, fo, q0);
  }
  // One voxel per lane, every lane's voxel counted: in = the lanes whose
  // voxel is in phase 1 (the caller counts them).
  __device__ __forceinline__ void addSplit(float v, uint64_t in) {
    sums((double)selectLanes(v, in), (double)selectOtherLanes(v, in));
  }
  // The same for the voxels of the lanes in own only.
  __device__ __forceinline__ void addOwned(float v, uint64_t in, uint64_t own) {
    sums((double)selectLanes(v, in & own), (double)selectLanes(v, ~in & own));
  }
  // As addOwned; returns the number of lanes counted in phase 1.
  __device__ __forceinline__ uint32_t add(float v, uint64_t in, uint64_t own) {
    addOwned(v, in, own);
    return (uint32_t)__builtin_popcountll(in & own);
  }
};

constexpr int kMomN = 6;  // {sum0, sumsq0, sum1, sumsq1, n0, n1}

// Writes the workgroup's moments to part[blockIdx.x * kMomN ..] after a
// fixed-order reduction (lanes by xor shuffles, then waves in order).
// Every thread of the workgroup must call it. n0 and n1 are further counts
// of the calling lane (a wave's counts come in on one lane).
__device__ __forceinline__ void momBlockPartial(const MomAcc& a, double* __restrict__ part, double n0 = 0.0,
                                                double n1 = 0.0) {
  __shared__ double red[16][kMomN];
  double t[kMomN] = {a.s0, a.q0, a.s1, a.q1, (double)(a.nAll - a.n1) + n0, (double)a.n1 + n1};
#pragma unroll
  for (int j = 0; j < kMomN; ++j) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) t[j] += __shfl_xor(t[j], o);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
    for (int j = 0; j < kMomN; ++j) red[w][j] = t[j];
  }
  __syncthreads();
  if (threadIdx.x < kMomN) {
    const int nw = blockDim.x >> 6;
    double s = red[0][threadIdx.x];
    for (int k = 1; k < nw; ++k) s += red[k][threadIdx.x];
    part[(size_t)blockIdx.x * kMomN + threadIdx.x] = s;
  }
}

// Moments off: every call compiles away.
struct NoMom {
  static constexpr bool kOn = false;
};
struct OnMom {
  static constexpr bool kOn = true;
  MomAcc acc;
  uint32_t vkeep;     // the lane's voxels inside the field (bit k: voxel k)
  bool lastRow;       // the stack's row nY is the field's last row
  bool lastPlane;     // the slab's last plane is the field's last plane
  bool haloLast;      // the voxel right of the wave is the field's last column, read from memory
};

// Sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t waveSum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Byte-wise arithmetic on a word of 4 cell codes (byte k = code of cell k).
// Byte k of the result is (byte k of C) - 1 mod 256, with no borrow between
// bytes: setting each byte's top bit keeps every byte's subtraction inside
// it, and the XOR restores the top bit.
__device__ __forceinline__ uint32_t bytesMinusOne(uint32_t C) {
  constexpr uint32_t H = 0x80808080u;
  return ((C | H) - 0x01010101u) ^ (~C & H);
}
// Bit 7 of byte k is set iff byte k of C is 0xFF (the low 7 bits + 1 carry
// into bit 7, which stays inside the byte).
__device__ __forceinline__ uint32_t allOneBytes(uint32_t C) {
  return ((C & 0x7F7F7F7Fu) + 0x01010101u) & C & 0x80808080u;
}

// Byte k of the result is nonzero iff byte k of C is mixed (neither 0 nor
// 255): it holds the XOR of each pair of neighbouring bits of byte k, which
// are all equal only for 0x00 and 0xFF (bit 7 is cleared: it would pair with
// the next byte's bit 0).
__device__ __forceinline__ uint32_t mixedBits(uint32_t C) { return (C ^ (C >> 1)) & 0x7F7F7F7Fu; }
// Bit 7 of byte k is set iff byte k of C is 0xFF, given X = mixedBits(C):
// X + 0x7F7F7F7F sets bit 7 of exactly the bytes with X nonzero (no byte
// carries: X is at most 0x7F per byte), and a byte that is not mixed is 0xFF
// iff its bit 7 is set.
__device__ __forceinline__ uint32_t allOneBytesX(uint32_t C, uint32_t X) {
  return C & ~(X + 0x7F7F7F7Fu) & 0x80808080u;
}

// The lanes whose byte K of D is below 254, i.e. whose cell K is mixed when D
// = bytesMinusOne(C) (codes 0 and 255 become 255 and 254): one compare with
// an SDWA byte select.
template <int K>
__device__ __forceinline__ uint64_t mixedLanes(uint32_t D) {
  uint64_t s;
  const uint32_t lim = 254u;
  if constexpr (K == 0) asm("v_cmp_gt_u32_sdwa %0, %1, %2 src0_sel:DWORD src1_sel:BYTE_0" : "=s"(s) : "v"(lim), "v"(D));
  if constexpr (K == 1) asm("v_cmp_gt_u32_sdwa %0, %1, %2 src0_sel:DWORD src1_sel:BYTE_1" : "=s"(s) : "v"(lim), "v"(D));
  if constexpr (K == 2) asm("v_cmp_gt_u32_sdwa %0, %1, %2 src0_sel:DWORD src1_sel:BYTE_2" : "=s"(s) : "v"(lim), "v"(D));
  if constexpr (K == 3) asm("v_cmp_gt_u32_sdwa %0, %1, %2 src0_sel:DWORD src1_sel:BYTE_3" : "=s"(s) : "v"(lim), "v"(D));
  return s;
}
// 4 x byte K of C (the byte offset of its histogram bin): one shift with an
// SDWA byte select.
template <int K>
__device__ __forceinline__ uint32_t byteTimes4(uint32_t C) {
  uint32_t a;
  if constexpr (K == 0)
    asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0" : "=v"(a) : "v"(C));
  if constexpr (K == 1)
    asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_1" : "=v"(a) : "v"(C));
  if constexpr (K == 2)
    asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_2" : "=v"(a) : "v"(C));
  if constexpr (K == 3)
    asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_3" : "=v"(a) : "v"(C));
  return a;
}

// The lanes whose byte K of M is nonzero: one compare with an SDWA byte
// select.
template <int K>
__device__ __forceinline__ uint64_t nonZeroByteLanes(uint32_t M) {
  uint64_t s;
  const uint32_t zero = 0u;
  if constexpr (K == 0) asm("v_cmp_lt_u32_sdwa %0, %1, %2 src0_sel:DWORD src1_sel:BYTE_0" : "=s"(s) : "v"(zero), "v"(M));
  if constexpr (K == 1) asm("v_cmp_lt_u32_sdwa %0, %1, %2 src0_sel:DWORD src1_sel:BYTE_1" : "=s"(s) : "v"(zero), "v"(M));
  if constexpr (K == 2) asm("v_cmp_lt_u32_sdwa %0, %1, %2 src0_sel:DWORD src1_sel:BYTE_2" : "=s"(s) : "v"(zero), "v"(M));
  if constexpr (K == 3) asm("v_cmp_lt_u32_sdwa %0, %1, %2 src0_sel:DWORD src1_sel:BYTE_3" : "=s"(s) : "v"(zero), "v"(M));
  return s;
}

// Per-plane tests of a row-pair word Hc (byte k: the 4-bit half code of
// cell k in one plane, bits 4..7 zero), made once per plane and used for
// both plane pairs the plane belongs to:
//   full  = Hc + 0x01010101: bit 4 of byte k set iff nibble k is 15
//   some  = (Hc + 0x0F0F0F0F) & 0x10101010: bit 4 set iff nibble k is not 0
// No byte carries into the next (15 + 15 < 256).
struct HalfCodes {
  uint32_t h, full, some;
};
__device__ __forceinline__ HalfCodes halfCodes(uint32_t Hc) {
  return HalfCodes{Hc, Hc + 0x01010101u, (Hc + 0x0F0F0F0Fu) & 0x10101010u};
}

// Tally of cell codes into the workgroup's LDS histogram. Cells that do not
// exist have code 0 (their pair bits are cleared), so no per-cell validity
// test is needed. All-one cells are counted per lane, four at a time, by a
// byte-wise test and a popcount; only mixed cells (code 1..254) issue an LDS
// atomic; all-zero cells are what remains of the cell count.
struct Tally {
  uint32_t ones = 0;  // this lane's all-one cells

  template <int K>
  __device__ __forceinline__ void addByte(uint32_t C, uint32_t D, uint32_t* hist) {
    const uint64_t mix = mixedLanes<K>(D);
    const uint32_t a = byteTimes4<K>(C);
    if (laneBit(mix))
      __hip_atomic_fetch_add((uint32_t*)((char*)hist + a), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  // The V cells between planes a (below) and b (above), from their half
  // codes: cell k is all-one iff both nibbles are 15, all-zero iff both are
  // 0, and mixed otherwise; its code is a.h | b.h << 4.
  template <int V>
  __device__ __forceinline__ void addPlanes(const HalfCodes& a, const HalfCodes& b, uint32_t* hist) {
    const uint32_t ones4 = a.full & b.full & 0x10101010u;
    ones += (uint32_t)__builtin_popcount(ones4);
    const uint32_t mixed = (a.some | b.some) & ~ones4;
    const uint32_t C = a.h | (b.h << 4);
    addMixed<0>(C, mixed, hist);
    if constexpr (V > 1) addMixed<1>(C, mixed, hist);
    if constexpr (V > 2) {
      addMixed<2>(C, mixed, hist);
      addMixed<3>(C, mixed, hist);
    }
  }
  template <int K>
  __device__ __forceinline__ void addMixed(uint32_t C, uint32_t mixed, uint32_t* hist) {
    const uint64_t mix = nonZeroByteLanes<K>(mixed);
    const uint32_t a = byteTimes4<K>(C);
    if (laneBit(mix))
      __hip_atomic_fetch_add((uint32_t*)((char*)hist + a), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  // The V cells of code word C (byte k: cell k; bytes V.. are zero).
  template <int V>
  __device__ __forceinline__ void add(uint32_t C, uint32_t* hist) {
    ones += (uint32_t)__builtin_popcount(allOneBytes(C));
    const uint32_t D = bytesMinusOne(C);
    addByte<0>(C, D, hist);
    if constexpr (V > 1) addByte<1>(C, D, hist);
    if constexpr (V > 2) {
      addByte<2>(C, D, hist);
      addByte<3>(C, D, hist);
    }
  }
};

// Clears the workgroup's LDS histogram and cell counter; every thread calls
// it before its first tally.
__device__ __forceinline__ void clearHistogram(uint32_t* hist, uint32_t* cnt) {
  for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
  if (threadIdx.x == 0) *cnt = 0;
  __syncthreads();
}

// Adds the workgroup's histogram to the result with one 64-bit atomic per
// nonzero bin. Bin 255 receives the waves' all-one cells, bin 0 the
// workgroup's cells (each wave passes its own count) minus every other bin.
// Every thread of the workgroup calls it once, after its tally; hist[256]
// and cnt are the workgroup's LDS histogram and a counter, cleared at start.
// With rows, the workgroup instead stores its 256 bins to row blockIdx.x of
// rows (plain stores), for kHistRowsReduce to add up: the device bins then
// take a few atomics per bin instead of one per bin and workgroup.
__device__ __forceinline__ void flushHistogram(uint32_t* hist, uint32_t* cnt, const Tally& tally, uint32_t cells,
                                               unsigned long long* __restrict__ bpfd,
                                               uint32_t* __restrict__ rows = nullptr) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t ones = waveSum(tally.ones);
  __syncthreads();
  if (lane == 0) {
    if (ones) atomicAdd(hist + 255, ones);
    atomicAdd(cnt, cells);
  }
  __syncthreads();
  if (threadIdx.x < kWave) {
    uint32_t rest = 0;
    for (int i = 1 + lane; i < 256; i += kWave) rest += hist[i];
    rest = waveSum(rest);
    if (lane == 0) hist[0] = *cnt - rest;
  }
  __syncthreads();
  if (rows) {
    uint32_t* const row = rows + (size_t)blockIdx.x * 256;
    for (int i = threadIdx.x; i < 256; i += blockDim.x) row[i] = hist[i];
    return;
  }
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    const uint32_t v = hist[i];
    if (v) atomicAdd(bpfd + i, (unsigned long long)v);
  }
}

// How a wave obtains the bit right of its last voxel.
enum HaloFrom : int { kHaloNone = 0, kHaloLds = 1, kHaloMem = 2 };

// Row-stack walker. The workgroup's waves sit side by side along x; each
// holds rows y0 .. y0+nY (nY <= TY) and marches z from z0 to z0+nZ, so one
// z-step of the workgroup reads nY+1 consecutive rows of one plane. At every
// step each wave publishes the low word of its voxel-0 masks in LDS (double
// buffered by z parity); one barrier later lane 63 takes bit 0 of its right
// neighbour's word as the right bit of its last voxel. Every wave of the
// workgroup runs the same number of steps (barriers).
// Rows read with non-temporal loads: rows 1 .. TY-1 of a stack are read by
// that stack only, so they need not stay in the caches; rows 0 and TY are
// also read by the neighbouring stacks (from L2) and use ordinary loads.
// At 1024^3 this took the fused kernel from 6.55 to 7.27 TB/s (f64, 2
// voxels per lane) and from 5.97 to 6.17 TB/s (f32); non-temporal loads of
// the shared rows too were slower than none.
template <int TY>
__device__ constexpr bool ntRow(int r) {
  return r > 0 && r < TY;
}

// Loads rows 0 .. TY at offsets yo[r] + zo, each with its load kind.
template <int TY, class Src, int... R>
__device__ __forceinline__ void loadRows(const Src& src, typename Src::Raw* out, const long long* yo, long long zo,
                                         std::integer_sequence<int, R...>) {
  ((out[R] = src.template load<ntRow<TY>(R)>(yo[R] + zo)), ...);
}

// With Mom::kOn the walker also adds the phase moments of the voxels this
// wave owns: its lanes' voxels inside the field (and, for the wave left of
// the field's last column when that column is read from memory, lane 63 adds
// it), in rows 0 .. nY-1 (and nY when that is the field's last row) and
// planes z0 .. z0+nZ-1 (and z0+nZ when that is the field's last plane), so
// every voxel is counted once over the grid.
template <int TY, int V, class Src, class Mom = NoMom>
__device__ __forceinline__ void walkStack(const Src& src, uint32_t (*xch)[TY + 1][16], int w, HaloFrom halo,
                                          int y0, int nY, int z0, int nZ, uint32_t* hist, Tally& tally,
                                          Mom* mom = nullptr) {
  const int lane = threadIdx.x & (kWave - 1);
  const int nb = min(w + 1, 15);  // neighbour slot, clamped inside the array
  const uint32_t ldsKeep = (halo == kHaloLds) ? 1u : 0u;
  const uint32_t lane63 = (lane == kWave - 1) ? 1u : 0u;
  long long yo[TY + 1];
#pragma unroll
  for (int r = 0; r <= TY; ++r) yo[r] = (long long)(y0 + min(r, nY)) * src.sy;
  // One z-step: loads plane z+1 into nxt, forms the pair words of plane z
  // (in cur) into Hc, and tallies the cells between Hp (plane z-1) and Hc.
  auto step = [&](int z, typename Src::Raw(&cur)[TY + 1], typename Src::Raw(&nxt)[TY + 1],
                  const HalfCodes(&Hp)[TY], HalfCodes(&Hc)[TY]) {
    const long long zon = (long long)min(z + 1, z0 + nZ) * src.sz;
    loadRows<TY>(src, nxt, yo, zon, std::make_integer_sequence<int, TY + 1>{});
    const int par = z & 1;
    uint32_t P[TY + 1];
#pragma unroll
    for (int r = 0; r <= TY; ++r) {
      uint64_t R0;
      P[r] = src.pairs(cur[r], R0);
      if (!Src::kSelfHalo && lane == 0) xch[par][r][w] = (uint32_t)R0;
    }
    if constexpr (Mom::kOn) {
      if (z < z0 + nZ || mom->lastPlane) {
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
          if (r < nY || (r == nY && mom->lastRow)) src.moments(cur[r], mom->vkeep, mom->acc);
        }
      }
    }
    if constexpr (!Src::kSelfHalo) {
      __syncthreads();
      if (halo == kHaloMem) {
        const long long zo = (long long)z * src.sz;
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
          if constexpr (Mom::kOn) {
            const auto h = src.haloVal(yo[r] + zo);
            const uint32_t hb = (uint32_t)(relMask<Src::kM>(h, src.thr) & 1ull);
            P[r] |= (hb & lane63) << kRightBit<V>;
            const bool own = mom->haloLast && lane63 && (z < z0 + nZ || mom->lastPlane) &&
                             (r < nY || (r == nY && mom->lastRow));
            mom->acc.add(h, hb != 0, own);
          } else {
            P[r] |= (src.haloBit(yo[r] + zo) & lane63) << kRightBit<V>;
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
          const uint32_t hb = __builtin_amdgcn_readfirstlane(xch[par][r][nb]) & ldsKeep;
          P[r] |= (hb & lane63) << kRightBit<V>;
        }
      }
    }
    // The map histogram (PBM rows, issue-bound) tests each plane's half
    // codes once for both plane pairs it belongs to; the field kernels tally
    // each plane pair's code word (2100 x 1000 x 500 f32 at 4 voxels per
    // lane: 0.652 ms against 0.783 with half codes; 1030 x 1000 x 1000: 0.687
    // against 0.743).
#pragma unroll
    for (int r = 0; r < TY; ++r) {
      uint32_t h = P[r] | (P[r + 1] << 2);
      if constexpr (HasPairKeep<Src>::value) h &= src.pairKeep(r);
      if constexpr (Src::kSelfHalo) {
        Hc[r] = halfCodes(h);
      } else {
        if (z > z0 && r < nY) tally.template add<V>(Hp[r].h | (h << 4), hist);
        Hc[r] = HalfCodes{h, 0u, 0u};
      }
    }
    if constexpr (Src::kSelfHalo) {
      if (z > z0) {
#pragma unroll
        for (int r = 0; r < TY; ++r) {
          if (r < nY) tally.template addPlanes<V>(Hp[r], Hc[r], hist);
        }
      }
    }
  };
  typename Src::Raw ra[TY + 1], rb[TY + 1];
  HalfCodes ha[TY], hb[TY];
  loadRows<TY>(src, ra, yo, (long long)z0 * src.sz, std::make_integer_sequence<int, TY + 1>{});
  if constexpr (Src::kSelfHalo) {
    // Unrolled by two: the next plane's rows and the previous plane's words
    // alternate between two register sets instead of being copied each step
    // (the map histogram, which is issue-bound, saves those moves).
    for (int z = z0; z <= z0 + nZ;) {
      step(z, ra, rb, hb, ha);
      if (++z > z0 + nZ) break;
      step(z, rb, ra, ha, hb);
      if (++z > z0 + nZ) break;
    }
  } else {
    for (int z = z0; z <= z0 + nZ; ++z) {
      step(z, ra, rb, hb, ha);
#pragma unroll
      for (int r = 0; r < TY; ++r) hb[r] = ha[r];
#pragma unroll
      for (int r = 0; r <= TY; ++r) ra[r] = rb[r];
    }
  }
}

}  // namespace mkf
