// mkf_kernels.hip - gfx950 kernels of the hot path and their launchers.
//
//   kPBM         field -> packed binary map      (replaces mapField /
//                mapStrideField, src/binMapCUDA.cuh:282-323)
//   kBPFD4/kBPFD PBM -> 256-bin pattern histogram, 4 or 1 voxels per lane
//                (replaces addMultiPlaneSeqBPFD / addPlaneBPFD,
//                src/bpfdCUDA.cuh:158-206)
//   kFusedBPFD   field -> histogram, reading the field once (no reference
//                counterpart; the reference round-trips the PBM through memory)
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "mkf_device.hpp"
#include "mkf_internal.h"

// The file is compiled once per part (-DMKF_KPART=0..8), each part holding
// only its own launchers and so instantiating only its own kernels; the
// parts compile in parallel. Without MKF_KPART one compile holds them all.
//   0 tiling choice and the fused dispatcher   5 kFusedBPFD with moments
//   1 kPBM and the moment reduction            6 kBPFD4 / kBPFD16 / kBPFD
//   2 kFusedBPFD for float fields              7 kPBM for double fields
//   3 kFusedBPFD for double fields             8 kFused16 with moments
//   4 kFused16
#ifdef MKF_KPART
#define MKF_PART(p) (MKF_KPART == (p))
#else
#define MKF_PART(p) 1
#endif

namespace mkf {

// ---------------------------------------------------------------------------
// Packed binary map. A lane holds 16 voxels of a row as E-voxel pieces
// (E = 4 floats or 2 doubles, one 16-byte load each) spaced 64 E voxels apart,
// so each load of the wave reads 1 KiB of consecutive bytes -- whole lines,
// read with non-temporal loads since every voxel is read once -- and a wave
// covers 1024 voxels of a row. PBM word 2 E q + g of the wave's segment comes
// from load q of the 32 / E lanes of group g: each lane packs its E compare
// bits (v_addc doubling steps with the ballot as carry-in), shifts them to
// its place, and three or four DPP OR steps (quad xor 1, quad xor 2,
// half-row mirror, row mirror) give the whole word to every lane of the
// group. 32 writer lanes then store the row's 32 words as one 128-byte line.
// (Each lane holding 16 consecutive voxels, i.e. four lane-strided 16-byte
// loads, was 7-10 % slower at 1024^3 and could not use non-temporal loads;
// 256 voxels per wave with 32-byte stores was slower still.) A wave owns one
// 1024-voxel segment and every C-th row of it (rows counted y fastest, then
// z; C waves per segment), kPbmAhead rows at a time; rows advance by adding
// (C mod Y, C / Y) to (y, z), with no division.
#ifndef MKF_PBM_AHEAD
#define MKF_PBM_AHEAD 2
#endif
#ifndef MKF_PBM_WAVE_TARGET
#define MKF_PBM_WAVE_TARGET 16384
#endif
#ifndef MKF_PBM_INTERLEAVE
#define MKF_PBM_INTERLEAVE 1
#endif
#ifndef MKF_PBM_STAGE
#define MKF_PBM_STAGE 128
#endif
constexpr int kPbmWaves = 8;
constexpr int kPbmAhead = MKF_PBM_AHEAD;
constexpr int kPbmVoxels = 16;  // per lane
constexpr int kPbmStage = MKF_PBM_STAGE;  // map rows staged per wave

// Phase moments (MOM): each lane also sums, in double, the field values of
// its voxels and their squares per phase (map bit 0 or 1) and counts them
// (MomAcc). Each workgroup writes its totals {s0, q0, s1, q1, n0, n1} to
// part[blockIdx.x] after a fixed-order reduction; kMomentsReduce adds the
// partials in a fixed order, so the result does not depend on scheduling.
// With pW == nullptr no map is written.

// Rows of a segment are dealt to its waves either as runs (wave c takes rows
// c R .. c R + R - 1, rowStep 1) or interleaved (wave c takes rows c, c + C,
// c + 2 C, ... for C waves, rowStep C), so that the waves resident at one time
// read neighbouring rows and write neighbouring map lines. A wave steps its
// row index r = y + Y z by rowStep = dy + Y dz.
struct RowSched {
  long long nRows, rowsPerWave, nChunks, chunkStep, rowStep;
  int dy, dz;
  int Z;  // planes of the field (rows of row groups past the last plane are not there)
};

// VPL voxels per lane (16, 8 or 4): segments of 1024, 512 or 256 voxels, so
// that narrow and ragged rows keep most lanes busy. Row groups (GRP = 2, 4
// or 8; float fields at 4 voxels per lane): the wave's lanes form GRP groups
// of 64 / GRP lanes, group k reading row GRP R + k of row group R, so rows of
// at most 256 / GRP voxels fill the wave. The schedule then counts row
// groups, and each lane keeps its own row; lanes whose row is past the last
// plane load the last row, keep no voxel and write nothing.
template <typename T, int M, bool G, bool MOM, int VPL = kPbmVoxels, int GRP = 1>
__global__ __launch_bounds__(kPbmWaves* kWave) void kPBM(const FieldArgs<T> fa, T thr, int X, int Y, int nSegX,
                                                           const RowSched rs, uint32_t* __restrict__ pW,
                                                           uint32_t rowWS, uint32_t planeWS,
                                                           double* __restrict__ part) {
  constexpr int V = VPL;
  constexpr int E = 16 / sizeof(T);  // elements per 16-byte load
  static_assert(V % E == 0 && V >= E && V <= kPbmVoxels, "whole 16-byte loads, at most 16 voxels per lane");
  static_assert(GRP == 1 || (V == E && E == 4 && !G), "row groups: one 16-byte load of floats per lane and row");
  constexpr int SL = kWave / GRP;  // lanes per row
  constexpr int W = kWave * V;
  typedef T Vec __attribute__((ext_vector_type(E), aligned(sizeof(T))));
  const int lane = threadIdx.x & (kWave - 1);
  const int ll = lane % SL;  // lane within its row
  const long long gw = (long long)blockIdx.x * kPbmWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int seg = (int)(gw % nSegX);
  // Waves past the last chunk (the grid is whole workgroups) take no rows:
  // interleaved, their first row would be another wave's.
  const long long chunk = gw / nSegX;
  const long long r0 = chunk * rs.chunkStep;
  if (!MOM && chunk >= rs.nChunks) return;  // with moments every wave reaches the block reduction
  const long long n =
      chunk >= rs.nChunks ? 0ll : min(rs.rowsPerWave, (rs.nRows - r0 + rs.rowStep - 1) / rs.rowStep);
  const long long rl = GRP > 1 ? r0 * GRP + lane / SL : r0;  // this lane's first row
  int y = (int)(rl % Y);
  long long z = rl / Y;
  const int x0 = seg * W;
  // Lane-contiguous layout: load q of a row gives lane l the E voxels at
  // x0 + 64 E q + E l, so each 16-byte load of the wave covers 1 KiB of
  // consecutive bytes (whole lines) and can bypass the caches. PBM word
  // 2 E q + g of the segment is built from the LPW = 32 / E lanes of group g,
  // each contributing E bits.
  constexpr int NQ = V / E;    // 16-byte loads per lane and row
  constexpr int LPW = 32 / E;  // lanes per PBM word
  typedef const __attribute__((address_space(1))) Vec* GVec;
  const bool vec = !G && x0 + W <= X;
  uint32_t kq[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int xs = x0 + kWave * E * q + E * ll;
    kq[q] = (xs >= X) ? 0u : ((X - xs >= E) ? ((1u << E) - 1u) : ((1u << (X - xs)) - 1u));
  }
  const int jl = lane % LPW;  // lane within its word group; writers take word q = jl
  const uint32_t word = (uint32_t)(x0 >> 5) + (uint32_t)(2 * E * jl + ll / LPW);
  const bool writer = pW != nullptr && jl < NQ && word < rowWS;
  const uint32_t sh = (uint32_t)(E * jl);
  MomAcc acc;
  // Map staging (forms without moments): each writer lane keeps its word of
  // up to kPbmStage rows in the wave's LDS slice and stores them back to
  // back when the slice is full or the run ends, instead of one store per
  // step. Only the lane that wrote a slot reads it back, so no barrier is
  // needed. Gathering a wave's map writes in time makes them cheaper for HBM
  // to interleave with the field's reads (the stores themselves cost
  // nothing: folded into one L2-resident plane they leave the pass at the
  // read roof).
  constexpr bool STAGE = !MOM;
  constexpr int kSt = STAGE ? kPbmStage : kPbmAhead;
  static_assert(kSt % kPbmAhead == 0 && (!STAGE || kSt % 8 == 0), "staged rows must be whole steps");
  __shared__ uint32_t stage[STAGE ? kPbmWaves * kSt * 32 : 1];
  uint32_t* const myStage = stage + (STAGE ? (threadIdx.x >> 6) * kSt * 32 : 0);
  // The writer's slot in its wave's staging row: its word in the segment, or
  // with row groups its group of LPW lanes.
  const uint32_t widx = GRP > 1 ? (uint32_t)(lane / LPW) : ((word - (uint32_t)(x0 >> 5)) & 31u);
  int yb = y;  // first staged row
  long long zb = z;
  // Partial segment (not the general form): load q of a row is read from
  // min(64 E q + E lane, X - x0 - E), so it ends inside the row (the host
  // sends rows shorter than E voxels to the general form), as an unsigned
  // offset from E voxels before the segment; shq[q] (0..E-1) shifts the
  // voxels back into place where they are used.
  uint32_t offq[NQ];
  int shq[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int start = kWave * E * q + E * ll;
    offq[q] = (uint32_t)(min(start, X - x0 - E) + E);
    shq[q] = start + E - (int)offq[q];
  }
  // The row loop, compiled once for whole segments (VEC) and once for the
  // partial or general ones, and chosen per wave, so that neither carries
  // the other's loads or branches.
  auto walk = [&](auto vecTag) {
    constexpr bool VEC = decltype(vecTag)::value;
    for (long long i = 0; i < n; i += kPbmAhead) {
      T v[kPbmAhead][V];
      size_t wo[kPbmAhead];
      bool okr[kPbmAhead];  // the lane's row is in the field (row groups)
#pragma unroll
      for (int u = 0; u < kPbmAhead; ++u) {
        if (i + u < n) {
          // Several fields are summed in table order (general form only).
          long long ro = (long long)y * fa.sy + z * fa.sz;
          if constexpr (GRP > 1) {
            okr[u] = z < rs.Z;
            if (!okr[u]) ro = (long long)(Y - 1) * fa.sy + (long long)(rs.Z - 1) * fa.sz;
          }
          for (int fi = 0; fi < (G ? fa.nF : 1); ++fi) {
            const T* row = fa.p[fi] + ro;
            T t[V];
            if constexpr (VEC) {
#pragma unroll
              for (int q = 0; q < NQ; ++q) {
                const Vec w4 = __builtin_nontemporal_load((GVec)(row + x0 + kWave * E * q + E * lane));
#pragma unroll
                for (int e = 0; e < E; ++e) t[E * q + e] = w4[e];
              }
            } else if constexpr (G) {
#pragma unroll
              for (int q = 0; q < NQ; ++q) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                  const int x = min(x0 + kWave * E * q + E * lane + e, X - 1);
                  t[E * q + e] = row[(long long)x * fa.sx];
                }
              }
            } else {
              const char* rb = (const char*)(row + x0 - E);
#pragma unroll
              for (int q = 0; q < NQ; ++q) {
                const Vec w4 = __builtin_nontemporal_load((GVec)(rb + offq[q] * (uint32_t)sizeof(T)));
#pragma unroll
                for (int e = 0; e < E; ++e) t[E * q + e] = w4[e];
              }
            }
#pragma unroll
            for (int k = 0; k < V; ++k) v[u][k] = fi ? v[u][k] + t[k] : t[k];
          }
          wo[u] = (size_t)y * rowWS + (size_t)z * planeWS + word;
          y += rs.dy;
          z += rs.dz;
          if (y >= Y) {
            y -= Y;
            ++z;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kPbmAhead; ++u) {
        if (i + u < n) {
          if constexpr (!VEC && !G) {
            // Shift each load's voxels back by shq (voxels past the row's
            // end are cleared by kq and owned by nobody).
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
              const bool s1 = (shq[q] & 1) != 0;
              if constexpr (E == 4) {
                const bool s2 = (shq[q] & 2) != 0;
                const T a0 = v[u][4 * q], a1 = v[u][4 * q + 1], a2 = v[u][4 * q + 2], a3 = v[u][4 * q + 3];
                const T t0 = s1 ? a1 : a0, t1 = s1 ? a2 : a1, t2 = s1 ? a3 : a2;
                v[u][4 * q] = s2 ? t2 : t0;
                v[u][4 * q + 1] = s2 ? a3 : t1;
                v[u][4 * q + 2] = t2;
              } else {
                v[u][2 * q] = s1 ? v[u][2 * q + 1] : v[u][2 * q];
              }
            }
          }
          uint32_t wq[NQ];
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const uint32_t kqu = (GRP == 1 || okr[u]) ? kq[q] : 0u;
            uint32_t nib = laneBit(relMask<M>(v[u][E * q + E - 1], thr)) ? 1u : 0u;
#pragma unroll
            for (int e = E - 2; e >= 0; --e) nib = dblAddLaneBit(nib, relMask<M>(v[u][E * q + e], thr));
            nib &= kqu;
            uint32_t pw = orHalfMirror(orQuadXor2(orQuadXor1(nib << sh)));
            if constexpr (LPW == 16) pw = orRowMirror(pw);
            wq[q] = pw;
            if constexpr (MOM) {
#pragma unroll
              for (int e = 0; e < E; ++e) acc.add(v[u][E * q + e], ((nib >> e) & 1u) != 0, ((kqu >> e) & 1u) != 0);
            }
          }
          uint32_t w32 = wq[0];
#pragma unroll
          for (int q = 1; q < NQ; ++q) w32 = (jl == q) ? wq[q] : w32;
          if constexpr (STAGE) {
            if (writer) myStage[(int)((i + u) % kSt) * 32 + widx] = w32;
          } else if (writer && (GRP == 1 || okr[u])) {
            pW[wo[u]] = w32;
          }
        }
      }
      if constexpr (STAGE) {
        if ((i + kPbmAhead) % kSt == 0 || i + kPbmAhead >= n) {
          // Eight slots are read before their eight stores are issued.
          const int nb = (int)(min(i + kPbmAhead, n) - (i / kSt) * kSt);
          for (int k0 = 0; k0 < nb; k0 += 8) {
            uint32_t t[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) t[j] = writer ? myStage[(k0 + j) * 32 + widx] : 0u;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              if (k0 + j < nb) {
                if (writer && (GRP == 1 || zb < rs.Z)) pW[(size_t)yb * rowWS + (size_t)zb * planeWS + word] = t[j];
                yb += rs.dy;
                zb += rs.dz;
                if (yb >= Y) {
                  yb -= Y;
                  ++zb;
                }
              }
            }
          }
        }
      }
    }
  };
  if (vec)
    walk(std::true_type{});
  else
    walk(std::false_type{});
  if constexpr (MOM) momBlockPartial(acc, part);
}

#if MKF_PART(1)
// Adds the per-workgroup moment partials to out[kMomN] in a fixed order:
// thread t sums partials t, t + 256, ... and a pairwise tree sums the threads.
__global__ __launch_bounds__(256) void kMomentsReduce(const double* __restrict__ part, long long nParts,
                                                      double* __restrict__ out) {
  __shared__ double red[256];
  for (int j = 0; j < kMomN; ++j) {
    double a = 0;
    for (long long i = threadIdx.x; i < nParts; i += 256) a += part[i * kMomN + j];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[j] += red[0];
    __syncthreads();
  }
}
#endif

// ---------------------------------------------------------------------------
// Histogram kernels. A tile is (x-tile, y-range, z-range): a workgroup of
// `waves` waves side by side along x (one 64-voxel segment each, so a
// workgroup reads whole rows), TY row pairs and up to TZ plane pairs. Block
// b runs tile (b & 7) * chunk + (b >> 3): blocks are dealt to the 8 XCDs
// round-robin, so consecutive tiles -- which share their halo rows -- run on
// the same XCD at about the same time and the shared row is an L2 hit.
// ---------------------------------------------------------------------------
struct TileGrid {
  int nXT, nYR, nZR, nTiles, chunk;
  int TZ, waves, nSegX;
};

__device__ __forceinline__ bool tileOf(const TileGrid& g, int& xt, int& yr, int& zr) {
  const int b = blockIdx.x;
  const int L = (b & 7) * g.chunk + (b >> 3);
  if (L >= g.nTiles) return false;
  xt = L % g.nXT;
  const int r = L / g.nXT;
  yr = r % g.nYR;
  zr = r / g.nYR;
  return true;
}

// Geometry of one wave inside its tile: the wave holds V voxels per lane, so
// it covers W = 64 V voxels starting at x0.
struct WaveGeom {
  int w, seg, x0, W, y0, nY, z0, nZ;
  HaloFrom halo;
};

__device__ __forceinline__ WaveGeom waveGeom(const TileGrid& g, int xt, int yr, int zr, int TY, int V, int X,
                                             int Y, int Z) {
  WaveGeom q;
  q.w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  q.seg = xt * g.waves + q.w;
  q.W = kWave * V;
  q.x0 = q.seg * q.W;
  q.y0 = yr * TY;
  q.nY = min(TY, Y - 1 - q.y0);
  q.z0 = zr * g.TZ;
  q.nZ = min(g.TZ, Z - 1 - q.z0);
  // The next wave's first voxel is x0 + W; that wave is in this workgroup
  // unless this wave is the last of its tile.
  if (q.x0 + q.W >= X) {
    q.halo = kHaloNone;
  } else if (q.w + 1 < g.waves) {
    q.halo = kHaloLds;
  } else {
    q.halo = kHaloMem;
  }
  return q;
}

// Lane mask of pair-word bytes whose cell exists, and the wave's cell count
// per plane pair and row pair.
template <int V>
__device__ __forceinline__ uint32_t cellKeep(int x0, int X, uint32_t& cellsPerRow) {
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t keep = 0, n = 0;
#pragma unroll
  for (int k = 0; k < V; ++k) {
    const bool ok = x0 + V * lane + k < X - 1;
    keep |= ok ? (0xFFu << (8 * k)) : 0u;
    n += ok ? 1u : 0u;
  }
  cellsPerRow = waveSum(n);
  return keep;
}

// With MOM the kernel also writes the workgroup's phase moments of the
// voxels it owns (see walkStack) to part[blockIdx.x]; blocks past the last
// tile write zeros. PW compiles the walk for partial waves into the plain
// 2-voxel double kernel too (for rows whose last segment is partial). With
// rows (forms without moments), each workgroup stores its histogram to
// rows[blockIdx.x] (zeros past the last tile) instead of adding it to bpfd.
template <typename T, int M, int V, int TY, bool G, bool MOM = false, bool PW = false>
__global__ __launch_bounds__(1024) void kFusedBPFD(const FieldArgs<T> fa, T thr, int X, int Y, int Z, TileGrid g,
                                                    unsigned long long* __restrict__ bpfd,
                                                    double* __restrict__ part = nullptr,
                                                    uint32_t* __restrict__ rows = nullptr) {
  __shared__ uint32_t hist[256];
  __shared__ uint32_t cnt;
  __shared__ uint32_t xch[2][TY + 1][16];
  int xt, yr, zr;
  if (!tileOf(g, xt, yr, zr)) {
    if (MOM && threadIdx.x < kMomN) part[(size_t)blockIdx.x * kMomN + threadIdx.x] = 0.0;
    if (!MOM && rows) {
      for (int i = threadIdx.x; i < 256; i += blockDim.x) rows[(size_t)blockIdx.x * 256 + i] = 0u;
    }
    return;
  }
  const int lane = threadIdx.x & (kWave - 1);
  const WaveGeom q = waveGeom(g, xt, yr, zr, TY, V, X, Y, Z);
  clearHistogram(hist, &cnt);

  const int xb = min(q.x0, X - 1);
  const bool vec = !G && (V == 1 || q.x0 + q.W <= X);
  uint32_t cellsPerRow;
  const uint32_t keep = cellKeep<V>(q.x0, X, cellsPerRow);
  // The walk, compiled for whole waves (and the general form) and, for the
  // moments forms, 4 voxels per lane and PW, once more for partial waves;
  // each wave runs one, and both pass the same barriers. (The plain 2-voxel
  // double kernel for rows of whole segments keeps one walk: the second copy
  // cost it 4 % at 1024^3.)
  auto run = [&](auto partTag) {
    constexpr bool PART = decltype(partTag)::value;
    FieldRows<T, M, V, G, PART> src;
    if constexpr (G) {
#pragma unroll
      for (int i = 0; i < kMaxFieldsDev; ++i) src.fw[i] = fa.p[i] + (long long)xb * fa.sx;
      src.nF = fa.nF;
      src.sx = fa.sx;
    } else {
      src.fw[0] = fa.p[0] + xb;  // one field, unit x stride
      src.nF = 1;
      src.sx = 1;
    }
    src.thr = thr;
    src.sy = fa.sy;
    src.sz = fa.sz;
    src.lx = (uint32_t)(min(q.x0 + V * lane, X - 1) - xb);
    src.last = X - 1 - xb;
    src.vec = vec;
    src.hxo = min(q.x0 + q.W, X - 1) - xb;
    src.vs = min(q.x0 + V * lane, X - V) - xb;
    src.shv = (uint32_t)(q.x0 + V * lane - min(q.x0 + V * lane, X - V));
    src.keep = keep;
    Tally tally;
    if constexpr (MOM) {
      OnMom mom;
      mom.vkeep = 0;
#pragma unroll
      for (int k = 0; k < V; ++k) mom.vkeep |= (q.x0 + V * lane + k < X) ? (1u << k) : 0u;
      mom.lastRow = q.y0 + q.nY == Y - 1;
      mom.lastPlane = q.z0 + q.nZ == Z - 1;
      mom.haloLast = q.halo == kHaloMem && q.x0 + q.W == X - 1;
      walkStack<TY, V, FieldRows<T, M, V, G, PART>, OnMom>(src, xch, q.w, q.halo, q.y0, q.nY, q.z0, q.nZ, hist,
                                                           tally, &mom);
      flushHistogram(hist, &cnt, tally, cellsPerRow * q.nY * q.nZ, bpfd);
      momBlockPartial(mom.acc, part);
    } else {
      walkStack<TY, V>(src, xch, q.w, q.halo, q.y0, q.nY, q.z0, q.nZ, hist, tally);
      flushHistogram(hist, &cnt, tally, cellsPerRow * q.nY * q.nZ, bpfd, rows);
    }
  };
  if constexpr (!G && V > 1 && (MOM || V == 4 || PW)) {
    if (vec)
      run(std::false_type{});
    else
      run(std::true_type{});
  } else {
    run(std::false_type{});
  }
}

template <int TY>
__global__ __launch_bounds__(1024) void kBPFD(const uint32_t* __restrict__ pW, int X, int Y, int Z,
                                              uint32_t rowWS, uint32_t planeWS, TileGrid g,
                                              unsigned long long* __restrict__ bpfd) {
  __shared__ uint32_t hist[256];
  __shared__ uint32_t cnt;
  __shared__ uint32_t xch[2][TY + 1][16];
  int xt, yr, zr;
  if (!tileOf(g, xt, yr, zr)) return;
  const WaveGeom q = waveGeom(g, xt, yr, zr, TY, 1, X, Y, Z);
  clearHistogram(hist, &cnt);

  PbmRows src;
  const uint32_t lastWord = rowWS - 1u;
  const uint32_t wi0 = min(2u * q.seg, lastWord);
  src.ws = pW + wi0;
  src.sy = rowWS;
  src.sz = planeWS;
  src.o1 = (2u * q.seg + 1u <= lastWord) ? 1u : 0u;
  src.hw = min(2u * q.seg + 2u, lastWord) - wi0;
  uint32_t cellsPerRow;
  src.keep = cellKeep<1>(q.x0, X, cellsPerRow);
  Tally tally;
  walkStack<TY, 1>(src, xch, q.w, q.halo, q.y0, q.nY, q.z0, q.nZ, hist, tally);
  flushHistogram(hist, &cnt, tally, cellsPerRow * q.nY * q.nZ, bpfd);
}

// Histogram of an existing PBM with 4 voxels per lane (PbmRows4): each lane
// reads its PBM word and the next, so there is no halo exchange.
template <int TY>
__global__ __launch_bounds__(1024) void kBPFD4(const uint32_t* __restrict__ pW, int X, int Y, int Z,
                                               uint32_t rowWS, uint32_t planeWS, TileGrid g,
                                               unsigned long long* __restrict__ bpfd) {
  __shared__ uint32_t hist[256];
  __shared__ uint32_t cnt;
  int xt, yr, zr;
  if (!tileOf(g, xt, yr, zr)) return;
  const int lane = threadIdx.x & (kWave - 1);
  const WaveGeom q = waveGeom(g, xt, yr, zr, TY, 4, X, Y, Z);
  clearHistogram(hist, &cnt);

  PbmRows4 src;
  src.row = pW;
  src.sy = rowWS;
  src.sz = planeWS;
  const uint32_t wl = (uint32_t)(q.x0 >> 5) + (uint32_t)(lane >> 3);
  src.b0 = 4u * min(wl, rowWS - 1u);
  src.b1 = 4u * min(wl + 1u, rowWS - 1u);
  src.sh = 4u * (uint32_t)(lane & 7);
  uint32_t cellsPerRow;
  src.keep = cellKeep<4>(q.x0, X, cellsPerRow);
  Tally tally;
  walkStack<TY, 4>(src, (uint32_t(*)[TY + 1][16]) nullptr, q.w, kHaloNone, q.y0, q.nY, q.z0, q.nZ, hist, tally);
  flushHistogram(hist, &cnt, tally, cellsPerRow * q.nY * q.nZ, bpfd);
}

// Histogram of a narrow PBM (rows of at most GW <= 8 words, X <= 256): the
// lanes form G = 8 / GW groups of 8 GW lanes, group k walking row pairs
// yT + k TY .. yT + k TY + TY - 1 of the wave's tile, so a wave covers G TY
// row pairs of 32 GW voxels instead of TY row pairs of a 256-voxel segment
// that is mostly past the row's end. A group whose rows would run past the
// field reads rows Y - 1 - TY .. Y - 1 instead and masks off the pairs that
// belong to the group before it (kp), so every load stays inside the map and
// every cell is counted once. The host requires Y - 1 >= TY. Each of the
// workgroup's g.waves waves takes a tile of its own (consecutive tiles, which
// share their halo rows), and the waves share one LDS histogram, so the
// device histogram takes one flush per workgroup instead of one per tile.
template <int TY, int GW>
__global__ __launch_bounds__(1024) void kBPFD4R(const uint32_t* __restrict__ pW, int X, int Y, int Z,
                                                uint32_t rowWS, uint32_t planeWS, TileGrid g,
                                                unsigned long long* __restrict__ bpfd) {
  constexpr int G = 8 / GW;
  static_assert(GW == 1 || GW == 2 || GW == 4 || GW == 8, "1, 2, 4 or 8 words per row group");
  __shared__ uint32_t hist[256];
  __shared__ uint32_t cnt;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.x;
  const long long L = (long long)((b & 7) * g.chunk + (b >> 3)) * g.waves + w;
  clearHistogram(hist, &cnt);
  Tally tally;
  uint32_t cells = 0;
  if (L < g.nTiles) {
    const int yT = (int)(L % g.nYR) * (G * TY);
    const int z0 = (int)(L / g.nYR) * g.TZ;
    const int nZ = min(g.TZ, Z - 1 - z0);
    const int yWant = yT + (lane / (8 * GW)) * TY;  // the group's first row pair
    const int yBase = min(yT, Y - 1 - TY);           // group 0's rows (wave-uniform)
    const int yk = min(yWant, Y - 1 - TY);           // this group's rows
    const uint32_t wl = (uint32_t)((lane >> 3) & (GW - 1));
    PbmRowsR<TY> src;
    src.row = pW;
    src.sy = rowWS;
    src.sz = planeWS;
    const uint32_t ro = (uint32_t)(yk - yBase) * rowWS;
    src.b0 = 4u * (ro + min(wl, rowWS - 1u));
    src.b1 = 4u * (ro + min(wl + 1u, rowWS - 1u));
    src.sh = 4u * (uint32_t)(lane & 7);
    const int x0 = 32 * (int)wl + 4 * (lane & 7);
    uint32_t xkeep = 0, nx = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const bool ok = x0 + c < X - 1;
      xkeep |= ok ? (0xFFu << (8 * c)) : 0u;
      nx += ok ? 1u : 0u;
    }
    src.keep = xkeep;
    uint32_t np = 0;
#pragma unroll
    for (int r = 0; r < TY; ++r) {
      const bool ok = yk + r >= yWant && yk + r < Y - 1;
      src.kp[r] = ok ? xkeep : 0u;
      np += ok ? 1u : 0u;
    }
    cells = waveSum(nx * np) * (uint32_t)nZ;
    walkStack<TY, 4>(src, (uint32_t(*)[TY + 1][16]) nullptr, 0, kHaloNone, yBase, TY, z0, nZ, hist, tally);
  }
  flushHistogram(hist, &cnt, tally, cells, bpfd);
}

// ---------------------------------------------------------------------------
// Fused histogram, float fields, 4 NG voxels per lane (voxelsPerLane 16 for
// NG = 4, 8 for NG = 2). The lane layout of kPBM: lane l holds groups
// q = 0..NG-1 of 4 voxels at x0 + 256 q + 4 l, so each 16-byte load of the
// wave reads 1 KiB of consecutive bytes and a wave covers a 256 NG-voxel
// segment of its rows. The
// right neighbour of a group's last voxel is the next lane's first voxel of
// the group: bit l + 1 of that voxel's ballot; lane 63 takes bit 0 of the
// next group's ballot, or, for the last group, the voxel right of the
// segment, read by a scalar load issued with the row loads one plane ahead.
// A group's 4 voxel bits and its right bit enter a 5-bit value by carry-in
// adds; one multiply spreads it into the byte-pair word of its 4 cells. One
// wave per workgroup: no LDS exchange and no barrier.
// ---------------------------------------------------------------------------
// The fused 16-voxel kernel runs at most 2 waves per SIMD (8 per CU): more
// requests in flight lowered its throughput (1024^3: 3 waves per SIMD
// 0.66-0.73 ms against 0.65 at 2).
#ifndef MKF_FUSED16_PER_CU
#define MKF_FUSED16_PER_CU 8
#endif
constexpr int kFused16PerCU = MKF_FUSED16_PER_CU;

//
// With MOM the kernel also sums the phase moments of the voxels each tile
// owns: its segment's voxels below X, rows y0 .. y0 + nY - 1 (and row nY when
// it is the field's last row), planes z0 .. z0 + nZ - 1 (and the last plane
// likewise), and, when the segment ends one voxel short of the row's end,
// that last voxel (read as the halo). Every workgroup writes its moments to
// part[blockIdx.x] after its last tile.
// The moments form asks for 3 waves per SIMD (at most 168 VGPRs); the plain
// form's code is unchanged by the attribute (1: no bound).
template <int M, int TY, int NG, bool MOM = false>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(MOM ? 3 : 1))) void kFused16(const float* __restrict__ f, float thr, int X, int Y, int Z,
                                                  long long sy, long long sz, TileGrid g,
                                                  unsigned long long* __restrict__ bpfd,
                                                  double* __restrict__ part = nullptr) {
  typedef float Vec __attribute__((ext_vector_type(4), aligned(4)));
  typedef const __attribute__((address_space(1))) Vec* GVec;
  __shared__ uint32_t hist[256];
  __shared__ uint32_t cnt;
  const int lane = threadIdx.x;
  clearHistogram(hist, &cnt);
  constexpr int VPL = 4 * NG;  // voxels per lane
  Tally tally;
  uint32_t cells = 0;
  MomWave mw;
  uint32_t n1Lane = 0;  // the lane's voxels counted in phase 1 (wave counts are in mw)
  // Persistent: block b takes tiles b, b + gridDim.x, ... in the XCD-aware
  // order (L = (t & 7) * chunk + (t >> 3)), so all of a block's tiles lie in
  // its XCD's share of the field (the grid is a multiple of 8).
  for (int t = blockIdx.x; t < g.chunk * 8; t += gridDim.x) {
    const int L = (t & 7) * g.chunk + (t >> 3);
    if (L >= g.nTiles) continue;
    const int xt = L % g.nXT;
    const int rr = L / g.nXT;
    const int yr = rr % g.nYR;
    const int zr = rr / g.nYR;
    const int x0 = xt * (kWave * VPL);
    const int y0 = yr * TY;
    const int nY = min(TY, Y - 1 - y0);
    const int z0 = zr * g.TZ;
    const int nZ = min(g.TZ, Z - 1 - z0);
    const bool vec = x0 + kWave * VPL <= X;
    const bool halo = x0 + kWave * VPL < X;  // the voxel right of the segment exists
    const bool lastRow = y0 + nY == Y - 1, lastPlane = z0 + nZ == Z - 1;
    const bool haloLast = x0 + kWave * VPL == X - 1;  // the halo is the row's last voxel
    // cells that exist: x < X - 1
    uint32_t keep[NG], nCells = 0;
#pragma unroll
    for (int q = 0; q < NG; ++q) {
      keep[q] = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const bool ok = x0 + 256 * q + 4 * lane + e < X - 1;
        keep[q] |= ok ? (0xFFu << (8 * e)) : 0u;
        nCells += ok ? 1u : 0u;
      }
    }
    cells += waveSum(nCells) * (uint32_t)nY * (uint32_t)nZ;
    long long yo[TY + 1];
#pragma unroll
    for (int r = 0; r <= TY; ++r) yo[r] = (long long)(y0 + min(r, nY)) * sy;
    const float* base = f + x0;
    // Partial segment: group q is read from min(256 q + 4 lane, X - x0 - 4)
    // (X >= 4, checked by the host, so the 4 voxels before the row's end
    // exist), as an unsigned offset from 4 voxels before the segment, which
    // keeps the loads in the scalar-base + 32-bit-offset form.
    uint32_t ga[NG];
#pragma unroll
    for (int q = 0; q < NG; ++q) ga[q] = (uint32_t)(min(256 * q + 4 * lane, X - x0 - 4) + 4);

    // The slab walk, compiled once for whole segments (VEC) and once for the
    // last, partial segment of a row, and chosen per tile, so that neither
    // carries the other's loads or branches.
    auto walk = [&](auto vecTag) {
      constexpr bool VEC = decltype(vecTag)::value;
      auto loadPlane = [&](long long zo, float (&v)[TY + 1][VPL], float (&h)[TY + 1]) {
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
          const float* row = base + yo[r] + zo;
#pragma unroll
          for (int q = 0; q < NG; ++q) {
            Vec w;
            if constexpr (VEC) {
              w = (r > 0 && r < TY) ? __builtin_nontemporal_load((GVec)(row + 256 * q + 4 * lane))
                                    : *(const Vec*)(row + 256 * q + 4 * lane);
            } else {
              // Partial segment: group q's 16 bytes are read from at most 4
              // voxels before the row's end (ga[q], counted from 4 voxels
              // before the segment); the walk shifts them into place when it
              // uses them, so the loads stay in flight during the tally.
              const char* rowb = (const char*)(row - 4);
              const uint32_t byteOff = ga[q] * 4u;
              w = (r > 0 && r < TY) ? __builtin_nontemporal_load((GVec)(rowb + byteOff))
                                    : *(const Vec*)(rowb + byteOff);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) v[r][4 * q + e] = w[e];
          }
          h[r] = halo ? uniformLoad(row + kWave * VPL) : 0.f;
        }
      };

      // One plane of registers: the pair words of plane z are computed
      // first, then plane z+1 is loaded into the same registers while plane
      // z's cells are tallied.
      float cur[TY + 1][VPL];
      float hc[TY + 1];
      loadPlane((long long)z0 * sz, cur, hc);
      uint32_t Hp[TY][NG];
      // Moments, partial segment: the row ends in group qs (the groups below
      // it lie inside the row, those above it outside); pm[e] holds the lanes
      // whose voxel e of group qs lies inside, ob the lane's voxels of group
      // qs inside as 4 bits.
      const int rowLen = min(X - x0, kWave * VPL);  // voxels of the segment inside the row
      int qs = NG;
      uint64_t pm[4] = {~0ull, ~0ull, ~0ull, ~0ull};
      uint32_t ob = 15u;
      if constexpr (MOM && !VEC) {
        qs = rowLen >> 8;
#pragma unroll
        for (int e = 0; e < 4; ++e) pm[e] = ballot(256 * qs + 4 * lane + e < rowLen);
        ob = (1u << min(max(rowLen - 256 * qs - 4 * lane, 0), 4)) - 1u;
      }
      for (int z = z0; z <= z0 + nZ; ++z) {
        const bool planeOwn = z < z0 + nZ || lastPlane;
        uint32_t P[TY + 1][NG];
#pragma unroll
        for (int r = 0; r <= TY; ++r) {
          if constexpr (!VEC) {
            // Shift each group's voxels back by sh = 0..3 (voxels past the
            // row's end only meet cells that keep[] drops).
#pragma unroll
            for (int q = 0; q < NG; ++q) {
              const int sh = 256 * q + 4 * lane + 4 - (int)ga[q];
              const bool s1 = (sh & 1) != 0, s2 = (sh & 2) != 0;
              float* g4 = &cur[r][4 * q];
              const float t0 = s1 ? g4[1] : g4[0], t1 = s1 ? g4[2] : g4[1], t2 = s1 ? g4[3] : g4[2];
              g4[0] = s2 ? t2 : t0;
              g4[1] = s2 ? g4[3] : t1;
              g4[2] = t2;
            }
          }
          const uint64_t hb = (halo && (relMask<M>(hc[r], thr) & 1ull)) ? 1ull : 0ull;
          if constexpr (MOM) {
            // Group by group from the right, so that each group's masks die
            // with it: its pair word, then the moments of its voxels. Rows
            // past nY repeat row nY (the loads are clamped); row nY below TY
            // is the field's last row, row TY only counts as the last row.
            const bool rowOwn = planeOwn && (r < TY ? r <= nY : (nY == TY && lastRow));
            uint64_t nextFirst = hb;
#pragma unroll
            for (int q = NG - 1; q >= 0; --q) {
              uint64_t mq[4];
#pragma unroll
              for (int e = 0; e < 4; ++e) mq[e] = relMask<M>(cur[r][4 * q + e], thr);
              uint32_t v5 = laneBit((mq[0] >> 1) | (nextFirst << 63)) ? 1u : 0u;
#pragma unroll
              for (int e = 3; e >= 0; --e) v5 = dblAddLaneBit(v5, mq[e]);
              P[r][q] = ((v5 * 0x204081u) & 0x03030303u) & keep[q];
              nextFirst = mq[0] & 1ull;
              // v5 holds the lane's 4 phase bits in bits 0..3.
              if (rowOwn && (VEC || q < qs)) {
#pragma unroll
                for (int e = 0; e < 4; ++e) mw.addSplit(cur[r][4 * q + e], mq[e]);
                n1Lane += __builtin_popcount(v5 & 15u);
              } else if (!VEC && rowOwn && q == qs) {
#pragma unroll
                for (int e = 0; e < 4; ++e) mw.addOwned(cur[r][4 * q + e], mq[e], pm[e]);
                n1Lane += __builtin_popcount(v5 & ob);
              }
            }
            if (rowOwn) {
              mw.nAll += (uint32_t)rowLen;
              if (haloLast) {
                mw.n1 += mw.add(hc[r], hb, ballot(lane == 0));
                mw.nAll += 1;
              }
            }
          } else {
            uint64_t m[NG][4];
#pragma unroll
            for (int q = 0; q < NG; ++q)
#pragma unroll
              for (int e = 0; e < 4; ++e) m[q][e] = relMask<M>(cur[r][4 * q + e], thr);
#pragma unroll
            for (int q = 0; q < NG; ++q) {
              const uint64_t nextFirst = q < NG - 1 ? (m[q + 1][0] & 1ull) : hb;
              uint32_t v5 = laneBit((m[q][0] >> 1) | (nextFirst << 63)) ? 1u : 0u;
#pragma unroll
              for (int e = 3; e >= 0; --e) v5 = dblAddLaneBit(v5, m[q][e]);
              P[r][q] = ((v5 * 0x204081u) & 0x03030303u) & keep[q];
            }
          }
        }
        if (z < z0 + nZ) loadPlane((long long)(z + 1) * sz, cur, hc);
#pragma unroll
        for (int r = 0; r < TY; ++r) {
#pragma unroll
          for (int q = 0; q < NG; ++q) {
            const uint32_t Hc = P[r][q] | (P[r + 1][q] << 2);
            if (z > z0 && r < nY) {
              tally.add<4>(Hp[r][q] | (Hc << 4), hist);
            }
            Hp[r][q] = Hc;
          }
        }
      }
    };
    if (vec)
      walk(std::true_type{});
    else
      walk(std::false_type{});
  }
  flushHistogram(hist, &cnt, tally, cells, bpfd);
  if constexpr (MOM) {
    MomAcc acc;
    acc.s0 = mw.s0;
    acc.q0 = mw.q0;
    acc.s1 = mw.s1;
    acc.q1 = mw.q1;
    momBlockPartial(acc, part, (lane == 0 ? (double)(mw.nAll - mw.n1) : 0.0) - (double)n1Lane,
                    (lane == 0 ? (double)mw.n1 : 0.0) + (double)n1Lane);
  }
}

// ---------------------------------------------------------------------------
// Histogram of a PBM, 16 voxels per lane (voxelsPerLane 16). Lane l holds
// voxels x0 + 16 l .. x0 + 16 l + 15 of a 1024-voxel segment: half of PBM
// word (x0 >> 5) + (l >> 1). It reads that word and the next one (both
// clamped to the row) and one funnel shift leaves its 16 voxels and the voxel
// right of them in bits 0..16. Each group of 4 cells takes a 5-bit window;
// one multiply spreads it into the group's byte-pair word, as in kFused16.
// One wave per workgroup, persistent over the tiles: no LDS exchange and no
// barrier.
// ---------------------------------------------------------------------------
template <int TY>
__global__ __launch_bounds__(kWave) void kBPFD16(const uint32_t* __restrict__ pW, int X, int Y, int Z,
                                                 uint32_t rowWS, uint32_t planeWS, TileGrid g,
                                                 unsigned long long* __restrict__ bpfd) {
  __shared__ uint32_t hist[256];
  __shared__ uint32_t cnt;
  const int lane = threadIdx.x;
  clearHistogram(hist, &cnt);
  Tally tally;
  uint32_t cells = 0;
  for (int t = blockIdx.x; t < g.chunk * 8; t += gridDim.x) {
    const int L = (t & 7) * g.chunk + (t >> 3);
    if (L >= g.nTiles) continue;
    const int xt = L % g.nXT;
    const int rr = L / g.nXT;
    const int yr = rr % g.nYR;
    const int zr = rr / g.nYR;
    const int x0 = xt * (kWave * 16);
    const int y0 = yr * TY;
    const int nY = min(TY, Y - 1 - y0);
    const int z0 = zr * g.TZ;
    const int nZ = min(g.TZ, Z - 1 - z0);
    // cells that exist: x < X - 1 (cell x0 + 16 lane + 4 q + k is byte k of group q)
    uint32_t keep[4], nCells = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      keep[q] = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const bool ok = x0 + 16 * lane + 4 * q + k < X - 1;
        keep[q] |= ok ? (0xFFu << (8 * k)) : 0u;
        nCells += ok ? 1u : 0u;
      }
    }
    cells += waveSum(nCells) * (uint32_t)nY * (uint32_t)nZ;
    // The next word is needed only when the lane's last cell exists, and
    // then it lies inside the row; clamping keeps every read in the row.
    const uint32_t wlo = min((uint32_t)(x0 >> 5) + (uint32_t)(lane >> 1), rowWS - 1u);
    const uint32_t whi = min(wlo + 1u, rowWS - 1u);
    const uint32_t sh = 16u * (uint32_t)(lane & 1);
    size_t yo[TY + 1];
#pragma unroll
    for (int r = 0; r <= TY; ++r) yo[r] = (size_t)(y0 + min(r, nY)) * rowWS;
    auto loadPlane = [&](size_t zo, uint32_t (&lo)[TY + 1], uint32_t (&hi)[TY + 1]) {
#pragma unroll
      for (int r = 0; r <= TY; ++r) {
        const uint32_t* row = pW + zo + yo[r];
        lo[r] = row[wlo];
        hi[r] = row[whi];
      }
    };
    uint32_t lo[TY + 1], hi[TY + 1];
    loadPlane((size_t)z0 * planeWS, lo, hi);
    uint32_t Hp[TY][4];
    for (int z = z0; z <= z0 + nZ; ++z) {
      uint32_t P[TY + 1][4];
#pragma unroll
      for (int r = 0; r <= TY; ++r) {
        const uint32_t b = __builtin_amdgcn_alignbit(hi[r], lo[r], sh);
#pragma unroll
        for (int q = 0; q < 4; ++q) P[r][q] = (((b >> (4 * q)) & 31u) * 0x204081u & 0x03030303u) & keep[q];
      }
      if (z < z0 + nZ) loadPlane((size_t)(z + 1) * planeWS, lo, hi);
#pragma unroll
      for (int r = 0; r < TY; ++r) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t Hc = P[r][q] | (P[r + 1][q] << 2);
          if (z > z0 && r < nY) tally.add<4>(Hp[r][q] | (Hc << 4), hist);
          Hp[r][q] = Hc;
        }
      }
    }
  }
  flushHistogram(hist, &cnt, tally, cells, bpfd);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline long long ceilDiv(long long a, long long b) { return (a + b - 1) / b; }

// Launchers defined in other parts of this file.
hipError_t launchMomentsReduce(const double* part, long long nParts, double* mom, hipStream_t s);  // part 1
hipError_t launchFusedF32(const FieldView& fv, float thr, uint32_t m, unsigned long long* bpfd, const Tiling& tl,
                          hipStream_t s);  // part 2
hipError_t launchFusedF64(const FieldView& fv, double thr, uint32_t m, unsigned long long* bpfd, const Tiling& tl,
                          hipStream_t s);  // part 3
hipError_t launchFused16(const FieldView& fv, float thr, uint32_t m, unsigned long long* bpfd, const Tiling& tl,
                         hipStream_t s, LaunchInfo* info);  // part 4
unsigned fused16Slots(int V, int X);      // part 4: resident slots of kFused16 with V voxels per lane
unsigned fused16MomSlots(int V);          // part 8: the same for its moments form
// part 8: kFused16 with moments (4 row pairs); with nParts set it only
// reports its workgroup count (= moment partials) and launches nothing.
hipError_t launchFused16Moments(const FieldView& fv, float thr, uint32_t m, unsigned long long* bpfd, double* part,
                                const Tiling& tl, hipStream_t s, long long* nParts);

// Kernel view of a field description; unused table slots repeat field 0.
template <typename T>
static FieldArgs<T> fieldArgs(const FieldView& fv) {
  FieldArgs<T> a;
  for (int i = 0; i < 4; ++i) a.p[i] = (const T*)fv.p[i < fv.nF ? i : 0];
  a.nF = fv.nF;
  a.sx = fv.sx;
  a.sy = fv.sy;
  a.sz = fv.sz;
  return a;
}

static inline TileGrid makeGrid(int X, int Y, int Z, const Tiling& tl) {
  TileGrid g;
  g.nSegX = (int)ceilDiv(X - 1, (long long)kWave * tl.V);
  // As few x-tiles as tl.ZW waves allow, with the segments spread evenly
  // over them (17 segments: 2 tiles of 9 waves, not 16 + 1).
  g.nXT = (int)ceilDiv(g.nSegX, std::min<long long>(tl.ZW, g.nSegX));
  g.waves = (int)ceilDiv(g.nSegX, g.nXT);
  g.TZ = tl.TP;
  g.nYR = (int)ceilDiv(Y - 1, tl.Ty);
  g.nZR = (int)ceilDiv(Z - 1, tl.TP);
  g.nTiles = g.nXT * g.nYR * g.nZR;
  g.chunk = (int)ceilDiv(g.nTiles, 8);
  return g;
}

// Fused launches of more workgroups than this store per-workgroup histograms
// and add them up with kHistRowsReduce.
[[maybe_unused]] constexpr long long kHistRowsMin = 4096;

// Waves per workgroup of the narrow PBM histogram, each on a tile of its own.
#ifndef MKF_NARROW_WAVES
#define MKF_NARROW_WAVES 8
#endif
[[maybe_unused]] constexpr int kNarrowWaves = MKF_NARROW_WAVES;

// Words per row group of the narrow PBM histogram (kBPFD4R) for a map of
// rowWS-word rows, or 0 where the map takes kBPFD4 (rows of more than 8
// words, or fewer row pairs than one group's stack).
static inline int narrowGroupWords(uint32_t rowWS, int Y, const Tiling& tl) {
  if (tl.V != 4 || rowWS < 1 || rowWS > 8 || Y - 1 < tl.Ty) return 0;
  return rowWS == 1 ? 1 : (rowWS == 2 ? 2 : (rowWS <= 4 ? 4 : 8));
}

// Its grid: tiles of 8 / GW stacks of Ty row pairs by TP plane pairs, one per
// wave, tl.ZW waves per workgroup.
static inline TileGrid narrowGrid(int GW, int Y, int Z, const Tiling& tl) {
  TileGrid g;
  g.nSegX = 1;
  g.nXT = 1;
  g.waves = tl.ZW;
  g.TZ = tl.TP;
  g.nYR = (int)ceilDiv(Y - 1, (long long)(8 / GW) * tl.Ty);
  g.nZR = (int)ceilDiv(Z - 1, tl.TP);
  g.nTiles = g.nYR * g.nZR;
  g.chunk = (int)ceilDiv(ceilDiv(g.nTiles, tl.ZW), 8);
  return g;
}

#if MKF_PART(0)
bool tilingSupported(const Tiling& tl, int elemBytes) {
  const bool vOk =
      tl.V == 1 || (elemBytes == 4 && (tl.V == 4 || tl.V == 8 || tl.V == 16)) || (elemBytes == 8 && tl.V == 2);
  const bool tyOk = tl.Ty == 2 || tl.Ty == 4 || (tl.Ty == 8 && tl.V < 8);
  // A workgroup counts its cells in 32 bits: a tile holds at most
  // 4096 voxels x 8 rows x 65536 planes = 2^31 cells.
  return vOk && tyOk && tl.TP >= 1 && tl.TP <= 65536 && tl.ZW >= 1 && tl.ZW <= 16;
}
#endif

// Fused kernel: workgroups as wide as the row (up to 16 waves) and the
// longest slab that still gives the grid a dtype's target number of waves.
// Measured on MI355X at 1024^3 with the interior rows of each stack read by
// non-temporal loads (HIP events, warm clocks, best of 3 interleaved rounds
// of 20 launches):
//   f32: 1 voxel/lane, 4 row pairs, 128 plane pairs   0.700 ms (6.13 TB/s)
//        4 voxels/lane, 4 row pairs, 256 plane pairs  0.696 ms
//        1 voxel/lane, 4 row pairs, 64 plane pairs    0.712 ms
//        1 voxel/lane, 8 row pairs, 128 plane pairs   0.816 ms
//   f64: 2 voxels/lane, 4 row pairs, 512 plane pairs  1.181 ms (7.27 TB/s)
//        1 voxel/lane, 8 row pairs, 512 plane pairs   1.280 ms
//        1 voxel/lane, 8 row pairs, 256 plane pairs   1.293 ms
// The f32 leaders are within 1 % and swap between machines; the library's
// autotuner times them on the caller's device.
// Resident one-wave workgroups of a kernel on the current device, at most
// capPerCU per CU when capPerCU > 0, rounded down to a multiple of 8
// (computed once per kernel, device and cap).
static inline unsigned persistentSlots(const void* kernel, int capPerCU = 0) {
  static std::mutex mu;
  static std::map<std::tuple<const void*, int, int>, unsigned> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find({kernel, dev, capPerCU});
  if (it != cache.end()) return it->second;
  int perCU = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, kWave, 0) != hipSuccess || perCU < 1) perCU = 1;
  if (capPerCU > 0) perCU = std::min(perCU, capPerCU);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 8;
  const unsigned n = std::max(8u, ((unsigned)(perCU * cus)) / 8u * 8u);
  cache[{kernel, dev, capPerCU}] = n;
  return n;
}


// Rows of X voxels fill their segments of seg voxels at least to frac.
static inline bool segmentsFilled(int X, double frac, int seg = kWave * 16) {
  return X >= 2 && (double)(X - 1) >= frac * (double)seg * (double)ceilDiv(X - 1, (long long)seg);
}

// Per-CU cap of the persistent fused kernel with V voxels per lane on rows
// of X voxels: 2 waves per SIMD, or 4 for 8 voxels per lane on rows that
// fill less than 90 % of their 512-voxel segments (their waves hold fewer
// loads in flight). Measured on MI355X, 8 voxels per lane, caps 8 / 16 per CU:
// 300 x 1000 x 3000 0.828 / 0.654 ms, 400 x 1000 x 2500 0.713 / 0.697,
// 512 x 1024 x 2048 0.657 / 0.714.
static inline int fused16PerCU(int V, int X) {
  return V == 8 && !segmentsFilled(X, 0.9, kWave * 8) ? 2 * kFused16PerCU : kFused16PerCU;
}

#if MKF_PART(0)
// Slab depth of the persistent 16-voxel kernels: their waves take tiles
// round robin, so the time goes as the most tiles any wave takes times the
// planes a tile reads (TP + 1, its first plane read twice). The depth with
// the lowest ceil(tiles / slots) * (TP + 1) wins (ties: the deeper slab). At
// 1024^3 on MI355X (2048 slots for the fused kernel) that is 128 plane pairs
// (one tile per wave): 0.651 ms; 64: 0.653, 32: 0.666, 96: 0.833 (some waves
// take two tiles, most one).
static int persistentDepth(int X, int Y, int Z, int TY, int seg, long long slots) {
  const long long cols = ceilDiv(X - 1, (long long)seg) * ceilDiv(Y - 1, TY);
  int best = 4;
  long long bestCost = -1;
  for (int tp = 4; tp <= 512; ++tp) {
    const long long tiles = cols * ceilDiv(Z - 1, tp);
    const long long cost = ceilDiv(tiles, slots) * (tp + 1);
    if (bestCost < 0 || cost <= bestCost) {
      bestCost = cost;
      best = tp;
    }
  }
  return best;
}

// Workgroup width (waves) for rows of nSeg segments. A row of at most 16
// segments is one workgroup. Longer rows are split into x-tiles of 16, 8 or
// 4 waves (spread evenly over the tiles they need); the split with the
// fewest waves wins, each width weighted by how it fills a CU: waves are
// dealt to 4 SIMDs (a width that is not a multiple of 4 counts 1.25x) and a
// CU holds 32 waves of this kernel (12 waves: 2 workgroups, 24 of 32). Ties
// go to the wider. Measured on MI355X (f32, 1 voxel per lane, 1000 x 1000
// rows of the given segments):
//   17: 5 x 4 waves 0.99 ms; 3 x 6 1.09; 2 x 9 1.25; 16 + 1 1.48
//   24: 3 x 8 0.51; 2 x 12 0.57; 6 x 4 0.58
//   65: 9 x 8 0.737; 17 x 4 0.739; 5 x 13 0.87
//   9, 11, 16: one workgroup 0.50, 0.58, 0.72 against 0.57, 0.63, 0.85+ split
static int pickWaves(long long nSeg) {
  if (nSeg <= 16) return 16;
  int best = 16;
  double bestCost = -1;
  for (int zw : {16, 8, 4}) {
    const long long nXT = ceilDiv(nSeg, zw);
    const long long w = ceilDiv(nSeg, nXT);
    const double perCU = (double)((32 / w) * w);
    double cost = (double)(nXT * w) * 32.0 / perCU;
    if (w % 4 != 0) cost *= 1.25;
    if (bestCost < 0 || cost < bestCost - 1e-9) {
      bestCost = cost;
      best = zw;
    }
  }
  return best;
}

// Slab depth of the fused kernel (kFusedBPFD) for float fields at 4 voxels
// per lane and for double fields: of maxTP, maxTP / 2, ..., 8 plane pairs,
// the depths whose grid holds at least 4096 waves (16 per CU, to hide
// latency), and of those the one with the least work on the busiest CU.
// Workgroups are dealt to the CUs, each costing its plane pairs plus 8 for
// its set-up, first plane and flush: ceil(workgroups / CUs) * (TP + 8). 8
// when no depth gives 4096 waves. Measured on MI355X, floats: 1030 x 722 x
// 722 at 256 plane pairs (543 workgroups, 3 on some CUs, 2 on most) 0.570 ms,
// at 32 0.405; 300 x 669 x 669 at 8 0.320, at 32 0.150; 1030 x 1000 x 1000 and
// 2100 x 1000 x 500 best at 256. Doubles (the earlier rule halved 512 until
// the grid held 4096 waves): 2000 x 200 x 300 at 128 0.275, at 32 0.166;
// 1000 x 366 x 366 at 128 0.268, at 32 0.192; 1024 x 362 x 362 at 64 0.204,
// at 32 0.187; 1024^3 best at 512 (1.18), 700 x 500 x 500 at 128 (0.221).
static int balancedDepth(int X, int Y, int Z, Tiling tl, int maxTP) {
  int cus = 0, dev = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
  int best = 8;
  long long bestCost = -1;
  for (int tp = maxTP; tp >= 8; tp >>= 1) {
    tl.TP = tp;
    const TileGrid g = makeGrid(X, Y, Z, tl);
    if ((long long)g.nTiles * g.waves < 4096) continue;
    const long long cost = ceilDiv(g.nTiles, (long long)cus) * (tp + 8);
    if (bestCost < 0 || cost < bestCost) {
      bestCost = cost;
      best = tp;
    }
  }
  return best;
}

// At least 2 workgroups per CU for row stacks of 4 row pairs and slabs of
// 32 plane pairs (one workgroup per row stack: rows of up to 16 segments).
static bool grid32Filled(int X, int Y, int Z) {
  int cus = 0, dev = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
  (void)X;
  return ceilDiv(Y - 1, 4) * ceilDiv(Z - 1, 32) >= 2LL * cus;
}

// Plane pairs each wave of a persistent fused tiling walks: the tiles it
// takes times their depth.
static long long persistentWalk(int X, int Y, int Z, const Tiling& tl) {
  const long long slots = fused16Slots(tl.V, X);
  const long long tiles =
      ceilDiv(X - 1, (long long)kWave * tl.V) * ceilDiv(Y - 1, tl.Ty) * ceilDiv(Z - 1, tl.TP);
  return ceilDiv(tiles, slots) * tl.TP;
}

// The persistent fused kernel with V = 8 or 16 voxels per lane, 4 row pairs
// and the slab depth of the cost model (any float field with X >= 4).
Tiling choose16Tiling(int X, int Y, int Z, int V, int quarters) {
  Tiling tl;
  tl.V = V == 8 ? 8 : 16;
  tl.Ty = 4;
  tl.ZW = 16;
  const long long slots = std::max<long long>(1, (long long)fused16Slots(tl.V, X) * quarters / 4);
  tl.TP = persistentDepth(X, Y, Z, 4, kWave * tl.V, slots);
  return tl;
}

// Float fields at 4 voxels per lane: workgroups as wide as pickWaves gives,
// slab depth from the busiest-CU model over 256, 128, ... plane pairs.
Tiling choose4Tiling(int X, int Y, int Z) {
  X = std::max(X, 2);
  Tiling tl;
  tl.V = 4;
  tl.Ty = 4;
  tl.ZW = pickWaves(ceilDiv(X - 1, (long long)kWave * 4));
  tl.TP = balancedDepth(X, Y, Z, tl, 256);
  return tl;
}

Tiling chooseTiling(int X, int Y, int Z, int elemBytes) {
  X = std::max(X, 2);  // a row without cells takes the tiling of one cell
  Tiling tl;
  long long targetWaves;
  // Float fields (HIP events, x 1000 x 1000 unless given, best of 3 rounds
  // of 10-20 launches on MI355X):
  // * rows that fill their 1024-voxel segments at least 80 %, or their one
  //   segment at least 68 %: the persistent 16-voxel kernel. X = 1024
  //   (x 1024 x 1024): 0.669 ms against 0.714 for 4 voxels per lane; 1000:
  //   0.686 against 0.714; 900: 0.591 against 0.633; 1900 x 1000 x 560: 0.722
  //   against 0.888. One segment (GRF, HIP events, best of 3 means of 10):
  //   800: 0.500 against 0.553; 769: 0.495 against 0.575; 700: 0.488 against
  //   0.505; 650: 0.487 against 0.472; 600: 0.486 against 0.444 (the
  //   16-voxel kernel's time barely depends on the fill).
  // * rows that fill their 512-voxel segments at least 90 %: the 8-voxel
  //   kernel. 1536 x 1000 x 700: 0.695 against 0.762 (16) and 0.759 (4).
  // * rows of 257 to 513 voxels: the 8-voxel kernel. 280 x 1000 x 3600:
  //   0.817 against 0.891 (1) and 1.027 (4); 300 x 1000 x 3000: 0.653 against
  //   0.750 and 0.897; 400 x 1000 x 2500: 0.697 against 0.859 and 0.849.
  // * the persistent kernels only where each wave walks at least 48 plane
  //   pairs: on smaller fields the pipeline's start and end are not paid for.
  //   At 2^27 voxels 4 voxels per lane win (1000 x 366 x 366: 0.104 against
  //   0.119 (16); 1536 x 296 x 296: 0.099 against 0.111 (8); waves walk 17 and
  //   33 plane pairs), at 2^29 the persistent kernels (1000 x 733 x 733: 0.363
  //   against 0.402; 1536 x 591 x 591: 0.357 against 0.419; 67 and 132).
  // * other rows longer than 513 voxels: 4 voxels per lane, slab depth from
  //   balancedDepth (with the per-word tally of the field walk: 1030 0.686 ms,
  //   2100 x 1000 x 500 0.650). 1030: 0.736 against 1.008 (1 voxel), 1.105
  //   (16) and 0.792 (8);
  //   1240: 0.854 against 1.097 (16) and 0.871 (8); 640: 0.454 against 0.549
  //   (16); 576: 0.420 against 0.510 (1); 2100 x 1000 x 500: 0.694 against
  //   0.818 (16); 700 x 500 x 500: 0.124 against 0.167 (16); 2000 x 200 x 300:
  //   0.087 against 0.111 (16).
  // * rows of 258..513 voxels filling 72 % of their 256-voxel segments: 4
  //   voxels per lane at 32 plane pairs (below).
  // * other rows: 1 voxel per lane. 200 x 1000 x 5000: 0.797 against 0.977
  //   (4) and 1.028 (8); 256 x 1000 x 4000: 0.721 against 0.841;
  //   300 x 669 x 669: 0.133 against 0.149 (4).
  if (elemBytes == 4) {
    if (segmentsFilled(X, 0.8) || (X - 1 <= kWave * 16 && segmentsFilled(X, 0.68))) {
      const Tiling t16 = choose16Tiling(X, Y, Z, 16);
      if (persistentWalk(X, Y, Z, t16) >= 48) return t16;
    }
    if (segmentsFilled(X, 0.9, kWave * 8) || (X - 1 > kWave * 4 && X - 1 <= kWave * 8)) {
      const Tiling t8 = choose16Tiling(X, Y, Z, 8);
      if (persistentWalk(X, Y, Z, t8) >= 48) return t8;
    }
  }
  if (elemBytes >= 8) {
    // 2 voxels per lane (one 16-byte load per row). Rows whose last
    // 128-voxel segment is partial run the kernel form with a second walk for
    // the partial waves where they span more than 3 segments and the slabs
    // are at least 32 plane pairs deep; against 1 voxel per lane (GRF, HIP
    // events): 1000^3 1.152-1.175 ms (256-512 plane pairs) against 1.347;
    // 1030 x 1000 x 1000 1.282 against 1.907; 600 x 400 x 400 0.134 (32)
    // against 0.147; 400^3 0.097 (32) against 0.100. Otherwise 1 voxel per
    // lane: 250 x 300 x 350 0.091 ms (2 voxels, 8 plane pairs) against 0.061;
    // 386 x 300 x 300 0.068 (16) against 0.056; 450 x 300 x 300 0.072 (16)
    // against 0.063. (Before the second walk, partial waves loaded voxel by
    // voxel: 1000^3 2.13 ms.)
    tl.V = 2;
    tl.Ty = 4;
    tl.ZW = pickWaves(ceilDiv(X - 1, (long long)kWave * tl.V));
    tl.TP = balancedDepth(X, Y, Z, tl, 512);
    if (X % (kWave * 2) > 1 && (X - 1 <= 3 * kWave * 2 || tl.TP < 32)) {
      tl.V = 1;
      tl.ZW = pickWaves(ceilDiv(X - 1, (long long)kWave));
      tl.TP = balancedDepth(X, Y, Z, tl, 512);
    }
    return tl;
  } else if (X - 1 > kWave * 8) {
    return choose4Tiling(X, Y, Z);
  } else if (X - 1 > kWave * 4 && segmentsFilled(X, 0.72, kWave * 4) && grid32Filled(X, Y, Z)) {
    // Rows of 258..513 voxels that fill their two 256-voxel segments at
    // least 72 %, on grids of at least 2 workgroups per CU at 32 plane
    // pairs: 4 voxels per lane at 32 plane pairs (2-wave workgroups, whose
    // grids the busiest-CU model sizes poorly). Against 1 voxel per lane:
    // 400 x 819 x 819 0.214 ms against 0.251; 400 x 290 x 290 (730
    // workgroups) 0.054 against 0.073; 390 x 1000 x 700 0.219 against 0.258;
    // 500^3 0.104 against 0.117; 376 x 363 x 212 (637) 0.052 against 0.064.
    // It loses below 72 % (350 x 600 x 600, 68 %: 0.137 against 0.122;
    // 300 x 669 x 669, 58 %: 0.148 against 0.133) and on smaller grids
    // (383 x 92 x 375, 276 workgroups: 0.048 against 0.034; 370 x 321 x 57,
    // 160: 0.040 against 0.022).
    tl.V = 4;
    tl.Ty = 4;
    tl.ZW = pickWaves(ceilDiv(X - 1, (long long)kWave * 4));
    tl.TP = 32;
    return tl;
  } else {
    tl.V = 1;
    tl.Ty = 4;
    tl.TP = 128;
    targetWaves = 16384;
  }
  tl.ZW = pickWaves(ceilDiv(X - 1, (long long)kWave * tl.V));
  while (tl.TP > 8) {
    const TileGrid g = makeGrid(X, Y, Z, tl);
    if ((long long)g.nTiles * g.waves >= targetWaves) break;
    tl.TP >>= 1;
  }
  return tl;
}

// Fused kernel with phase moments: the moments add ~9 instructions per voxel,
// so the kernel needs more voxels per lane and more tiles in flight than the
// plain one. Measured on MI355X at 1024^3 (best of 3 rounds of 20 launches):
//   f32: 4 voxels/lane, 64 plane pairs  0.963 ms (128: 1.015, 256: 1.123;
//        1 voxel/lane, 128: 1.523); plain histogram + moments-only map
//        pass: 0.715 + 0.670 ms
//   f64: 2 voxels/lane, 128 plane pairs 1.595 ms (512: 1.609); two passes
//        1.173 + 1.245 ms
// Float fields take the persistent kernel's moments form where each of its
// waves walks at least 48 plane pairs: 16 voxels per lane for rows of whole
// 1024-voxel segments (a partial one costs the 16-voxel form more: 1000 x 733
// x 733 0.501 ms against 0.448 at 8, 1000 x 600 x 600 0.348 against 0.298,
// 1000^3 0.795 against 0.804), else 8 voxels per lane with 32 plane
// pairs for rows that fill their 512-voxel segments at least half (narrower
// rows idle too many lanes: X = 64, 100, 200 took 4.55, 2.60, 1.44 ms at 8
// against 3.74, 1.99, 1.24 at 4; X = 300 1.02 against 1.23). The 16-voxel
// form takes 64 plane pairs, halved while the grid holds fewer tiles than
// resident workgroups; the depth matters more than the busiest-CU model says
// (1024^3: 48 0.858 ms, 56 0.795, 64 0.714, 72 0.787, 86 (one tile per slot)
// 0.752; 2048x1024x512: 32 0.709, 64 0.719, 86 0.757; 1025x1000x1000: 64
// 0.724, 86 0.752; 1024x700x700: 32 0.374, 64 0.424).
// GRF fields, HIP events, warm clocks (the 16-voxel column with 3 waves per
// SIMD; the others from the first, 2-wave build of the kernel):
//   1024^3         16: 0.714 ms (64 plane pairs)   8: 0.880  4: 0.861
//   1000^3          8: 0.802 (16: 0.788)           4: 0.846
//   2049x600x600   16: 0.596 (32)  8: 0.617        4: 0.861
//   1536x1000x700   8: 0.791 (16 plane pairs: 0.803) 4: 1.240
//   1030x1000x1000  8: 0.983                       4: 1.662
//   300x1000x3000   8: 1.020 (64: 1.109)           4: 1.227
//   2100x1000x500   8: 0.874 (16: 1.181)           4: 1.562
//   1030x722x722    8: 0.541                       4: 0.890
// Smaller fields walk too few planes per wave to fill the pipeline:
//   1000x366x366    8: 0.184  4: 0.151;  700x500x500  8: 0.227  4: 0.213;
//   2000x200x300    8: 0.159  4: 0.141;  256^3        8: 0.091  4: 0.070.
// Elsewhere 4 voxels per lane with 32 plane pairs, or 16 where that gives
// fewer than 8 workgroups per CU (best of 3 rounds of 10 launches):
// 1000 x 733 x 733 0.478 (64: 0.507); 300 x 669 x 669 0.256 (8, the earlier
// choice: 0.418); at 16: 1000 x 366 x 366 0.153 (32: 0.166); 700 x 500 x 500
// 0.216 (0.218); 2000 x 200 x 300 0.144 (0.142).
static long long momTiles(int X, int Y, int Z, const Tiling& tl) {
  return ceilDiv(X - 1, (long long)kWave * tl.V) * ceilDiv(Y - 1, tl.Ty) * ceilDiv(Z - 1, tl.TP);
}

static long long momWalk(int X, int Y, int Z, const Tiling& tl) {
  return ceilDiv(momTiles(X, Y, Z, tl), (long long)fused16MomSlots(tl.V)) * tl.TP;
}

Tiling chooseMomTiling(int X, int Y, int Z, int elemBytes) {
  X = std::max(X, 2);  // a row without cells takes the tiling of one cell
  Tiling tl;
  tl.V = elemBytes >= 8 ? 2 : 4;
  tl.Ty = 4;
  tl.TP = elemBytes >= 8 ? 128 : 64;
  tl.ZW = 16;
  if (elemBytes == 4 && X >= 4 && Y >= 2 && Z >= 2) {
    Tiling tp = tl;
    if (X >= kWave * 16 && X % (kWave * 16) <= 1) {
      tp.V = 16;
      tp.TP = 64;
      while (tp.TP > 16 && momTiles(X, Y, Z, tp) < (long long)fused16MomSlots(16)) tp.TP >>= 1;
      if (momWalk(X, Y, Z, tp) >= 48) return tp;
    }
    tp.V = 8;
    tp.TP = 32;
    if (segmentsFilled(X, 0.5, kWave * 8) && momWalk(X, Y, Z, tp) >= 48) return tp;
  }
  if (elemBytes == 4) {
    int cus = 0, dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    tl.TP = 32;
    if ((long long)makeGrid(X, Y, Z, tl).nTiles < 8LL * cus) tl.TP = 16;
    return tl;
  }
  // Double fields: the busiest-CU model from 128 plane pairs. Against the
  // earlier rule (128 halved until the grid held 16K waves): 2000 x 200 x 300
  // 0.190 ms (32) against 0.245; 300 x 669 x 669 0.218 (32) against 0.239;
  // 700 x 500 x 500 0.294 (64) against 0.323; 1024^3 and 1000^3 unchanged
  // (128).
  tl.TP = balancedDepth(X, Y, Z, tl, 128);
  return tl;
}

// Histogram of an existing PBM: kBPFD4 with 4 row pairs.
Tiling chooseBpfdTiling(int X, int Y, int Z) {
  X = std::max(X, 2);  // a row without cells takes the tiling of one cell
  // 4 voxels per lane, 4 row pairs; slabs of 32 plane pairs, or 16 where
  // that gives fewer than 4 workgroups per CU, and deeper (up to 128) while
  // the grid has more than 4096 workgroups. Each workgroup adds its LDS
  // histogram to the device's 256 bins with one atomic per bin, all on the
  // same 256 addresses, which costs about 22 ns per workgroup: at 1024^3 GRF
  // 8192 workgroups (32 plane pairs) took 0.219 ms, the same kernel with its
  // atomics spread over 64 copies of the bins 0.166 ms, and 4096 workgroups
  // (64 plane pairs) 0.175 ms. Other GRF maps (HIP events, 20-launch means),
  // 32 / 64 / 128 plane pairs: 1000^3 0.217 / 0.176 / 0.187 ms; 2048 x 512 x
  // 1024 0.174 / 0.178 / 0.189; 500 x 1000 x 2000 0.369 / 0.240 / 0.202;
  // 300 x 1000 x 3000 0.539 / 0.325 / 0.280. Against the 16-voxel kernel
  // (kBPFD16) with its resident-slot slab depth, at 32 plane pairs: 1000 x
  // 733 x 733 0.141 against 0.194; 1000 x 366 x 366 0.057 against 0.137;
  // 2048 x 256 x 256 (16) 0.044 against 0.138. Rows of at most 8 words take
  // the row-group kernel (kBPFD4R) with kNarrowWaves waves per workgroup.
  int cus = 0, dev = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
  Tiling tl;
  tl.V = 4;
  tl.Ty = 4;
  tl.TP = 32;
  tl.ZW = pickWaves(ceilDiv(X - 1, (long long)kWave * tl.V));
  const int GW = narrowGroupWords((uint32_t)((X + 31) / 32), Y, tl);
  if (GW) tl.ZW = kNarrowWaves;
  auto groups = [&] {
    return GW ? ceilDiv(narrowGrid(GW, Y, Z, tl).nTiles, tl.ZW) : (long long)makeGrid(X, Y, Z, tl).nTiles;
  };
  if (groups() < 4LL * cus) {
    tl.TP = 16;
  } else {
    while (tl.TP < 128 && groups() > 4096) tl.TP *= 2;
  }
  return tl;
}

#endif

#if MKF_PART(1) || MKF_PART(7)
// About 16K waves in all, each streaming every C-th row of one 1024-voxel
// segment, so the waves resident at one time read neighbouring rows and
// write neighbouring map lines. (At 1024^3: runs of consecutive rows per
// wave were 2-3 % slower; 4K, 8K and 32K waves, and 4 rows in flight instead
// of 2, were within 1-4 %.)
struct PbmGrid {
  int nSegX;
  RowSched rs;
  long long nBlocks;
};

// Voxels per lane of the map kernel: rows of at most 256 or 512 voxels take
// the narrowest segment that still holds the whole row (4 or 8 voxels per
// lane: fewer loads past the row's end); longer rows keep 1024-voxel
// segments, whose waves store whole 128-byte map lines. Against 1024-voxel
// segments for every row (f32 GRF, HIP events, identical maps): 200 x 1000
// x 5000 1.46 ms against 1.99; 300 x 1000 x 3000 1.06 against 1.25; 509 x
// 511 x 517 0.122 against 0.138 (f64 alike). Rows of whole map words take
// the plane-row view below instead.
// Splitting a longer row over narrower segments splits its map stores too,
// and they cost more than the idle lanes: 256-voxel segments took 1030 x
// 1000 x 1000 to 1.45 ms against 0.93. The general form (several fields, an
// x stride, rows shorter than one load) keeps 16.
static int pbmVoxels(const FieldView& fv) {
  const int eb = fv.enc & ENC_MASK_NUM;
  if (fv.nF > 1 || fv.sx != 1 || fv.X < 16 / std::max(eb, 1)) return kPbmVoxels;
  if (fv.X <= kWave * 4) return 4;
  if (fv.X <= kWave * 8) return 8;
  return 16;
}

// The map pass on a contiguous planar field whose map rows are exactly its
// rows (X a multiple of 32, rowWS = X / 32, planeWS = Y rowWS) may take each
// plane as one row of X Y voxels: voxel (x, y, z) keeps map bit
// 32 planeWS z + X y + x, and rows narrower than a segment, or ragged ones,
// then fill whole segments (f32 GRF, identical maps and moments: 64 x 4000 x
// 4000 0.77 ms against 4.16; 96 x 2000 x 5000 0.72 against 2.68; 1056 x 1000
// x 1000 0.79 against 0.92). The reference launches such fields collapsed to
// 1-D (src/gridUtil.cuh:12-38, src/binMapCUDA.cuh:282-294). Whole-word rows
// are asked for also when no map is written (moments only), so the moments
// come out the same to the bit either way. Rows of whole 1024-voxel segments
// keep their own shape.
static FieldView pbmView(const FieldView& fv, const uint32_t* pW, uint32_t& rowWS, uint32_t planeWS) {
  const long long XY = (long long)fv.X * fv.Y;
  const bool planar = fv.nF == 1 && fv.sx == 1 && fv.sy == fv.X && fv.sz == XY;
  const bool sameBits =
      fv.X % 32 == 0 && (!pW || ((long long)rowWS * 32 == fv.X && (long long)planeWS * 32 == XY));
  if (!planar || !sameBits || fv.Y < 2 || fv.Z < 1 || fv.X % (kWave * kPbmVoxels) == 0 || XY > (1ll << 30))
    return fv;
  FieldView v = fv;
  v.X = (int)XY;
  v.Y = fv.Z;
  v.Z = 1;
  v.sy = XY;
  v.sz = XY * fv.Z;
  rowWS = (uint32_t)(XY / 32);
  return v;
}

// Row groups of the map kernel: float rows of at most 128, 64 or 32 voxels
// (contiguous rows, one field) take 2, 4 or 8 rows per wave step.
static int pbmGroups(const FieldView& fv) {
  if ((fv.enc & ENC_MASK_NUM) != 4 || fv.nF > 1 || fv.sx != 1 || fv.X < 4 || fv.X > 2 * kWave) return 1;
  return fv.X <= 32 ? 8 : (fv.X <= 64 ? 4 : 2);
}

static PbmGrid pbmGrid(const FieldView& fv) {
  PbmGrid g;
  g.nSegX = (int)ceilDiv(fv.X, kWave * pbmVoxels(fv));
  const int grp = pbmGroups(fv);
  RowSched& rs = g.rs;
  rs.nRows = ceilDiv((long long)fv.Y * fv.Z, grp);  // rows, or row groups
  const long long wantChunks = std::max<long long>(1, (MKF_PBM_WAVE_TARGET + g.nSegX - 1) / g.nSegX);
  rs.rowsPerWave = ceilDiv(rs.nRows, wantChunks);
  const long long nChunks = ceilDiv(rs.nRows, rs.rowsPerWave);
  rs.nChunks = nChunks;
  if (MKF_PBM_INTERLEAVE) {
    rs.chunkStep = 1;
    rs.rowStep = nChunks;
  } else {
    rs.chunkStep = rs.rowsPerWave;
    rs.rowStep = 1;
  }
  rs.dy = (int)(rs.rowStep * grp % fv.Y);
  rs.dz = (int)(rs.rowStep * grp / fv.Y);
  rs.Z = fv.Z;
  g.nBlocks = ceilDiv(g.nSegX * nChunks, kPbmWaves);
  return g;
}

template <typename T, int M, bool MOM>
static hipError_t launchPBMT(const FieldView& fv, T thr, uint32_t* pW, uint32_t rowWS, uint32_t planeWS,
                             double* part, hipStream_t s) {
  const PbmGrid g = pbmGrid(fv);
  const dim3 grid((unsigned)g.nBlocks), block(kPbmWaves * kWave);
  // The general form also takes rows shorter than one 16-byte load.
  if (fv.nF > 1 || fv.sx != 1 || fv.X < (int)(16 / sizeof(T))) {
    hipLaunchKernelGGL((kPBM<T, M, true, MOM>), grid, block, 0, s, fieldArgs<T>(fv), thr, fv.X, fv.Y, g.nSegX, g.rs,
                       pW, rowWS, planeWS, part);
  } else {
    switch (pbmVoxels(fv)) {
      case 16:
        hipLaunchKernelGGL((kPBM<T, M, false, MOM, 16>), grid, block, 0, s, fieldArgs<T>(fv), thr, fv.X, fv.Y,
                           g.nSegX, g.rs, pW, rowWS, planeWS, part);
        break;
      case 8:
        hipLaunchKernelGGL((kPBM<T, M, false, MOM, 8>), grid, block, 0, s, fieldArgs<T>(fv), thr, fv.X, fv.Y,
                           g.nSegX, g.rs, pW, rowWS, planeWS, part);
        break;
      default:
        if constexpr (sizeof(T) == 4) {
          switch (pbmGroups(fv)) {
            case 8:
              hipLaunchKernelGGL((kPBM<T, M, false, MOM, 4, 8>), grid, block, 0, s, fieldArgs<T>(fv), thr, fv.X,
                                 fv.Y, g.nSegX, g.rs, pW, rowWS, planeWS, part);
              return hipGetLastError();
            case 4:
              hipLaunchKernelGGL((kPBM<T, M, false, MOM, 4, 4>), grid, block, 0, s, fieldArgs<T>(fv), thr, fv.X,
                                 fv.Y, g.nSegX, g.rs, pW, rowWS, planeWS, part);
              return hipGetLastError();
            case 2:
              hipLaunchKernelGGL((kPBM<T, M, false, MOM, 4, 2>), grid, block, 0, s, fieldArgs<T>(fv), thr, fv.X,
                                 fv.Y, g.nSegX, g.rs, pW, rowWS, planeWS, part);
              return hipGetLastError();
            default:
              break;
          }
        }
        hipLaunchKernelGGL((kPBM<T, M, false, MOM, 4>), grid, block, 0, s, fieldArgs<T>(fv), thr, fv.X, fv.Y,
                           g.nSegX, g.rs, pW, rowWS, planeWS, part);
        break;
    }
  }
  return hipGetLastError();
}

template <typename T, bool MOM>
static hipError_t dispatchPBM(const FieldView& fv, T thr, uint32_t m, uint32_t* pW, uint32_t rowWS,
                              uint32_t planeWS, double* part, hipStream_t s) {
  switch (m & 7u) {
    case 0: return launchPBMT<T, 0, MOM>(fv, thr, pW, rowWS, planeWS, part, s);
    case 1: return launchPBMT<T, 1, MOM>(fv, thr, pW, rowWS, planeWS, part, s);
    case 2: return launchPBMT<T, 2, MOM>(fv, thr, pW, rowWS, planeWS, part, s);
    case 3: return launchPBMT<T, 3, MOM>(fv, thr, pW, rowWS, planeWS, part, s);
    case 4: return launchPBMT<T, 4, MOM>(fv, thr, pW, rowWS, planeWS, part, s);
    case 5: return launchPBMT<T, 5, MOM>(fv, thr, pW, rowWS, planeWS, part, s);
    case 6: return launchPBMT<T, 6, MOM>(fv, thr, pW, rowWS, planeWS, part, s);
    default: return launchPBMT<T, 7, MOM>(fv, thr, pW, rowWS, planeWS, part, s);
  }
}

// The double-field instantiations, compiled as a part of their own.
hipError_t dispatchPBMf64(const FieldView& fv, double thr, uint32_t m, uint32_t* pW, uint32_t rowWS,
                          uint32_t planeWS, double* part, bool mom, hipStream_t s);
#endif

#if MKF_PART(7)
hipError_t dispatchPBMf64(const FieldView& fv, double thr, uint32_t m, uint32_t* pW, uint32_t rowWS,
                          uint32_t planeWS, double* part, bool mom, hipStream_t s) {
  return mom ? dispatchPBM<double, true>(fv, thr, m, pW, rowWS, planeWS, part, s)
             : dispatchPBM<double, false>(fv, thr, m, pW, rowWS, planeWS, part, s);
}
#endif

#if MKF_PART(1)
hipError_t launchPBM(const FieldView& fv0, const BinMapF64& m, uint32_t* pW, uint32_t rowWS, uint32_t planeWS,
                     hipStream_t s) {
  if (fv0.X < 1 || fv0.Y < 1 || fv0.Z < 1) return hipErrorInvalidValue;
  const FieldView fv = pbmView(fv0, pW, rowWS, planeWS);
  // f32 fields compare against the threshold rounded to float
  // (src/binMapCUDA.cuh:138-139).
  if (fv.enc == ENC_F32) return dispatchPBM<float, false>(fv, (float)m.t[0], m.m, pW, rowWS, planeWS, nullptr, s);
  if (fv.enc == ENC_F64) return dispatchPBMf64(fv, m.t[0], m.m, pW, rowWS, planeWS, nullptr, false, s);
  return hipErrorInvalidValue;
}

long long pbmMomentParts(const FieldView& fv, const uint32_t* pW, uint32_t rowWS, uint32_t planeWS) {
  return pbmGrid(pbmView(fv, pW, rowWS, planeWS)).nBlocks;
}

hipError_t launchPBMMoments(const FieldView& fv0, const BinMapF64& m, uint32_t* pW, uint32_t rowWS,
                            uint32_t planeWS, double* part, double* mom, hipStream_t s) {
  if (fv0.X < 1 || fv0.Y < 1 || fv0.Z < 1) return hipErrorInvalidValue;
  const FieldView fv = pbmView(fv0, pW, rowWS, planeWS);
  // Per-lane counts are 32-bit: at most 16 voxels per row.
  if (pbmGrid(fv).rs.rowsPerWave > (1ll << 27)) return hipErrorInvalidValue;
  hipError_t e;
  if (fv.enc == ENC_F32)
    e = dispatchPBM<float, true>(fv, (float)m.t[0], m.m, pW, rowWS, planeWS, part, s);
  else if (fv.enc == ENC_F64)
    e = dispatchPBMf64(fv, m.t[0], m.m, pW, rowWS, planeWS, part, true, s);
  else
    return hipErrorInvalidValue;
  if (e != hipSuccess) return e;
  return launchMomentsReduce(part, pbmGrid(fv).nBlocks, mom, s);
}

// Adds the per-workgroup histograms rows[0 .. nRows - 1] (256 bins each) to
// bpfd: block j sums rows j, j + gridDim.x, ... bin by bin in 64 bits and
// adds each nonzero sum with one atomic.
__global__ __launch_bounds__(256) void kHistRowsReduce(const uint32_t* __restrict__ rows, long long nRows,
                                                       unsigned long long* __restrict__ bpfd) {
  constexpr int U = 8;  // rows in flight per thread
  unsigned long long a = 0;
  long long r = blockIdx.x;
  const long long step = gridDim.x;
  for (; r + (U - 1) * step < nRows; r += U * step) {
    uint32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = rows[(size_t)(r + u * step) * 256 + threadIdx.x];
#pragma unroll
    for (int u = 0; u < U; ++u) a += v[u];
  }
  for (; r < nRows; r += step) a += rows[(size_t)r * 256 + threadIdx.x];
  if (a) atomicAdd(bpfd + threadIdx.x, a);
}

hipError_t launchHistRowsReduce(const uint32_t* rows, long long nRows, unsigned long long* bpfd, hipStream_t s) {
  if (nRows <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<long long>(nRows, 256);
  hipLaunchKernelGGL(kHistRowsReduce, dim3(blocks), dim3(256), 0, s, rows, nRows, bpfd);
  return hipGetLastError();
}

hipError_t launchMomentsReduce(const double* part, long long nParts, double* mom, hipStream_t s) {
  hipLaunchKernelGGL(kMomentsReduce, dim3(1), dim3(256), 0, s, part, nParts, mom);
  return hipGetLastError();
}
#endif

template <typename T, int M, int V, int TY>
static hipError_t launchFusedT(const FieldView& fv, T thr, unsigned long long* bpfd, const TileGrid& g,
                               uint32_t* rows, hipStream_t s) {
  // Several fields or an x stride take the general form, one voxel per lane.
  const dim3 grid((unsigned)g.chunk * 8u), block(g.waves * kWave);
  if (fv.nF > 1 || fv.sx != 1) {
    if constexpr (V == 1)
      hipLaunchKernelGGL((kFusedBPFD<T, M, 1, TY, true>), grid, block, 0, s, fieldArgs<T>(fv), thr, fv.X, fv.Y,
                         fv.Z, g, bpfd, nullptr, rows);
    else
      return hipErrorInvalidValue;
  } else if (V == 2 && fv.X % (kWave * V) > 1) {
    // Rows whose last segment is partial: the form with a second walk for
    // the partial waves (PW).
    hipLaunchKernelGGL((kFusedBPFD<T, M, V, TY, false, false, V == 2>), grid, block, 0, s, fieldArgs<T>(fv), thr,
                       fv.X, fv.Y, fv.Z, g, bpfd, nullptr, rows);
  } else {
    hipLaunchKernelGGL((kFusedBPFD<T, M, V, TY, false>), grid, block, 0, s, fieldArgs<T>(fv), thr, fv.X, fv.Y,
                       fv.Z, g, bpfd, nullptr, rows);
  }
  return hipGetLastError();
}

template <typename T, int V, int TY>
static hipError_t dispatchFusedM(const FieldView& fv, T thr, uint32_t m, unsigned long long* bpfd,
                                 const TileGrid& g, uint32_t* rows, hipStream_t s) {
  switch (m & 7u) {
    case 0: return launchFusedT<T, 0, V, TY>(fv, thr, bpfd, g, rows, s);
    case 1: return launchFusedT<T, 1, V, TY>(fv, thr, bpfd, g, rows, s);
    case 2: return launchFusedT<T, 2, V, TY>(fv, thr, bpfd, g, rows, s);
    case 3: return launchFusedT<T, 3, V, TY>(fv, thr, bpfd, g, rows, s);
    case 4: return launchFusedT<T, 4, V, TY>(fv, thr, bpfd, g, rows, s);
    case 5: return launchFusedT<T, 5, V, TY>(fv, thr, bpfd, g, rows, s);
    case 6: return launchFusedT<T, 6, V, TY>(fv, thr, bpfd, g, rows, s);
    default: return launchFusedT<T, 7, V, TY>(fv, thr, bpfd, g, rows, s);
  }
}

template <typename T, int V>
static hipError_t dispatchFusedTY(const FieldView& fv, T thr, uint32_t m, unsigned long long* bpfd,
                                  const TileGrid& g, int Ty, uint32_t* rows, hipStream_t s) {
  switch (Ty) {
    case 2: return dispatchFusedM<T, V, 2>(fv, thr, m, bpfd, g, rows, s);
    case 4: return dispatchFusedM<T, V, 4>(fv, thr, m, bpfd, g, rows, s);
    default: return dispatchFusedM<T, V, 8>(fv, thr, m, bpfd, g, rows, s);
  }
}

// The effective tiling of the workgroup kernels for fv: several fields, an x
// stride, V >= 8 (the persistent kernels' widths) or rows shorter than V run
// one voxel per lane (the general form gathers one voxel per lane).
static inline Tiling groupTiling(const FieldView& fv, const Tiling& tl0) {
  Tiling tl = tl0;
  if (fv.nF > 1 || fv.sx != 1 || tl.V >= 8 || fv.X < tl.V) tl.V = 1;
  return tl;
}

static inline void fillInfo(LaunchInfo* info, const TileGrid& g, long long groups, const Tiling& tl) {
  info->tiles = g.nTiles;
  info->workgroups = g.nTiles > 0 ? groups : 0;
  info->wavesPerGroup = g.waves;
  info->used = tl;
}

template <typename T, int VW>
static hipError_t dispatchFused(const FieldView& fv, T thr, uint32_t m, unsigned long long* bpfd, const Tiling& tl0,
                                hipStream_t s) {
  const Tiling tl = groupTiling(fv, tl0);
  const TileGrid g = makeGrid(fv.X, fv.Y, fv.Z, tl);
  if (g.nTiles <= 0) return hipSuccess;
  // Grids of many workgroups (narrow rows: one wave per workgroup) store
  // per-workgroup histograms for kHistRowsReduce rather than each adding its
  // 256 bins to the same device bins (about 22 ns of atomics per workgroup).
  const long long nGroups = (long long)g.chunk * 8;
  uint32_t* rows = nullptr;
  if (nGroups > kHistRowsMin && hipMallocAsync((void**)&rows, (size_t)nGroups * 256 * sizeof(uint32_t), s) != hipSuccess) {
    (void)hipGetLastError();
    rows = nullptr;  // the atomic flush needs no scratch
  }
  hipError_t e = tl.V == 1 ? dispatchFusedTY<T, 1>(fv, thr, m, bpfd, g, tl.Ty, rows, s)
                           : dispatchFusedTY<T, VW>(fv, thr, m, bpfd, g, tl.Ty, rows, s);
  if (rows) {
    if (e == hipSuccess) e = launchHistRowsReduce(rows, nGroups, bpfd, s);
    const hipError_t ef = hipFreeAsync(rows, s);
    if (e == hipSuccess) e = ef;
  }
  return e;
}

#if MKF_PART(2)
hipError_t launchFusedF32(const FieldView& fv, float thr, uint32_t m, unsigned long long* bpfd, const Tiling& tl,
                          hipStream_t s) {
  return dispatchFused<float, 4>(fv, thr, m, bpfd, tl, s);
}
#endif

#if MKF_PART(3)
hipError_t launchFusedF64(const FieldView& fv, double thr, uint32_t m, unsigned long long* bpfd, const Tiling& tl,
                          hipStream_t s) {
  return dispatchFused<double, 2>(fv, thr, m, bpfd, tl, s);
}
#endif

// Tiles of the persistent 8- and 16-voxel kernels: 64 V voxels x Ty row
// pairs x TP plane pairs.
static inline TileGrid grid16(int X, int Y, int Z, const Tiling& tl) {
  TileGrid g{};
  g.nSegX = (int)ceilDiv(X - 1, (long long)kWave * tl.V);
  g.waves = 1;
  g.nXT = g.nSegX;
  g.TZ = tl.TP;
  g.nYR = (int)ceilDiv(Y - 1, tl.Ty);
  g.nZR = (int)ceilDiv(Z - 1, tl.TP);
  g.nTiles = g.nXT * g.nYR * g.nZR;
  g.chunk = (int)ceilDiv(g.nTiles, 8);
  return g;
}

// One wave per resident slot of the kernel (rounded down to a multiple of 8
// XCDs), no more than there are tiles (rounded up to 8).
static inline unsigned blocks16(const TileGrid& g, const Tiling& tl, const void* kernel, int capPerCU) {
  unsigned nBlocks = std::min<unsigned>(persistentSlots(kernel, capPerCU), (unsigned)g.chunk * 8u);
  // A workgroup's histogram and cell count are 32-bit: no workgroup takes
  // more than 2^31 cells (only fields of far more than 2^40 cells need more
  // workgroups than slots for that).
  const long long tileCells = (long long)kWave * tl.V * tl.Ty * tl.TP;
  const long long maxTiles = std::max(1LL, (1LL << 31) / tileCells);
  const long long need = ceilDiv(ceilDiv((long long)g.chunk * 8, maxTiles), 8) * 8;
  if (need > (long long)nBlocks) nBlocks = (unsigned)std::min<long long>(need, (long long)g.chunk * 8);
  return nBlocks;
}

#if MKF_PART(4)
unsigned fused16Slots(int V, int X) {
  const void* k = V == 8 ? (const void*)kFused16<6, 4, 2> : (const void*)kFused16<6, 4, 4>;
  return persistentSlots(k, fused16PerCU(V, X));
}

template <int M, int TY, int NG>
static hipError_t launchFused16K(const FieldView& fv, float thr, unsigned long long* bpfd, const Tiling& tl,
                                 hipStream_t s, LaunchInfo* info) {
  const TileGrid g = grid16(fv.X, fv.Y, fv.Z, tl);
  if (g.nTiles <= 0 && !info) return hipSuccess;
  const unsigned nBlocks = blocks16(g, tl, (const void*)kFused16<M, TY, NG>, fused16PerCU(4 * NG, fv.X));
  if (info) {
    fillInfo(info, g, nBlocks, tl);
    return hipSuccess;
  }
  hipLaunchKernelGGL((kFused16<M, TY, NG>), dim3(nBlocks), dim3(kWave), 0, s, (const float*)fv.p[0], thr, fv.X,
                     fv.Y, fv.Z, fv.sy, fv.sz, g, bpfd);
  return hipGetLastError();
}

template <int M>
static hipError_t launchFused16M(const FieldView& fv, float thr, unsigned long long* bpfd, const Tiling& tl,
                                 hipStream_t s, LaunchInfo* info) {
  if (tl.V == 8)
    return tl.Ty == 2 ? launchFused16K<M, 2, 2>(fv, thr, bpfd, tl, s, info)
                      : launchFused16K<M, 4, 2>(fv, thr, bpfd, tl, s, info);
  return tl.Ty == 2 ? launchFused16K<M, 2, 4>(fv, thr, bpfd, tl, s, info)
                    : launchFused16K<M, 4, 4>(fv, thr, bpfd, tl, s, info);
}

hipError_t launchFused16(const FieldView& fv, float thr, uint32_t m, unsigned long long* bpfd, const Tiling& tl,
                         hipStream_t s, LaunchInfo* info) {
  switch (m & 7u) {
    case 0: return launchFused16M<0>(fv, thr, bpfd, tl, s, info);
    case 1: return launchFused16M<1>(fv, thr, bpfd, tl, s, info);
    case 2: return launchFused16M<2>(fv, thr, bpfd, tl, s, info);
    case 3: return launchFused16M<3>(fv, thr, bpfd, tl, s, info);
    case 4: return launchFused16M<4>(fv, thr, bpfd, tl, s, info);
    case 5: return launchFused16M<5>(fv, thr, bpfd, tl, s, info);
    case 6: return launchFused16M<6>(fv, thr, bpfd, tl, s, info);
    default: return launchFused16M<7>(fv, thr, bpfd, tl, s, info);
  }
}

#endif

#if MKF_PART(0)
hipError_t launchFused(const FieldView& fv, const BinMapF64& m, unsigned long long* bpfd, const Tiling& tl,
                       hipStream_t s, LaunchInfo* info) {
  if (info) *info = LaunchInfo{0, 0, 0, tl};
  if (fv.X < 2 || fv.Y < 2 || fv.Z < 2) return hipSuccess;  // no cells
  const int eb = fv.enc & ENC_MASK_NUM;
  if (!tilingSupported(tl, eb)) return hipErrorInvalidValue;
  if ((tl.V == 16 || tl.V == 8) && fv.enc == ENC_F32 && fv.nF == 1 && fv.sx == 1 && fv.X >= 4)
    return launchFused16(fv, (float)m.t[0], m.m, bpfd, tl, s, info);
  if (info) {
    const Tiling gt = groupTiling(fv, tl);
    const TileGrid g = makeGrid(fv.X, fv.Y, fv.Z, gt);
    fillInfo(info, g, (long long)g.chunk * 8, gt);
    return hipSuccess;
  }
  // f32 fields compare against the threshold rounded to float
  // (src/binMapCUDA.cuh:138-139).
  if (fv.enc == ENC_F32) return launchFusedF32(fv, (float)m.t[0], m.m, bpfd, tl, s);
  if (fv.enc == ENC_F64) return launchFusedF64(fv, m.t[0], m.m, bpfd, tl, s);
  return hipErrorInvalidValue;
}
#endif

#if MKF_PART(5)
// Fused histogram + phase moments: 4 row pairs per wave (the tiling's slab
// depth, wave count and voxels per lane are kept).
template <typename T, int M, int V>
static hipError_t launchFusedMomT(const FieldView& fv, T thr, unsigned long long* bpfd, double* part,
                                  const TileGrid& g, hipStream_t s) {
  const dim3 grid((unsigned)g.chunk * 8u), block(g.waves * kWave);
  if (fv.nF > 1 || fv.sx != 1) {
    if constexpr (V == 1)
      hipLaunchKernelGGL((kFusedBPFD<T, M, 1, 4, true, true>), grid, block, 0, s, fieldArgs<T>(fv), thr, fv.X, fv.Y,
                         fv.Z, g, bpfd, part);
    else
      return hipErrorInvalidValue;
  } else {
    hipLaunchKernelGGL((kFusedBPFD<T, M, V, 4, false, true>), grid, block, 0, s, fieldArgs<T>(fv), thr, fv.X, fv.Y,
                       fv.Z, g, bpfd, part);
  }
  return hipGetLastError();
}

template <typename T, int V>
static hipError_t dispatchFusedMomM(const FieldView& fv, T thr, uint32_t m, unsigned long long* bpfd, double* part,
                                    const TileGrid& g, hipStream_t s) {
  switch (m & 7u) {
    case 0: return launchFusedMomT<T, 0, V>(fv, thr, bpfd, part, g, s);
    case 1: return launchFusedMomT<T, 1, V>(fv, thr, bpfd, part, g, s);
    case 2: return launchFusedMomT<T, 2, V>(fv, thr, bpfd, part, g, s);
    case 3: return launchFusedMomT<T, 3, V>(fv, thr, bpfd, part, g, s);
    case 4: return launchFusedMomT<T, 4, V>(fv, thr, bpfd, part, g, s);
    case 5: return launchFusedMomT<T, 5, V>(fv, thr, bpfd, part, g, s);
    case 6: return launchFusedMomT<T, 6, V>(fv, thr, bpfd, part, g, s);
    default: return launchFusedMomT<T, 7, V>(fv, thr, bpfd, part, g, s);
  }
}

// The persistent 8- and 16-voxel kernel takes one float field of unit x
// stride and at least 4 voxels per row.
static bool momentsBy16(const FieldView& fv, const Tiling& tl) {
  return (tl.V == 8 || tl.V == 16) && fv.enc == ENC_F32 && fv.nF == 1 && fv.sx == 1 && fv.X >= 4;
}

static Tiling fusedMomTiling(const FieldView& fv, const Tiling& tl0) {
  Tiling tl = tl0;
  tl.Ty = 4;
  if (momentsBy16(fv, tl)) return tl;
  if (tl.V >= 8) tl.V = 4;
  if (fv.nF > 1 || fv.sx != 1 || fv.X < tl.V) tl.V = 1;
  return tl;
}

long long fusedMomentParts(const FieldView& fv, const Tiling& tl0, uint32_t m) {
  if (fv.X < 2 || fv.Y < 2 || fv.Z < 2) return 0;
  const Tiling tl = fusedMomTiling(fv, tl0);
  if (momentsBy16(fv, tl)) {
    // The workgroup count follows the resident slots of the relation's own
    // kernel.
    long long n = 0;
    return launchFused16Moments(fv, 0.f, m, nullptr, nullptr, tl, nullptr, &n) == hipSuccess ? n : 0;
  }
  const TileGrid g = makeGrid(fv.X, fv.Y, fv.Z, tl);
  return g.nTiles > 0 ? (long long)g.chunk * 8 : 0;
}

hipError_t launchFusedMoments(const FieldView& fv, const BinMapF64& m, unsigned long long* bpfd, double* part,
                              double* mom, const Tiling& tl0, hipStream_t s) {
  const int eb = fv.enc & ENC_MASK_NUM;
  if (!tilingSupported(tl0, eb)) return hipErrorInvalidValue;
  const long long nParts = fusedMomentParts(fv, tl0, m.m);
  if (nParts == 0) return hipErrorInvalidValue;  // no cells: the caller takes the map path
  const Tiling tl = fusedMomTiling(fv, tl0);
  const TileGrid g = makeGrid(fv.X, fv.Y, fv.Z, tl);
  hipError_t e;
  if (momentsBy16(fv, tl))
    e = launchFused16Moments(fv, (float)m.t[0], m.m, bpfd, part, tl, s, nullptr);
  else if (fv.enc == ENC_F32)
    e = tl.V == 1 ? dispatchFusedMomM<float, 1>(fv, (float)m.t[0], m.m, bpfd, part, g, s)
                  : dispatchFusedMomM<float, 4>(fv, (float)m.t[0], m.m, bpfd, part, g, s);
  else if (fv.enc == ENC_F64)
    e = tl.V == 1 ? dispatchFusedMomM<double, 1>(fv, m.t[0], m.m, bpfd, part, g, s)
                  : dispatchFusedMomM<double, 2>(fv, m.t[0], m.m, bpfd, part, g, s);
  else
    return hipErrorInvalidValue;
  if (e != hipSuccess) return e;
  return launchMomentsReduce(part, nParts, mom, s);
}
#endif

#if MKF_PART(8)
// The moments form runs up to 16 waves per CU (4 per SIMD), as many as its
// registers allow: the 16-voxel kernel is held to 168 VGPRs (3 per SIMD;
// unbounded it took 181 and fitted 2), the 8-voxel one takes 113 (4).
// Measured on MI355X (f32 GRF): 8 voxels per lane, 1536 x 1000 x 700 with 32
// plane pairs 0.807 ms at 16 per CU against 0.885 at 8; 16 voxels per lane,
// 1024^3 with 64 plane pairs 0.717 ms at 3 per SIMD against 0.811 at 2
// (2048 x 1024 x 512: 0.726 against 0.815). At 4 per SIMD (128 VGPRs) the
// 16-voxel kernel spills 75 VGPRs, and the 8-voxel one at 5 spills 18.
constexpr int kFused16MomPerCU = 16;

unsigned fused16MomSlots(int V) {
  return persistentSlots(V == 8 ? (const void*)kFused16<6, 4, 2, true> : (const void*)kFused16<6, 4, 4, true>,
                         kFused16MomPerCU);
}

template <int M, int NG>
static hipError_t launchFused16MomK(const FieldView& fv, float thr, unsigned long long* bpfd, double* part,
                                    const Tiling& tl, hipStream_t s, long long* nParts) {
  const TileGrid g = grid16(fv.X, fv.Y, fv.Z, tl);
  const unsigned nBlocks = blocks16(g, tl, (const void*)kFused16<M, 4, NG, true>, kFused16MomPerCU);
  if (nParts) {
    *nParts = nBlocks;
    return hipSuccess;
  }
  hipLaunchKernelGGL((kFused16<M, 4, NG, true>), dim3(nBlocks), dim3(kWave), 0, s, (const float*)fv.p[0], thr, fv.X,
                     fv.Y, fv.Z, fv.sy, fv.sz, g, bpfd, part);
  return hipGetLastError();
}

template <int M>
static hipError_t launchFused16MomM(const FieldView& fv, float thr, unsigned long long* bpfd, double* part,
                                    const Tiling& tl, hipStream_t s, long long* nParts) {
  return tl.V == 8 ? launchFused16MomK<M, 2>(fv, thr, bpfd, part, tl, s, nParts)
                   : launchFused16MomK<M, 4>(fv, thr, bpfd, part, tl, s, nParts);
}

hipError_t launchFused16Moments(const FieldView& fv, float thr, uint32_t m, unsigned long long* bpfd, double* part,
                                const Tiling& tl, hipStream_t s, long long* nParts) {
  if (fv.X < 4 || fv.Y < 2 || fv.Z < 2 || tl.Ty != 4 || (tl.V != 8 && tl.V != 16)) return hipErrorInvalidValue;
  switch (m & 7u) {
    case 0: return launchFused16MomM<0>(fv, thr, bpfd, part, tl, s, nParts);
    case 1: return launchFused16MomM<1>(fv, thr, bpfd, part, tl, s, nParts);
    case 2: return launchFused16MomM<2>(fv, thr, bpfd, part, tl, s, nParts);
    case 3: return launchFused16MomM<3>(fv, thr, bpfd, part, tl, s, nParts);
    case 4: return launchFused16MomM<4>(fv, thr, bpfd, part, tl, s, nParts);
    case 5: return launchFused16MomM<5>(fv, thr, bpfd, part, tl, s, nParts);
    case 6: return launchFused16MomM<6>(fv, thr, bpfd, part, tl, s, nParts);
    default: return launchFused16MomM<7>(fv, thr, bpfd, part, tl, s, nParts);
  }
}
#endif

#if MKF_PART(6)
// The reference's GPU row walk for rows narrower than one PBM word
// (src/bpfdCUDA.cuh:92-118 with n = X < 32), for callers that ask for its
// counts: four 64-bit registers hold the row pair's words in planes z and
// z+1; the word gives X - 1 patterns (bits 0..1 of every register, then all
// shift right by one), then the same word is ORed in again shifted up by one
// and gives X more. One lane per (row pair, plane pair); LDS histogram per
// workgroup, flushed with one 64-bit atomic per nonzero bin.
__global__ __launch_bounds__(256) void kBPFDNarrowRef(const uint32_t* __restrict__ pW, int X, int Y, int Z,
                                                      uint32_t rowWS, uint32_t planeWS,
                                                      unsigned long long* __restrict__ bpfd) {
  __shared__ uint32_t hist[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const long long pairs = (long long)(Y - 1) * (Z - 1);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < pairs;
       i += (long long)gridDim.x * blockDim.x) {
    const long long y = i % (Y - 1), z = i / (Y - 1);
    const uint32_t* p0 = pW + z * planeWS + y * rowWS;
    const uint32_t w[4] = {p0[0], p0[rowWS], p0[planeWS], p0[planeWS + rowWS]};
    uint64_t u[4] = {w[0], w[1], w[2], w[3]};
    for (int pass = 0; pass < 2; ++pass) {
      for (int j = 0; j < X - 1 + pass; ++j) {
        const uint32_t code =
            (uint32_t)((u[0] & 3u) | ((u[1] & 3u) << 2) | ((u[2] & 3u) << 4) | ((u[3] & 3u) << 6));
        atomicAdd(hist + code, 1u);
#pragma unroll
        for (int q = 0; q < 4; ++q) u[q] >>= 1;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) u[q] |= (uint64_t)w[q] << 1;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += blockDim.x)
    if (hist[i]) atomicAdd(bpfd + i, (unsigned long long)hist[i]);
}

hipError_t launchBPFDNarrowRef(const uint32_t* pW, int X, int Y, int Z, uint32_t rowWS, uint32_t planeWS,
                               unsigned long long* bpfd, hipStream_t s) {
  if (X < 2 || X >= 32 || Y < 2 || Z < 2) return X >= 32 ? hipErrorInvalidValue : hipSuccess;
  const long long pairs = (long long)(Y - 1) * (Z - 1);
  const unsigned blocks = (unsigned)std::min<long long>(4096, (pairs + 255) / 256);
  hipLaunchKernelGGL(kBPFDNarrowRef, dim3(blocks), dim3(256), 0, s, pW, X, Y, Z, rowWS, planeWS, bpfd);
  return hipGetLastError();
}

hipError_t launchBPFD(const uint32_t* pW, int X, int Y, int Z, uint32_t rowWS, uint32_t planeWS,
                      unsigned long long* bpfd, const Tiling& tl0, hipStream_t s, LaunchInfo* info) {
  if (info) *info = LaunchInfo{0, 0, 0, tl0};
  if (X < 2 || Y < 2 || Z < 2) return hipSuccess;
  // tl.V is 16 (half words), 4 (nibbles of the PBM words) or 1.
  const Tiling& tl = tl0;
  if (!tilingSupported(tl, 4)) return hipErrorInvalidValue;
  if (tl.V == 16) {
    const TileGrid g = grid16(X, Y, Z, tl);
    if (g.nTiles <= 0 && !info) return hipSuccess;
    const unsigned nBlocks = blocks16(g, tl, tl.Ty == 2 ? (const void*)kBPFD16<2> : (const void*)kBPFD16<4>, 0);
    if (info) {
      fillInfo(info, g, nBlocks, tl);
      return hipSuccess;
    }
    if (tl.Ty == 2)
      hipLaunchKernelGGL((kBPFD16<2>), dim3(nBlocks), dim3(kWave), 0, s, pW, X, Y, Z, rowWS, planeWS, g, bpfd);
    else
      hipLaunchKernelGGL((kBPFD16<4>), dim3(nBlocks), dim3(kWave), 0, s, pW, X, Y, Z, rowWS, planeWS, g, bpfd);
    return hipGetLastError();
  }
  if (const int GW = narrowGroupWords(rowWS, Y, tl)) {
    const TileGrid g = narrowGrid(GW, Y, Z, tl);
    if (info) {
      fillInfo(info, g, (long long)g.chunk * 8, tl);
      return hipSuccess;
    }
    const dim3 grid((unsigned)g.chunk * 8u), block(g.waves * kWave);
#define MKF_BPFD4R(TY, GWV) \
  hipLaunchKernelGGL((kBPFD4R<TY, GWV>), grid, block, 0, s, pW, X, Y, Z, rowWS, planeWS, g, bpfd)
    switch (tl.Ty * 8 + GW) {
      case 2 * 8 + 1: MKF_BPFD4R(2, 1); break;
      case 2 * 8 + 2: MKF_BPFD4R(2, 2); break;
      case 2 * 8 + 4: MKF_BPFD4R(2, 4); break;
      case 8 * 8 + 1: MKF_BPFD4R(8, 1); break;
      case 8 * 8 + 2: MKF_BPFD4R(8, 2); break;
      case 8 * 8 + 4: MKF_BPFD4R(8, 4); break;
      case 2 * 8 + 8: MKF_BPFD4R(2, 8); break;
      case 8 * 8 + 8: MKF_BPFD4R(8, 8); break;
      case 4 * 8 + 1: MKF_BPFD4R(4, 1); break;
      case 4 * 8 + 2: MKF_BPFD4R(4, 2); break;
      case 4 * 8 + 8: MKF_BPFD4R(4, 8); break;
      default: MKF_BPFD4R(4, 4); break;
    }
#undef MKF_BPFD4R
    return hipGetLastError();
  }
  const TileGrid g = makeGrid(X, Y, Z, tl);
  if (info) {
    fillInfo(info, g, (long long)g.chunk * 8, tl);
    return hipSuccess;
  }
  const dim3 grid((unsigned)g.chunk * 8u), block(g.waves * kWave);
  if (tl.V == 4) {
    switch (tl.Ty) {
      case 2: hipLaunchKernelGGL((kBPFD4<2>), grid, block, 0, s, pW, X, Y, Z, rowWS, planeWS, g, bpfd); break;
      case 8: hipLaunchKernelGGL((kBPFD4<8>), grid, block, 0, s, pW, X, Y, Z, rowWS, planeWS, g, bpfd); break;
      default: hipLaunchKernelGGL((kBPFD4<4>), grid, block, 0, s, pW, X, Y, Z, rowWS, planeWS, g, bpfd); break;
    }
    return hipGetLastError();
  }
  switch (tl.Ty) {
    case 2: hipLaunchKernelGGL((kBPFD<2>), grid, block, 0, s, pW, X, Y, Z, rowWS, planeWS, g, bpfd); break;
    case 8: hipLaunchKernelGGL((kBPFD<8>), grid, block, 0, s, pW, X, Y, Z, rowWS, planeWS, g, bpfd); break;
    default: hipLaunchKernelGGL((kBPFD<4>), grid, block, 0, s, pW, X, Y, Z, rowWS, planeWS, g, bpfd); break;
  }
  return hipGetLastError();
}
#endif

}  // namespace mkf
