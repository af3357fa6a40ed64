// mkf_kernels.hip - gfx950 kernels of the hot path and their launchers.
//
//   kPBM         field -> packed binary map      (replaces mapField /
//                mapStrideField, src/binMapCUDA.cuh:282-323)
//   kBPFD        PBM -> 256-bin pattern histogram (replaces
//                addMultiPlaneSeqBPFD / addPlaneBPFD, src/bpfdCUDA.cuh:158-206)
//   kFusedBPFD   field -> histogram, reading the field once (no reference
//                counterpart; the reference round-trips the PBM through memory)
#include "mkf_device.hpp"
#include "mkf_internal.h"

namespace mkf {

// ---------------------------------------------------------------------------
// Packed binary map: one wave per x-segment and RU consecutive rows.
// ---------------------------------------------------------------------------
constexpr int kPbmWaves = 4;
constexpr int kPbmRows = 8;

template <typename T, int M>
__global__ __launch_bounds__(kPbmWaves* kWave) void kPBM(const T* __restrict__ f, T thr, int X, int Y,
                                                           long long sy, long long sz, int nSegX, long long nRows,
                                                           uint32_t* __restrict__ pW, uint32_t rowWS,
                                                           uint32_t planeWS) {
  const int lane = threadIdx.x & (kWave - 1);
  const long long gw = (long long)blockIdx.x * kPbmWaves + (threadIdx.x >> 6);
  const int seg = (int)(gw % nSegX);
  const long long row0 = (gw / nSegX) * kPbmRows;
  if (row0 >= nRows) return;
  const int x0 = seg * kWave;
  const int xc = min(x0 + lane, X - 1);
  const uint64_t tail = lowBits(X - x0);
  const int nw = min(2, (int)rowWS - 2 * seg);
  T v[kPbmRows];
#pragma unroll
  for (int u = 0; u < kPbmRows; ++u) {
    const long long row = row0 + u;
    if (row < nRows) {
      const long long y = row % Y, z = row / Y;
      v[u] = f[y * sy + z * sz + xc];
    }
  }
#pragma unroll
  for (int u = 0; u < kPbmRows; ++u) {
    const long long row = row0 + u;
    if (row < nRows) {
      const long long y = row % Y, z = row / Y;
      const uint64_t R = relMask<M>(v[u], thr) & tail;
      const size_t wi = (size_t)y * rowWS + (size_t)z * planeWS + 2 * seg;
      if (lane < nw) pW[wi + lane] = lane ? (uint32_t)(R >> 32) : (uint32_t)R;
    }
  }
}

// ---------------------------------------------------------------------------
// Histogram kernels. Tiles are (x-segment, y-range, z-group); a block of ZW
// waves takes ZW consecutive z-ranges of TP plane pairs so the plane shared by
// neighbouring waves is fetched from HBM once and hit in cache by the other.
// Block b runs tile (b & 7) * chunk + (b >> 3): consecutive tiles land on the
// same XCD (blocks are dealt to the 8 XCDs round-robin), so the halo row a
// tile shares with the next y-range stays in that XCD's L2.
// ---------------------------------------------------------------------------
struct TileGrid {
  int nSegX, nYR, nZG, nTiles, chunk, Ty;
};

template <int ZW>
__device__ __forceinline__ void flushHistogram(uint32_t (*hist)[256], const Tally& tally,
                                               unsigned long long* __restrict__ bpfd) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x >> 6;
  if (lane == 0) {
    hist[w][0] += tally.c0;
    hist[w][255] += tally.c255;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += ZW * kWave) {
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < ZW; ++k) s += hist[k][i];
    if (s) atomicAdd(bpfd + i, s);
  }
}

__device__ __forceinline__ bool tileOf(const TileGrid& g, int& seg, int& yr, int& zg) {
  const int b = blockIdx.x;
  const int L = (b & 7) * g.chunk + (b >> 3);
  if (L >= g.nTiles) return false;
  seg = L % g.nSegX;
  const int r = L / g.nSegX;
  yr = r % g.nYR;
  zg = r / g.nYR;
  return true;
}

template <typename T, int M, int TP, int ZW>
__global__ __launch_bounds__(ZW* kWave) void kFusedBPFD(const T* __restrict__ f, T thr, int X, int Y, int Z,
                                                          long long sy, long long sz, TileGrid g,
                                                          unsigned long long* __restrict__ bpfd) {
  __shared__ uint32_t hist[ZW][256];
  int seg, yr, zg;
  if (!tileOf(g, seg, yr, zg)) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x >> 6;
  for (int i = lane; i < 256; i += kWave) hist[w][i] = 0;

  Tally tally;
  const int z0 = (zg * ZW + w) * TP;
  const int nT = min(TP, Z - 1 - z0);
  const int y0 = yr * g.Ty;
  const int nY = min(g.Ty, Y - 1 - y0);
  const int x0 = seg * kWave;
  const int nCells = min(kWave, X - 1 - x0);
  if (nT > 0 && nY > 0 && nCells > 0) {
    FieldRows<T, M> src;
    src.f = f;
    src.thr = thr;
    src.sy = sy;
    src.sz = sz;
    src.xc = min(x0 + lane, X - 1);
    src.haloX = min(x0 + kWave, X - 1);
    src.halo = x0 + kWave < X;
    src.tail = lowBits(X - x0);
    walkSegment<TP>(src, y0, nY, z0, nT, lane < nCells, lowBits(nCells), hist[w], tally);
  }
  flushHistogram<ZW>(hist, tally, bpfd);
}

template <int TP, int ZW>
__global__ __launch_bounds__(ZW* kWave) void kBPFD(const uint32_t* __restrict__ pW, int X, int Y, int Z,
                                                     uint32_t rowWS, uint32_t planeWS, TileGrid g,
                                                     unsigned long long* __restrict__ bpfd) {
  __shared__ uint32_t hist[ZW][256];
  int seg, yr, zg;
  if (!tileOf(g, seg, yr, zg)) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x >> 6;
  for (int i = lane; i < 256; i += kWave) hist[w][i] = 0;

  Tally tally;
  const int z0 = (zg * ZW + w) * TP;
  const int nT = min(TP, Z - 1 - z0);
  const int y0 = yr * g.Ty;
  const int nY = min(g.Ty, Y - 1 - y0);
  const int x0 = seg * kWave;
  const int nCells = min(kWave, X - 1 - x0);
  if (nT > 0 && nY > 0 && nCells > 0) {
    PbmRows src;
    src.w = pW;
    src.rowWS = rowWS;
    src.planeWS = planeWS;
    src.wi0 = 2u * seg;
    src.o1 = x0 + 32 < X ? 1u : 0u;
    src.o2 = x0 + kWave < X ? 2u : 0u;
    src.halo = x0 + kWave < X;
    src.tail = lowBits(X - x0);
    walkSegment<TP>(src, y0, nY, z0, nT, lane < nCells, lowBits(nCells), hist[w], tally);
  }
  flushHistogram<ZW>(hist, tally, bpfd);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline long long ceilDiv(long long a, long long b) { return (a + b - 1) / b; }

static TileGrid makeGrid(int X, int Y, int Z, const Tiling& tl) {
  TileGrid g;
  g.nSegX = (int)ceilDiv(X - 1, kWave);
  g.Ty = tl.Ty;
  g.nYR = (int)ceilDiv(Y - 1, tl.Ty);
  const long long nZR = ceilDiv(Z - 1, tl.TP);
  g.nZG = (int)ceilDiv(nZR, tl.ZW);
  g.nTiles = g.nSegX * g.nYR * g.nZG;
  g.chunk = (int)ceilDiv(g.nTiles, 8);
  return g;
}

Tiling chooseTiling(int X, int Y, int Z, int elemBytes) {
  (void)elemBytes;
  Tiling tl;
  tl.TP = 8;
  tl.ZW = 8;
  tl.Ty = 64;
  // Shrink the y-range until there are enough tiles to fill 256 CUs.
  while (tl.Ty > 4) {
    const TileGrid g = makeGrid(X, Y, Z, tl);
    if (g.nTiles >= 2048) break;
    tl.Ty >>= 1;
  }
  return tl;
}

template <typename T, int M>
static hipError_t launchPBMT(const FieldView& fv, T thr, uint32_t* pW, uint32_t rowWS, uint32_t planeWS,
                             hipStream_t s) {
  const int nSegX = (int)ceilDiv(fv.X, kWave);
  const long long nRows = (long long)fv.Y * fv.Z;
  const long long nWaves = nSegX * ceilDiv(nRows, kPbmRows);
  const long long nBlocks = ceilDiv(nWaves, kPbmWaves);
  hipLaunchKernelGGL((kPBM<T, M>), dim3((unsigned)nBlocks), dim3(kPbmWaves * kWave), 0, s,
                     (const T*)fv.p, thr, fv.X, fv.Y, fv.sy, fv.sz, nSegX, nRows, pW, rowWS, planeWS);
  return hipGetLastError();
}

template <typename T>
static hipError_t dispatchPBM(const FieldView& fv, T thr, uint32_t m, uint32_t* pW, uint32_t rowWS,
                              uint32_t planeWS, hipStream_t s) {
  switch (m & 7u) {
    case 0: return launchPBMT<T, 0>(fv, thr, pW, rowWS, planeWS, s);
    case 1: return launchPBMT<T, 1>(fv, thr, pW, rowWS, planeWS, s);
    case 2: return launchPBMT<T, 2>(fv, thr, pW, rowWS, planeWS, s);
    case 3: return launchPBMT<T, 3>(fv, thr, pW, rowWS, planeWS, s);
    case 4: return launchPBMT<T, 4>(fv, thr, pW, rowWS, planeWS, s);
    case 5: return launchPBMT<T, 5>(fv, thr, pW, rowWS, planeWS, s);
    case 6: return launchPBMT<T, 6>(fv, thr, pW, rowWS, planeWS, s);
    default: return launchPBMT<T, 7>(fv, thr, pW, rowWS, planeWS, s);
  }
}

hipError_t launchPBM(const FieldView& fv, const BinMapF64& m, uint32_t* pW, uint32_t rowWS, uint32_t planeWS,
                     hipStream_t s) {
  if (fv.X < 1 || fv.Y < 1 || fv.Z < 1) return hipErrorInvalidValue;
  // f32 fields compare against the threshold rounded to float
  // (src/binMapCUDA.cuh:138-139).
  if (fv.enc == ENC_F32) return dispatchPBM<float>(fv, (float)m.t[0], m.m, pW, rowWS, planeWS, s);
  if (fv.enc == ENC_F64) return dispatchPBM<double>(fv, m.t[0], m.m, pW, rowWS, planeWS, s);
  return hipErrorInvalidValue;
}

template <typename T, int M, int TP, int ZW>
static hipError_t launchFusedT(const FieldView& fv, T thr, unsigned long long* bpfd, const Tiling& tl,
                               hipStream_t s) {
  const TileGrid g = makeGrid(fv.X, fv.Y, fv.Z, tl);
  if (g.nTiles <= 0) return hipSuccess;
  hipLaunchKernelGGL((kFusedBPFD<T, M, TP, ZW>), dim3((unsigned)g.chunk * 8u), dim3(ZW * kWave), 0, s,
                     (const T*)fv.p, thr, fv.X, fv.Y, fv.Z, fv.sy, fv.sz, g, bpfd);
  return hipGetLastError();
}

template <typename T>
static hipError_t dispatchFused(const FieldView& fv, T thr, uint32_t m, unsigned long long* bpfd, const Tiling& tl,
                                hipStream_t s) {
  switch (m & 7u) {
    case 0: return launchFusedT<T, 0, 8, 8>(fv, thr, bpfd, tl, s);
    case 1: return launchFusedT<T, 1, 8, 8>(fv, thr, bpfd, tl, s);
    case 2: return launchFusedT<T, 2, 8, 8>(fv, thr, bpfd, tl, s);
    case 3: return launchFusedT<T, 3, 8, 8>(fv, thr, bpfd, tl, s);
    case 4: return launchFusedT<T, 4, 8, 8>(fv, thr, bpfd, tl, s);
    case 5: return launchFusedT<T, 5, 8, 8>(fv, thr, bpfd, tl, s);
    case 6: return launchFusedT<T, 6, 8, 8>(fv, thr, bpfd, tl, s);
    default: return launchFusedT<T, 7, 8, 8>(fv, thr, bpfd, tl, s);
  }
}

hipError_t launchFused(const FieldView& fv, const BinMapF64& m, unsigned long long* bpfd, const Tiling& tl,
                       hipStream_t s) {
  if (fv.X < 2 || fv.Y < 2 || fv.Z < 2) return hipSuccess;  // no cells
  if (tl.TP != 8 || tl.ZW != 8 || tl.Ty < 1) return hipErrorInvalidValue;
  if (fv.enc == ENC_F32) return dispatchFused<float>(fv, (float)m.t[0], m.m, bpfd, tl, s);
  if (fv.enc == ENC_F64) return dispatchFused<double>(fv, m.t[0], m.m, bpfd, tl, s);
  return hipErrorInvalidValue;
}

hipError_t launchBPFD(const uint32_t* pW, int X, int Y, int Z, uint32_t rowWS, uint32_t planeWS,
                      unsigned long long* bpfd, const Tiling& tl, hipStream_t s) {
  if (X < 2 || Y < 2 || Z < 2) return hipSuccess;
  if (tl.TP != 8 || tl.ZW != 8 || tl.Ty < 1) return hipErrorInvalidValue;
  const TileGrid g = makeGrid(X, Y, Z, tl);
  hipLaunchKernelGGL((kBPFD<8, 8>), dim3((unsigned)g.chunk * 8u), dim3(8 * kWave), 0, s, pW, X, Y, Z, rowWS,
                     planeWS, g, bpfd);
  return hipGetLastError();
}

}  // namespace mkf
