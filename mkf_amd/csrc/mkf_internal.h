// mkf_internal.h - launchers shared by the C entry points (host side only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mkf.h"

namespace mkf {

// Up to 4 device fields of identical dims and element strides, thresholded
// as their sum in table order (src/binMapCUDA.cuh:178-212: s = f0, then
// s += f1, f2, f3 in the element type).
constexpr int kMaxFields = 4;
struct FieldView {
  const void* p[kMaxFields];
  int nF;
  NumEnc enc;  // ENC_F32 or ENC_F64
  int X, Y, Z;
  long long sx, sy, sz;
};

// One planar field (x contiguous, y and z strides X and X*Y).
inline FieldView planarField(const void* p, NumEnc enc, int X, int Y, int Z) {
  FieldView fv{};
  fv.p[0] = p;
  fv.nF = 1;
  fv.enc = enc;
  fv.X = X;
  fv.Y = Y;
  fv.Z = Z;
  fv.sx = 1;
  fv.sy = X;
  fv.sz = (long long)X * Y;
  return fv;
}

// Work decomposition of the histogram kernels: a workgroup of up to ZW waves
// side by side along x (one 64-voxel segment each) holds Ty row pairs and
// marches over TP plane pairs.
struct Tiling {
  int TP;  // plane pairs per tile (>= 1)
  int ZW;  // waves per workgroup along x (1..16)
  int Ty;  // row pairs per wave (2, 4 or 8)
  int V;   // voxels per lane: 1, or 16 bytes' worth (4 floats / 2 doubles)
};

// Geometry of a histogram launch: tiles, workgroups launched (for the
// persistent kernels, each walks tiles b, b + workgroups, ...), waves per
// workgroup and the tiling the kernel runs with.
struct LaunchInfo {
  long long tiles, workgroups;
  int wavesPerGroup;
  Tiling used;
};

// Records the text of an error for mkfLastError() (this thread) and returns
// code.
int setError(int code, const char* fmt, ...);

// True when the named fault point, armed by mkfDebugInjectFault, is due to
// fail on this call (counts the calls of an armed point; thread-safe).
bool faultDue(const char* point);

Tiling chooseTiling(int X, int Y, int Z, int elemBytes);  // fused kernel
// Fused kernel, 8 or 16 voxels per lane (floats): the slab depth that spreads
// the tiles best over quarters / 4 of the kernel's resident slots.
Tiling choose16Tiling(int X, int Y, int Z, int V = 16, int quarters = 4);
Tiling choose4Tiling(int X, int Y, int Z);  // float fields at 4 voxels per lane
Tiling chooseBpfdTiling(int X, int Y, int Z);              // histogram of a PBM
Tiling chooseMomTiling(int X, int Y, int Z, int elemBytes);  // fused kernel with moments
bool tilingSupported(const Tiling& tl, int elemBytes);

// Thresholds fv into the PBM pW (rowWS / planeWS word strides).
hipError_t launchPBM(const FieldView& fv, const BinMapF64& m, uint32_t* pW, uint32_t rowWS, uint32_t planeWS,
                     hipStream_t s);

// As launchPBM (pW may be nullptr: no map), and ADDS the phase moments
// {sum0, sumsq0, sum1, sumsq1, n0, n1} of the thresholded values to mom
// (device, 6 doubles). part is device scratch of
// pbmMomentParts(fv, pW, rowWS, planeWS) x 6 doubles.
long long pbmMomentParts(const FieldView& fv, const uint32_t* pW, uint32_t rowWS, uint32_t planeWS);

// Adds nRows per-workgroup histograms (256 x u32 each, device) to bpfd.
hipError_t launchHistRowsReduce(const uint32_t* rows, long long nRows, unsigned long long* bpfd, hipStream_t s);
hipError_t launchPBMMoments(const FieldView& fv, const BinMapF64& m, uint32_t* pW, uint32_t rowWS,
                            uint32_t planeWS, double* part, double* mom, hipStream_t s);

// Adds the BPFD of fv thresholded by m to bpfd (device, 256 x u64), reading
// the field once; no PBM is written. With info, nothing is launched and
// info receives the launch's geometry.
hipError_t launchFused(const FieldView& fv, const BinMapF64& m, unsigned long long* bpfd, const Tiling& tl,
                       hipStream_t s, LaunchInfo* info = nullptr);

// As launchFused, and ADDS the phase moments of fv (as launchPBMMoments) to
// mom, reading the field once. Needs X, Y, Z >= 2 (hipErrorInvalidValue
// otherwise); the tiling's row pairs are set to 4. part is device scratch of
// fusedMomentParts(fv, tl, m.m) x 6 doubles (m: the relation mask).
long long fusedMomentParts(const FieldView& fv, const Tiling& tl, uint32_t m);
hipError_t launchFusedMoments(const FieldView& fv, const BinMapF64& m, unsigned long long* bpfd, double* part,
                              double* mom, const Tiling& tl, hipStream_t s);

// Adds the BPFD of the PBM pW (dims X, Y, Z) to bpfd (with info: only the
// geometry, as launchFused).
hipError_t launchBPFD(const uint32_t* pW, int X, int Y, int Z, uint32_t rowWS, uint32_t planeWS,
                      unsigned long long* bpfd, const Tiling& tl, hipStream_t s, LaunchInfo* info = nullptr);

// Adds the counts of the reference's GPU row walk for rows of X < 32 (2X - 1
// patterns per row pair) of the PBM pW to bpfd.
hipError_t launchBPFDNarrowRef(const uint32_t* pW, int X, int Y, int Z, uint32_t rowWS, uint32_t planeWS,
                               unsigned long long* bpfd, hipStream_t s);

// How the drop-in histogram entry points treat rows of X < 32
// (mkfSetNarrowRowMode).
int narrowRowMode();

// Copies bytes (a multiple of 16) with a kernel on s; either side may be
// the device address of mapped pinned host memory.
hipError_t copyWords(void* dst, const void* src, size_t bytes, hipStream_t s);

// Reusable launcher of the batched histogram kernel for a set of fields that
// changes from launch to launch. Descriptions are validated as by
// mkfBatchCreate; each run stages them through pinned memory (waiting for the
// previous run's staging copy) and launches asynchronously on s.
struct BatchLauncher;
BatchLauncher* batchLauncherCreate();
int batchLauncherRun(BatchLauncher* L, const MkfBatchField* fields, int n, const MkfBatchShape* shape, int flags,
                     hipStream_t s);
void batchLauncherDestroy(BatchLauncher* L);

}  // namespace mkf
