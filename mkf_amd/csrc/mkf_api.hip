// mkf_api.hip - C entry points of libmkf_amd.so.
//
// The reference's blocking entry points (binMapCUDA, mkfCUDAGetBPFD[H],
// mkfCUDAGetBPFDautoCtx, mkfAcc*) keep their names, argument meaning, return
// conventions and library-owned buffers; the additive stream-aware forms
// (mkf_ext.h) return error codes. All GPU work is done by the kernels in
// mkf_kernels.hip; nothing here computes a result on the CPU.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "mkf_internal.h"

using namespace mkf;

namespace {

thread_local char gErr[256] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(gErr, sizeof(gErr), fmt, ap);
  va_end(ap);
  if (getenv("MKF_VERBOSE")) fprintf(stderr, "[mkf] %s\n", gErr);
  return code;
}

int hipFail(hipError_t e, const char* what) {
  return fail(MKF_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace

int mkf::setError(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(gErr, sizeof(gErr), fmt, ap);
  va_end(ap);
  if (getenv("MKF_VERBOSE")) fprintf(stderr, "[mkf] %s\n", gErr);
  return code;
}

namespace {

// Armed fault points: name -> calls left before the one that fails.
std::mutex gFaultLock;
std::map<std::string, int> gFaults;

}  // namespace

namespace {
std::atomic<int> gNarrowMode{MKF_NARROW_CELLS};
}  // namespace

int mkf::narrowRowMode() { return gNarrowMode.load(std::memory_order_relaxed); }

bool mkf::faultDue(const char* point) {
  std::lock_guard<std::mutex> g(gFaultLock);
  if (gFaults.empty()) return false;
  auto it = gFaults.find(point);
  if (it == gFaults.end()) return false;
  if (--it->second > 0) return false;
  gFaults.erase(it);
  return true;
}

namespace {

// Library-owned buffers of the blocking entry points (the reference keeps the
// same statics: gpDevW, src/binMapCUDA.cu:142; gHD, src/mkfCUDA.cu:69). The
// PBM buffer is regrown when a larger field arrives (the reference never
// resizes it).
std::mutex gLock;
BMPackWord* gDevW = nullptr;
size_t gDevWWords = 0;
size_t* gHostBPFD = nullptr;
unsigned long long* gDevBPFD = nullptr;

// Event timer on one stream: the reference's CTimerCUDA (src/utilCUDA.hpp:21-60).
struct EventTimer {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t s;
  explicit EventTimer(hipStream_t st) : s(st) {
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) hipEventRecord(e0, s);
  }
  ~EventTimer() {
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
  }
  float stopMs() {
    float ms = 0.f;
    if (!e0 || !e1) return ms;
    hipEventRecord(e1, s);
    hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    return ms;
  }
};

// Resolves pI into a field view: the first (up to) 4 valid entries of the
// table (summed by the kernels), the dims > 1 compacted to the front, and the
// element strides (pS, any positive x, y, z strides; planar by default)
// (src/gridUtil.cu:30-56 validate, src/binMapCUDA.cuh:149-212 for the
// pointer choice and the multi-field sum).
int resolveField(const BMFieldInfo* pI, FieldView* fv, FieldDef def[3]) {
  if (!pI) return fail(MKF_ERR_ARG, "NULL field info");
  int nField = 0;
  if (pI->pFieldDevPtrTable) nField = countValidPtrByMask(pI->pFieldDevPtrTable, pI->fieldTableMask);
  const int nDim = copyNonZeroDef(def, pI->pD, 3);
  if (nDim <= 0 || nField <= 0) return fail(MKF_ERR_ARG, "no valid field (nDim=%d nField=%d)", nDim, nField);
  if (pI->elemID != ENC_F32 && pI->elemID != ENC_F64)
    return fail(MKF_ERR_ARG, "element encoding 0x%02X is not F32/F64", pI->elemID);
  // The first (up to) 4 valid entries of the table, summed in table order
  // (src/binMapCUDA.cuh:178-212; BMCU_FIELD_MAX, src/binMapUtil.h:105).
  ConstFieldPtr fp[kMaxFields];
  *fv = FieldView{};
  fv->nF = copyValidPtrByMask(fp, kMaxFields, pI->pFieldDevPtrTable, pI->fieldTableMask);
  for (int i = 0; i < fv->nF; ++i) fv->p[i] = fp[i].p;
  fv->enc = pI->elemID;
  fv->X = def[0];
  fv->Y = def[1];
  fv->Z = def[2];
  if (pI->pS) {
    // Strides follow pD's order; they are only meaningful here when no dim
    // was dropped by the compaction above.
    if (!pI->pD || pI->pD[0] != def[0] || pI->pD[1] != def[1] || pI->pD[2] != def[2])
      return fail(MKF_ERR_UNSUPPORTED, "strided field with dims <= 1");
    if (pI->pS[0] < 1 || pI->pS[1] < 1 || pI->pS[2] < 1)
      return fail(MKF_ERR_ARG, "strides must be positive (%ld, %ld, %ld)", (long)pI->pS[0], (long)pI->pS[1],
                  (long)pI->pS[2]);
    fv->sx = pI->pS[0];
    fv->sy = pI->pS[1];
    fv->sz = pI->pS[2];
  } else {
    fv->sx = 1;
    fv->sy = def[0];
    fv->sz = (long long)def[0] * def[1];
  }
  return MKF_OK;
}

// Fused-kernel tilings chosen by mkfFusedAutotune, per field shape.
struct ShapeKey {
  int X, Y, Z, enc, nF;
  long long sx;
  bool operator<(const ShapeKey& o) const {
    return std::tie(X, Y, Z, enc, nF, sx) < std::tie(o.X, o.Y, o.Z, o.enc, o.nF, o.sx);
  }
};
std::mutex gTuneLock;
std::map<ShapeKey, Tiling> gTuned;

ShapeKey keyOf(const FieldView& fv) { return ShapeKey{fv.X, fv.Y, fv.Z, (int)fv.enc, fv.nF, fv.sx}; }

// The fused kernel's tiling: the caller's, else a tuned one, else the
// built-in choice.
Tiling fusedTiling(const MkfTiling* pT, const FieldView& fv) {
  const int eb = fv.enc & ENC_MASK_NUM;
  if (pT) {
    Tiling tl{pT->planePairs, pT->waves, pT->rowPairs, pT->voxelsPerLane};
    if (tl.V == 0) {
      tl.V = chooseTiling(fv.X, fv.Y, fv.Z, eb).V;
      if (tl.V >= 8 && tl.Ty == 8) tl.V = 4;  // the 8- and 16-voxel kernel stacks 2 or 4 row pairs
    }
    return tl;
  }
  {
    std::lock_guard<std::mutex> g(gTuneLock);
    auto it = gTuned.find(keyOf(fv));
    if (it != gTuned.end()) return it->second;
  }
  return chooseTiling(fv.X, fv.Y, fv.Z, eb);
}

Tiling toTiling(const MkfTiling* pT, int X, int Y, int Z, int elemBytes) {
  if (pT) {
    Tiling tl{pT->planePairs, pT->waves, pT->rowPairs, pT->voxelsPerLane};
    if (tl.V == 0) {
      tl.V = chooseTiling(X, Y, Z, elemBytes).V;
      if (tl.V >= 8 && tl.Ty == 8) tl.V = 4;  // the 8- and 16-voxel kernel stacks 2 or 4 row pairs
    }
    return tl;
  }
  return chooseTiling(X, Y, Z, elemBytes);
}

int checkTiling(const Tiling& tl, int elemBytes) {
  if (!tilingSupported(tl, elemBytes))
    return fail(MKF_ERR_ARG, "unsupported tiling planePairs=%d waves=%d rowPairs=%d voxelsPerLane=%d", tl.TP, tl.ZW,
                tl.Ty, tl.V);
  return MKF_OK;
}

}  // namespace

extern "C" {

const char* mkfLastError(void) { return gErr; }

int mkfSetNarrowRowMode(int mode) {
  if (mode != MKF_NARROW_CELLS && mode != MKF_NARROW_REFERENCE && mode != MKF_NARROW_REJECT)
    return fail(MKF_ERR_ARG, "narrow-row mode %d: use MKF_NARROW_CELLS, _REFERENCE or _REJECT", mode);
  return gNarrowMode.exchange(mode);
}

int mkfDebugInjectFault(const char* point, int nth) {
  if (!point || nth < 0) return fail(MKF_ERR_ARG, "fault injection: NULL point or negative count");
  static const char* const known[] = {"batch.upload", "batch.capture"};
  bool ok = false;
  for (const char* k : known) ok |= strcmp(k, point) == 0;
  if (!ok) return fail(MKF_ERR_ARG, "fault injection: unknown point '%s'", point);
  std::lock_guard<std::mutex> g(gFaultLock);
  if (nth == 0)
    gFaults.erase(point);
  else
    gFaults[point] = nth;
  return MKF_OK;
}

MkfTiling mkfDefaultTiling(const FieldDef def[3], int elemBytes) {
  if (!def) {
    (void)fail(MKF_ERR_ARG, "NULL dims");
    return MkfTiling{0, 0, 0, 0};
  }
  const Tiling tl = chooseTiling(def[0], def[1], def[2], elemBytes);
  return MkfTiling{tl.TP, tl.ZW, tl.Ty, tl.V};
}

int mkfPBMAsync(BMPackWord* pW, BMOrg* pO, const BMFieldInfo* pI, const BinMapF64* pM, void* stream) {
  if (!pW || !pM) return fail(MKF_ERR_ARG, "NULL PBM or map");
  FieldView fv;
  FieldDef def[3];
  int rc = resolveField(pI, &fv, def);
  if (rc) return rc;
  BMOrg org;
  setBMO(&org, def, 0);
  if (pO) *pO = org;
  const hipError_t e = launchPBM(fv, *pM, pW, org.rowWS, org.planeWS, (hipStream_t)stream);
  return e == hipSuccess ? MKF_OK : hipFail(e, "PBM launch");
}

int mkfPBMMomentsAsync(BMPackWord* pW, double* pMom, BMOrg* pO, const BMFieldInfo* pI, const BinMapF64* pM,
                       void* stream) {
  if (!pMom || !pM) return fail(MKF_ERR_ARG, "NULL moments or map");
  FieldView fv;
  FieldDef def[3];
  int rc = resolveField(pI, &fv, def);
  if (rc) return rc;
  BMOrg org;
  setBMO(&org, def, 0);
  if (pO) *pO = org;
  const hipStream_t s = (hipStream_t)stream;
  // Workgroup partials come from the stream-ordered pool, so calls on
  // different streams never share scratch.
  double* part = nullptr;
  const size_t bytes = (size_t)pbmMomentParts(fv, pW, org.rowWS, org.planeWS) * 6 * sizeof(double);
  hipError_t e = hipMallocAsync((void**)&part, bytes, s);
  if (e != hipSuccess) return hipFail(e, "moments scratch");
  e = launchPBMMoments(fv, *pM, pW, org.rowWS, org.planeWS, part, pMom, s);
  const hipError_t ef = hipFreeAsync(part, s);
  if (e != hipSuccess) return hipFail(e, "PBM moments launch");
  return ef == hipSuccess ? MKF_OK : hipFail(ef, "moments scratch free");
}

namespace {

// The PBM histogram's tiling: 16 voxels per lane (half words), 4 (nibbles of
// the PBM words) or 1, as asked; 0 takes the built-in choice's (4 where it
// is 16 and the rows asked for are 8).
Tiling bpfdTiling(const MkfTiling* pT, int X, int Y, int Z) {
  Tiling tl = chooseBpfdTiling(X, Y, Z);
  if (pT) {
    const int v = pT->voxelsPerLane;
    tl = Tiling{pT->planePairs, pT->waves, pT->rowPairs, v == 1 || v == 16 ? v : v == 0 ? tl.V : 4};
    if (v == 0 && tl.V == 16 && tl.Ty == 8) tl.V = 4;
  }
  return tl;
}

void toLaunchInfo(const LaunchInfo& li, MkfLaunchInfo* out) {
  out->tiles = li.tiles;
  out->workgroups = li.workgroups;
  out->wavesPerGroup = li.wavesPerGroup;
  out->used = MkfTiling{li.used.TP, li.used.ZW, li.used.Ty, li.used.V};
}

}  // namespace

int mkfBPFDAsync(size_t* pBPFD, const BMOrg* pO, const BMPackWord* pW, const MkfTiling* pT, void* stream) {
  if (!pBPFD || !pO || !pW) return fail(MKF_ERR_ARG, "NULL argument");
  const int X = (int)pO->rowElem, Y = (int)pO->rowPairs + 1, Z = (int)pO->planePairs + 1;
  if (X >= 2 && X < 32 && Y >= 2 && Z >= 2) {
    const int mode = narrowRowMode();
    if (mode == MKF_NARROW_REJECT)
      return fail(MKF_ERR_UNSUPPORTED, "rows of %d voxels: the reference's walk counts 2X-1 patterns per row pair "
                                       "(narrow-row mode: reject)", X);
    if (mode == MKF_NARROW_REFERENCE) {
      const hipError_t e = launchBPFDNarrowRef(pW, X, Y, Z, pO->rowWS, pO->planeWS, (unsigned long long*)pBPFD,
                                               (hipStream_t)stream);
      return e == hipSuccess ? MKF_OK : hipFail(e, "narrow-row reference walk");
    }
    if (getenv("MKF_VERBOSE"))
      fprintf(stderr, "[mkf] rows of %d voxels: counting their X-1 cells (the reference's walk counts 2X-1 patterns; "
                      "see mkfSetNarrowRowMode)\n", X);
  }
  const Tiling tl = bpfdTiling(pT, X, Y, Z);
  int rc = checkTiling(tl, 4);
  if (rc) return rc;
  const hipError_t e = launchBPFD(pW, X, Y, Z, pO->rowWS, pO->planeWS, (unsigned long long*)pBPFD, tl,
                                  (hipStream_t)stream);
  return e == hipSuccess ? MKF_OK : hipFail(e, "BPFD launch");
}

int mkfFusedBPFD(KernInfo* pK, size_t* pBPFD, BMOrg* pO, const BMFieldInfo* pI, const BinMapF64* pM,
                 const MkfTiling* pT, void* stream) {
  if (!pBPFD || !pM) return fail(MKF_ERR_ARG, "NULL histogram or map");
  FieldView fv;
  FieldDef def[3];
  int rc = resolveField(pI, &fv, def);
  if (rc) return rc;
  if (pO) setBMO(pO, def, 0);
  const Tiling tl = fusedTiling(pT, fv);
  rc = checkTiling(tl, fv.enc & ENC_MASK_NUM);
  if (rc) return rc;
  const hipStream_t s = (hipStream_t)stream;
  if (pK) {
    EventTimer t(s);
    const hipError_t e = launchFused(fv, *pM, (unsigned long long*)pBPFD, tl, s);
    pK->dtms[0] = 0.f;
    pK->dtms[1] = t.stopMs();
    if (e != hipSuccess) return hipFail(e, "fused launch");
    const hipError_t e2 = hipGetLastError();
    return e2 == hipSuccess ? MKF_OK : hipFail(e2, "fused kernel");
  }
  const hipError_t e = launchFused(fv, *pM, (unsigned long long*)pBPFD, tl, s);
  return e == hipSuccess ? MKF_OK : hipFail(e, "fused launch");
}

int mkfFusedLaunchInfo(const BMFieldInfo* pI, const BinMapF64* pM, const MkfTiling* pT, MkfLaunchInfo* pInfo) {
  if (!pInfo) return fail(MKF_ERR_ARG, "NULL launch info");
  FieldView fv;
  FieldDef def[3];
  int rc = resolveField(pI, &fv, def);
  if (rc) return rc;
  const Tiling tl = fusedTiling(pT, fv);
  rc = checkTiling(tl, fv.enc & ENC_MASK_NUM);
  if (rc) return rc;
  BinMapF64 m{};
  m.m = 6;
  LaunchInfo li;
  const hipError_t e = launchFused(fv, pM ? *pM : m, nullptr, tl, 0, &li);
  if (e != hipSuccess) return hipFail(e, "fused launch geometry");
  toLaunchInfo(li, pInfo);
  return MKF_OK;
}

int mkfBPFDLaunchInfo(const BMOrg* pO, const MkfTiling* pT, MkfLaunchInfo* pInfo) {
  if (!pO || !pInfo) return fail(MKF_ERR_ARG, "NULL argument");
  const int X = (int)pO->rowElem, Y = (int)pO->rowPairs + 1, Z = (int)pO->planePairs + 1;
  const Tiling tl = bpfdTiling(pT, X, Y, Z);
  int rc = checkTiling(tl, 4);
  if (rc) return rc;
  LaunchInfo li;
  const hipError_t e = launchBPFD(nullptr, X, Y, Z, pO->rowWS, pO->planeWS, nullptr, tl, 0, &li);
  if (e != hipSuccess) return hipFail(e, "histogram launch geometry");
  toLaunchInfo(li, pInfo);
  return MKF_OK;
}

int mkfFusedBPFDMoments(size_t* pBPFD, double* pMom, BMOrg* pO, const BMFieldInfo* pI, const BinMapF64* pM,
                        const MkfTiling* pT, void* stream) {
  if (!pBPFD || !pMom || !pM) return fail(MKF_ERR_ARG, "NULL histogram, moments or map");
  FieldView fv;
  FieldDef def[3];
  int rc = resolveField(pI, &fv, def);
  if (rc) return rc;
  BMOrg org;
  setBMO(&org, def, 0);
  if (pO) *pO = org;
  // The plain kernel's autotuned tiling does not carry over: with moments
  // the kernel wants more voxels per lane and more tiles.
  const Tiling tl = pT ? Tiling{pT->planePairs, pT->waves, pT->rowPairs, pT->voxelsPerLane}
                       : chooseMomTiling(fv.X, fv.Y, fv.Z, fv.enc & ENC_MASK_NUM);
  rc = checkTiling(tl, fv.enc & ENC_MASK_NUM);
  if (rc) return rc;
  const hipStream_t s = (hipStream_t)stream;
  // A field without cells adds nothing to the histogram; its moments come
  // from the map kernel (no map written).
  const long long nParts = fusedMomentParts(fv, tl, pM->m);
  const long long need = nParts ? nParts : pbmMomentParts(fv, nullptr, org.rowWS, org.planeWS);
  const size_t bytes = (size_t)need * 6 * sizeof(double);
  double* part = nullptr;
  hipError_t e = hipMallocAsync((void**)&part, bytes, s);
  if (e != hipSuccess) return hipFail(e, "moments scratch");
  if (nParts)
    e = launchFusedMoments(fv, *pM, (unsigned long long*)pBPFD, part, pMom, tl, s);
  else
    e = launchPBMMoments(fv, *pM, nullptr, org.rowWS, org.planeWS, part, pMom, s);
  const hipError_t ef = hipFreeAsync(part, s);
  if (e != hipSuccess) return hipFail(e, "fused moments launch");
  return ef == hipSuccess ? MKF_OK : hipFail(ef, "moments scratch free");
}

}  // extern "C"

namespace {

// Scratch of the allocation-free moments entry point: the per-workgroup
// moment partials (6 doubles each). Each device holds its own copy at its own
// address, looked up once per device with hipGetSymbolAddress; host code
// passes that address, never the array's name, to kernels.
constexpr size_t kNoAllocDoubles = MKF_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[kNoAllocDoubles];

std::mutex gScratchLock;
std::map<int, double*> gScratchPtr;

int scratchOf(double** p) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hipFail(e, "no-alloc scratch: current device");
  std::lock_guard<std::mutex> g(gScratchLock);
  auto it = gScratchPtr.find(dev);
  if (it == gScratchPtr.end()) {
    double* ptr = nullptr;
    e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return hipFail(e, "no-alloc scratch: symbol address");
    it = gScratchPtr.emplace(dev, ptr).first;
  }
  *p = it->second;
  return MKF_OK;
}

}  // namespace

extern "C" {

int mkfFusedBPFDMomentsNoAlloc(size_t* pBPFD, double* pMom, BMOrg* pO, const BMFieldInfo* pI, const BinMapF64* pM,
                               const MkfTiling* pT, void* stream) {
  if (!pBPFD || !pMom || !pM) return fail(MKF_ERR_ARG, "NULL histogram, moments or map");
  FieldView fv;
  FieldDef def[3];
  int rc = resolveField(pI, &fv, def);
  if (rc) return rc;
  BMOrg org;
  setBMO(&org, def, 0);
  if (pO) *pO = org;
  const Tiling tl = pT ? Tiling{pT->planePairs, pT->waves, pT->rowPairs, pT->voxelsPerLane}
                       : chooseMomTiling(fv.X, fv.Y, fv.Z, fv.enc & ENC_MASK_NUM);
  rc = checkTiling(tl, fv.enc & ENC_MASK_NUM);
  if (rc) return rc;
  const long long nParts = fusedMomentParts(fv, tl, pM->m);
  const long long need = nParts ? nParts : pbmMomentParts(fv, nullptr, org.rowWS, org.planeWS);
  const long long maxParts = (long long)(kNoAllocDoubles / MKF_MOM_N);
  if (need > maxParts)
    return fail(MKF_ERR_BUDGET, "no-alloc moments: %lld workgroup partials exceed the %lld that fit the %d-byte scratch",
                need, maxParts, (int)MKF_NOALLOC_SCRATCH_BYTES);
  double* part = nullptr;
  rc = scratchOf(&part);
  if (rc) return rc;
  // Every workgroup writes its partial before the reduction reads it, so
  // the scratch needs no clearing between calls.
  const hipStream_t s = (hipStream_t)stream;
  const hipError_t e = nParts ? launchFusedMoments(fv, *pM, (unsigned long long*)pBPFD, part, pMom, tl, s)
                              : launchPBMMoments(fv, *pM, nullptr, org.rowWS, org.planeWS, part, pMom, s);
  return e == hipSuccess ? MKF_OK : hipFail(e, "no-alloc fused moments launch");
}

}  // extern "C"

namespace {

// Streaming read of n 16-byte words, for the practical HBM read roof. A
// block is U loads of 1 KiB (64 lanes x 16 bytes), all in flight at once, so
// a wave streams U KiB of consecutive bytes per step. Round robin (CH false):
// wave w takes blocks w, w + waves, ..., so all waves read one moving
// window. Chunked (CH true): wave w takes a contiguous run of blocks, so the
// waves stream as many separate regions at once. The words are folded into
// an XOR, so no load can be dropped; one atomic per wave publishes it.
template <bool NT, bool CH, int U>
__global__ __launch_bounds__(256) void kReadRoof(const uint4* __restrict__ p, size_t n, unsigned* __restrict__ out) {
  typedef unsigned V4 __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(1))) V4* GV4;
  const int lane = threadIdx.x & 63;
  const size_t waves = (size_t)gridDim.x * (blockDim.x >> 6);
  const size_t wave = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const size_t blk = (size_t)64 * U;
  const size_t full = n / blk;  // whole blocks
  auto ld = [&](size_t k) -> V4 {
    if constexpr (NT) return __builtin_nontemporal_load((GV4)(p + k));
    else return *(const V4*)(p + k);
  };
  size_t b0 = wave, b1 = full, step = waves;
  if constexpr (CH) {
    const size_t per = (full + waves - 1) / waves;
    b0 = std::min(full, wave * per);
    b1 = std::min(full, b0 + per);
    step = 1;
  }
  V4 acc = {0u, 0u, 0u, 0u};
  for (size_t b = b0; b < b1; b += step) {
    V4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = ld(b * blk + (size_t)u * 64 + lane);
#pragma unroll
    for (int u = 0; u < U; ++u) acc ^= v[u];
  }
  // the words after the last whole block: word by word, by the last wave
  if (wave == waves - 1)
    for (size_t k = full * blk + lane; k < n; k += 64) acc ^= ld(k);
  unsigned x = acc.x ^ acc.y ^ acc.z ^ acc.w;
  for (int o = 32; o > 0; o >>= 1) x ^= __shfl_xor(x, o);
  if (lane == 0) atomicXor(out, x);
}

template <bool NT, bool CH>
void launchReadRoof(int U, unsigned grid, hipStream_t s, const uint4* p, size_t n, unsigned* out) {
  if (U == 4)
    hipLaunchKernelGGL((kReadRoof<NT, CH, 4>), dim3(grid), dim3(256), 0, s, p, n, out);
  else if (U == 8)
    hipLaunchKernelGGL((kReadRoof<NT, CH, 8>), dim3(grid), dim3(256), 0, s, p, n, out);
  else
    hipLaunchKernelGGL((kReadRoof<NT, CH, 16>), dim3(grid), dim3(256), 0, s, p, n, out);
}

}  // namespace

extern "C" {

int mkfReadRoof(const void* p, size_t bytes, int flags, int loadsInFlight, int blocksPerCU, unsigned* pOut,
                void* stream) {
  if (!p || !pOut || bytes < 16 || (bytes & 15) || ((uintptr_t)p & 15) || blocksPerCU < 1 || blocksPerCU > 64 ||
      (loadsInFlight != 4 && loadsInFlight != 8 && loadsInFlight != 16) || (flags & ~3))
    return fail(MKF_ERR_ARG,
                "read roof: need 16-byte aligned memory, a multiple of 16 bytes, flags in 0..3, 4/8/16 loads in "
                "flight and 1..64 blocks per CU");
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
  const size_t n = bytes / 16;
  const unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>((size_t)cus * blocksPerCU, (n + 255) / 256));
  const hipStream_t s = (hipStream_t)stream;
  const uint4* q = (const uint4*)p;
  switch (flags) {
    case 0: launchReadRoof<false, false>(loadsInFlight, grid, s, q, n, pOut); break;
    case 1: launchReadRoof<true, false>(loadsInFlight, grid, s, q, n, pOut); break;
    case 2: launchReadRoof<false, true>(loadsInFlight, grid, s, q, n, pOut); break;
    default: launchReadRoof<true, true>(loadsInFlight, grid, s, q, n, pOut); break;
  }
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? MKF_OK : hipFail(e, "read roof launch");
}

int mkfFusedAutotune(const BMFieldInfo* pI, const BinMapF64* pM, MkfTiling* pChosen, void* stream) {
  if (!pM) return fail(MKF_ERR_ARG, "NULL map");
  FieldView fv;
  FieldDef def[3];
  int rc = resolveField(pI, &fv, def);
  if (rc) return rc;
  const int eb = fv.enc & ENC_MASK_NUM;
  // The built-in choice and the shapes that came out ahead of it on some
  // MI355X at 1024^3; several fields or an x stride run one voxel per lane.
  // Float fields also get the persistent 8- and 16-voxel kernels, whichever
  // the built-in choice is not, and each of them with its slab depth spread
  // over three quarters of its resident slots: fewer, longer waves per CU,
  // which streamed faster on some shapes (1024^3: 0.642 against 0.668 ms, one
  // tile per wave either way) and slower on others (1000^3: 0.633 against
  // 0.612).
  std::vector<Tiling> cand{chooseTiling(fv.X, fv.Y, fv.Z, eb)};
  if (eb == 4) {
    if (fv.X >= 4 && fv.nF == 1 && fv.sx == 1) {
      if (cand[0].V != 16) cand.push_back(choose16Tiling(fv.X, fv.Y, fv.Z, 16));
      if (cand[0].V != 8) cand.push_back(choose16Tiling(fv.X, fv.Y, fv.Z, 8));
      for (int v : {16, 8}) {
        const Tiling t = choose16Tiling(fv.X, fv.Y, fv.Z, v, 3);
        bool dup = false;
        for (const Tiling& c : cand) dup |= c.V == t.V && c.TP == t.TP && c.Ty == t.Ty && c.ZW == t.ZW;
        if (!dup) cand.push_back(t);
      }
    }
    // 4 voxels per lane at the built-in depth model (mid-size fields where
    // the persistent kernels are chosen can run faster there: 1024 x 512 x
    // 512 0.167 ms at 32 plane pairs against 0.183) and at 256 plane pairs.
    const Tiling t4 = choose4Tiling(fv.X, fv.Y, fv.Z);
    if (!(t4.V == cand[0].V && t4.TP == cand[0].TP && t4.ZW == cand[0].ZW)) cand.push_back(t4);
    if (!(t4.TP == 256 && t4.ZW == 16)) cand.push_back(Tiling{256, 16, 4, 4});
    cand.push_back(Tiling{128, 16, 4, 1});
  } else {
    cand.push_back(Tiling{512, 16, 8, 1});
    cand.push_back(Tiling{256, 16, 4, 2});
  }
  const hipStream_t s = (hipStream_t)stream;
  unsigned long long* scratch = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = MKF_OK;
  if (hipMalloc((void**)&scratch, MKF_BINS * 8) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    if (scratch) (void)hipFree(scratch);
    if (e0) (void)hipEventDestroy(e0);
    return fail(MKF_ERR_MEM, "autotune scratch");
  }
  // Warm-up of about 0.3 s so the clocks have left idle, then three rounds
  // over the candidates; a candidate scores its best round.
  hipError_t e = hipSuccess;
  {
    float spent = 0.f;
    (void)hipEventRecord(e0, s);
    while (e == hipSuccess && spent < 300.f) {
      for (int i = 0; i < 8 && e == hipSuccess; ++i) e = launchFused(fv, *pM, scratch, cand[0], s);
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) break;
      (void)hipEventElapsedTime(&spent, e0, e1);
    }
  }
  std::vector<float> score(cand.size(), 1e30f);
  for (int round = 0; round < 3 && e == hipSuccess; ++round) {
    for (size_t c = 0; c < cand.size() && e == hipSuccess; ++c) {
      if (!tilingSupported(cand[c], eb)) continue;
      e = launchFused(fv, *pM, scratch, cand[c], s);
      (void)hipEventRecord(e0, s);
      for (int i = 0; i < 5 && e == hipSuccess; ++i) e = launchFused(fv, *pM, scratch, cand[c], s);
      (void)hipEventRecord(e1, s);
      if (e != hipSuccess || hipEventSynchronize(e1) != hipSuccess) break;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      score[c] = std::min(score[c], ms);
    }
  }
  if (e != hipSuccess) rc = hipFail(e, "autotune launch");
  Tiling best = cand[0];
  float bestMs = 1e30f;
  for (size_t c = 0; c < cand.size(); ++c)
    if (score[c] < bestMs) {
      bestMs = score[c];
      best = cand[c];
    }
  (void)hipFree(scratch);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != MKF_OK) return rc;
  {
    std::lock_guard<std::mutex> g(gTuneLock);
    gTuned[keyOf(fv)] = best;
  }
  if (pChosen) *pChosen = MkfTiling{best.TP, best.ZW, best.Ty, best.V};
  return MKF_OK;
}

// src/binMapCUDA.cu:145-264.
BMPackWord* binMapCUDA(KernInfo* pK, BMPackWord* pW, BMOrg* pO, const BMFieldInfo* pI, const BinMapF64* pM) {
  FieldView fv;
  FieldDef def[3];
  if (!pM || resolveField(pI, &fv, def) != MKF_OK) return NULL;
  BMOrg org;
  const size_t words = setBMO(&org, def, (char)pI->profID);
  if (0 == words) return NULL;
  if (pI->profID & 0x0F) {
    // Partial organisation (src/binMapUtil.c:60-87): a buffer of min(2, Y)
    // rows of min(2, Z) planes. Map exactly those rows and planes of the
    // field at the table's pointer (same strides), so the words written fit
    // the buffer; stepping the pointer by rows and planes walks the field in
    // row-pair windows. (The reference's kernels write the whole field into
    // this buffer.)
    fv.Y = (int)org.rowPairs + 1;
    fv.Z = (int)org.planePairs + 1;
  }
  if (pO) *pO = org;
  if (NULL == pW) {
    std::lock_guard<std::mutex> g(gLock);
    if (gDevWWords < words) {
      if (gDevW) hipFree(gDevW);
      gDevW = nullptr;
      gDevWWords = 0;
      if (hipMalloc((void**)&gDevW, words * sizeof(BMPackWord)) != hipSuccess) {
        gDevW = nullptr;
        fail(MKF_ERR_MEM, "PBM buffer of %zu words", words);
        return NULL;
      }
      gDevWWords = words;
    }
    pW = gDevW;
  }
  EventTimer t(0);
  const hipError_t e = launchPBM(fv, *pM, pW, org.rowWS, org.planeWS, 0);
  const float ms = t.stopMs();
  if (pK) {
    pK->dtms[0] = ms;
  } else if (getenv("MKF_VERBOSE")) {
    fprintf(stderr, "[mkf] binMapCUDA() - dt= %Gms\n", ms);
  }
  const hipError_t e2 = hipGetLastError();
  if (e != hipSuccess || e2 != hipSuccess) {
    hipFail(e != hipSuccess ? e : e2, "binMapCUDA");
    return NULL;
  }
  return pW;
}

void binMapCUDACleanup(void) {
  std::lock_guard<std::mutex> g(gLock);
  if (gDevW) hipFree(gDevW);
  gDevW = nullptr;
  gDevWWords = 0;
}

// src/mkfCUDA.cu:13-61: every profile gives the same counts; all use the
// same kernel here.
size_t* mkfCUDAGetBPFD(KernInfo* pK, size_t* pBPFD, const BMOrg* pO, const BMPackWord* pW, const int profile) {
  (void)profile;
  if (pBPFD && pO && pW) {
    EventTimer t(0);
    const int rc = mkfBPFDAsync(pBPFD, pO, pW, NULL, 0);
    if (rc == MKF_ERR_UNSUPPORTED) return NULL;  // narrow rows refused: histogram untouched
    const float ms = t.stopMs();
    if (pK) {
      pK->dtms[1] = ms;
    } else if (getenv("MKF_VERBOSE")) {
      fprintf(stderr, "[mkf] mkfCUDAGetBPFD() - dt= %Gms\n", ms);
    }
    const hipError_t e = hipGetLastError();
    if (rc != MKF_OK || e != hipSuccess) fprintf(stderr, "[mkf] mkfCUDAGetBPFD: %s\n", gErr);
  }
  return pBPFD;
}

// src/mkfCUDA.cu:73-91.
size_t* mkfCUDAGetBPFDH(KernInfo* pK, size_t* pBPFDH, const BMOrg* pO, const BMPackWord* pW, const int profile) {
  const size_t bytes = sizeof(size_t) * MKF_BINS;
  {
    std::lock_guard<std::mutex> g(gLock);
    if (NULL == pBPFDH) {
      if (NULL == gHostBPFD && hipHostMalloc((void**)&gHostBPFD, bytes, hipHostMallocDefault) != hipSuccess)
        gHostBPFD = nullptr;
      pBPFDH = gHostBPFD;
    }
    if (NULL == gDevBPFD && hipMalloc((void**)&gDevBPFD, bytes) != hipSuccess) gDevBPFD = nullptr;
  }
  if (pBPFDH && gDevBPFD && pO && pW) {
    if (hipMemset(gDevBPFD, 0, bytes) != hipSuccess) return NULL;
    if (mkfCUDAGetBPFD(pK, (size_t*)gDevBPFD, pO, pW, profile)) {
      if (hipMemcpy(pBPFDH, gDevBPFD, bytes, hipMemcpyDeviceToHost) == hipSuccess) return pBPFDH;
    }
  }
  return NULL;
}

void mkfCUDACleanup(int lvl) {
  {
    std::lock_guard<std::mutex> g(gLock);
    if (gHostBPFD) hipHostFree(gHostBPFD);
    if (gDevBPFD) hipFree(gDevBPFD);
    gHostBPFD = nullptr;
    gDevBPFD = nullptr;
  }
  if (lvl > 0) {
    binMapCUDACleanup();
    hipDeviceReset();
  }
}

// src/mkfCUDA.cu:102-167 (deprecated all-in-one path), same buffer protocol:
// host field -> device, PBM (optionally copied back), histogram into pDZ
// (copied back to pHZ when given). Returns 1, or 0 when the PBM failed.
int mkfCUDAGetBPFDautoCtx(Context* pC, const FieldDef def[3], const BinMapF64* pM, const int profile) {
  if (!pC || !def || !pM) return fail(MKF_ERR_ARG, "NULL context, dims or map"), 0;
  if (pC->pHF) {
    if (NULL == pC->pDF && hipMalloc(&(pC->pDF), pC->bytesF) != hipSuccess) pC->pDF = NULL;
    if (pC->pDF) hipMemcpy(pC->pDF, pC->pHF, pC->bytesF, hipMemcpyHostToDevice);
    if (NULL == pC->pDU && hipMalloc((void**)&(pC->pDU), pC->bytesU) != hipSuccess) pC->pDU = NULL;
    if (pC->pDF && pC->pDU) {
      ConstFieldPtr fieldPtr[BMCU_FIELD_MAX];
      BMFieldInfo fi;
      setupFields(&fi, autoFieldPtr(fieldPtr, pC), pC->nField, def, pC->bytesElem, profile);
      if (NULL == binMapCUDA(&(pC->ki), pC->pDU, &(pC->bmo), &fi, pM)) return 0;
      if (pC->pHU) hipMemcpy(pC->pHU, pC->pDU, pC->bytesU, hipMemcpyDeviceToHost);
    }
  } else if (pC->pHU) {
    hipMemcpy(pC->pDU, pC->pHU, pC->bytesU, hipMemcpyHostToDevice);
    if (NULL == pC->pDZ && hipMalloc(&(pC->pDZ), pC->bytesZ) != hipSuccess) pC->pDZ = NULL;
  }
  if (pC->pDU && pC->pDZ) {
    hipMemset(pC->pDZ, 0, pC->bytesZ);
    mkfCUDAGetBPFD(&(pC->ki), (size_t*)(pC->pDZ), &(pC->bmo), pC->pDU, MKFCU_PROFILE_FAST);
    if (pC->pHZ) hipMemcpy(pC->pHZ, pC->pDZ, pC->bytesZ, hipMemcpyDeviceToHost);
  }
  return 1;
}

// src/mkfACC.c:206-239: host double field in; host PBM (full organisation,
// zero tail bits) and host histogram out; computed on the GPU.
Bool32 mkfAccGetBPFDSimple(size_t rBPFD[MKF_BINS], BMPackWord* pW, const double* pF, const FieldDef def[3],
                           const BinMapF64* pM) {
  if (!rBPFD || !pW || !pF || !def || !pM) return 0;
  if (def[0] < 32 && narrowRowMode() != MKF_NARROW_CELLS) {
    // The reference's CPU walk reads the next row's word for such rows (and
    // past the map on the last one): there is no defined result to reproduce.
    fail(MKF_ERR_UNSUPPORTED, "mkfAccGetBPFDSimple: rows of %d voxels (narrow-row mode %d)", def[0], narrowRowMode());
    return 0;
  }
  memset(rBPFD, 0, sizeof(size_t) * MKF_BINS);
  const size_t nF = (size_t)def[0] * def[1] * def[2];
  BMOrg org;
  const size_t words = setBMO(&org, def, 0);
  void* dF = nullptr;
  BMPackWord* dW = nullptr;
  unsigned long long* dH = nullptr;
  Bool32 ok = 0;
  if (hipMalloc(&dF, nF * sizeof(double)) == hipSuccess && hipMalloc((void**)&dW, words * 4) == hipSuccess &&
      hipMalloc((void**)&dH, MKF_BINS * 8) == hipSuccess &&
      hipMemcpy(dF, pF, nF * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
      hipMemset(dH, 0, MKF_BINS * 8) == hipSuccess) {
    const FieldView fv = planarField(dF, ENC_F64, def[0], def[1], def[2]);
    const Tiling tl = chooseBpfdTiling(fv.X, fv.Y, fv.Z);
    if (launchPBM(fv, *pM, dW, org.rowWS, org.planeWS, 0) == hipSuccess &&
        launchBPFD(dW, fv.X, fv.Y, fv.Z, org.rowWS, org.planeWS, dH, tl, 0) == hipSuccess &&
        hipMemcpy(pW, dW, words * 4, hipMemcpyDeviceToHost) == hipSuccess &&
        hipMemcpy(rBPFD, dH, MKF_BINS * 8, hipMemcpyDeviceToHost) == hipSuccess) {
      ok = 1;
    } else {
      hipFail(hipGetLastError(), "mkfAccGetBPFDSimple");
    }
  } else {
    fail(MKF_ERR_MEM, "mkfAccGetBPFDSimple: device buffers for %zu voxels", nF);
  }
  if (dF) hipFree(dF);
  if (dW) hipFree(dW);
  if (dH) hipFree(dH);
  return ok;
}

// src/mkfACC.c:248-318: host field of either encoding in, host histogram
// out. The PBM is not needed by the caller, so the field is streamed through
// the fused kernel in z-slabs (mkfHostStreamBPFD): bounded device memory,
// copies overlapping the kernel.
Bool32 mkfAccCUDAGetBPFD(size_t rBPFD[MKF_BINS], const void* pF, const FieldDef def[3], const NumEnc enc,
                         const BinMapF64* pM) {
  if (!rBPFD || !pF || !def || !pM || (enc != ENC_F32 && enc != ENC_F64)) return 0;
  FieldDef d[3];
  const int nd = copyNonZeroDef(d, def, 3);
  if (nd > 0 && d[0] < 32 && d[1] >= 2 && d[2] >= 2 && narrowRowMode() != MKF_NARROW_CELLS) {
    // The reference maps the field and walks the map on the GPU: for rows of
    // X < 32 its walk counts 2X-1 patterns per row pair. Reproduce (or refuse)
    // that with the whole field on the device.
    if (narrowRowMode() == MKF_NARROW_REJECT) {
      fail(MKF_ERR_UNSUPPORTED, "mkfAccCUDAGetBPFD: rows of %d voxels (narrow-row mode: reject)", d[0]);
      return 0;
    }
    memset(rBPFD, 0, sizeof(size_t) * MKF_BINS);
    const size_t eb = enc == ENC_F32 ? 4 : 8, nF = (size_t)d[0] * d[1] * d[2];
    BMOrg org;
    const size_t words = setBMO(&org, d, 0);
    void* dF = nullptr;
    BMPackWord* dW = nullptr;
    unsigned long long* dH = nullptr;
    Bool32 ok = 0;
    if (hipMalloc(&dF, nF * eb) == hipSuccess && hipMalloc((void**)&dW, words * 4) == hipSuccess &&
        hipMalloc((void**)&dH, MKF_BINS * 8) == hipSuccess &&
        hipMemcpy(dF, pF, nF * eb, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemset(dH, 0, MKF_BINS * 8) == hipSuccess) {
      const FieldView fv = planarField(dF, enc, d[0], d[1], d[2]);
      ok = launchPBM(fv, *pM, dW, org.rowWS, org.planeWS, 0) == hipSuccess &&
           launchBPFDNarrowRef(dW, d[0], d[1], d[2], org.rowWS, org.planeWS, dH, 0) == hipSuccess &&
           hipMemcpy(rBPFD, dH, MKF_BINS * 8, hipMemcpyDeviceToHost) == hipSuccess;
      if (!ok) hipFail(hipGetLastError(), "mkfAccCUDAGetBPFD");
    } else {
      fail(MKF_ERR_MEM, "mkfAccCUDAGetBPFD: device buffers for %zu voxels", nF);
    }
    if (dF) hipFree(dF);
    if (dW) hipFree(dW);
    if (dH) hipFree(dH);
    return ok;
  }
  memset(rBPFD, 0, sizeof(size_t) * MKF_BINS);
  return mkfHostStreamBPFD(rBPFD, pF, enc, def, pM, 0) == MKF_OK;
}

}  // extern "C"
