// mkf_stream.hip - host fields streamed through the device in z-slabs (host
// code; the kernels are the fused histogram kernels).
//
// Slab k holds planes [z0, z1] (z1 = z0 + slab plane pairs, clamped), so
// consecutive slabs share one plane and every plane pair lies in exactly one
// slab. Two device slab buffers alternate: the copy of slab k+1 on the copy
// stream overlaps the fused kernel on slab k on the compute stream; a copy
// into a buffer waits for the kernel that last read it. Device memory is two
// slabs and one histogram, whatever the field's size.
#include <string.h>

#include <algorithm>

#include "mkf_internal.h"

using namespace mkf;

namespace {

constexpr size_t kDefaultSlabBytes = 256u << 20;

struct StreamRes {
  void* buf[2] = {nullptr, nullptr};
  unsigned long long* dH = nullptr;
  hipStream_t compute = nullptr, copy = nullptr;
  hipEvent_t copied[2] = {nullptr, nullptr}, used[2] = {nullptr, nullptr};

  ~StreamRes() {
    if (compute) (void)hipStreamSynchronize(compute);
    if (copy) (void)hipStreamSynchronize(copy);
    for (int b = 0; b < 2; ++b) {
      if (buf[b]) (void)hipFree(buf[b]);
      if (copied[b]) (void)hipEventDestroy(copied[b]);
      if (used[b]) (void)hipEventDestroy(used[b]);
    }
    if (dH) (void)hipFree(dH);
    if (compute) (void)hipStreamDestroy(compute);
    if (copy) (void)hipStreamDestroy(copy);
  }
};

}  // namespace

extern "C" {

int mkfHostStreamBPFD(size_t* pBPFDH, const void* pF, NumEnc enc, const FieldDef def[3], const BinMapF64* pM,
                      int slabPlanes) {
  if (!pBPFDH || !pF || !def || !pM || (enc != ENC_F32 && enc != ENC_F64))
    return setError(MKF_ERR_ARG, "host stream: NULL argument or encoding 0x%02X", enc);
  FieldDef d[3];
  if (copyNonZeroDef(d, def, 3) <= 0) return setError(MKF_ERR_ARG, "host stream: no dims");
  const int X = d[0], Y = d[1], Z = d[2];
  if (X < 2 || Y < 2 || Z < 2) return MKF_OK;  // no cells
  const size_t eb = enc & ENC_MASK_NUM;
  const size_t planeBytes = (size_t)X * Y * eb;
  long long sp = slabPlanes > 0 ? slabPlanes : (long long)std::max<size_t>(1, kDefaultSlabBytes / planeBytes);
  sp = std::min<long long>(sp, Z - 1);
  const size_t slabBytes = planeBytes * (size_t)(sp + 1);

  StreamRes R;
  for (int b = 0; b < 2; ++b)
    if (hipMalloc(&R.buf[b], slabBytes) != hipSuccess ||
        hipEventCreateWithFlags(&R.copied[b], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&R.used[b], hipEventDisableTiming) != hipSuccess)
      return setError(MKF_ERR_MEM, "host stream: slab buffers of %zu bytes", slabBytes);
  if (hipMalloc((void**)&R.dH, MKF_BINS * sizeof(unsigned long long)) != hipSuccess ||
      hipStreamCreateWithFlags(&R.compute, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&R.copy, hipStreamNonBlocking) != hipSuccess ||
      hipMemsetAsync(R.dH, 0, MKF_BINS * sizeof(unsigned long long), R.compute) != hipSuccess)
    return setError(MKF_ERR_MEM, "host stream: histogram and streams");

  const char* src = (const char*)pF;
  int k = 0;
  for (long long z0 = 0; z0 < Z - 1; z0 += sp, ++k) {
    const int b = k & 1;
    const long long z1 = std::min<long long>(z0 + sp, Z - 1);
    const int np = (int)(z1 - z0 + 1);
    if (k >= 2 && hipStreamWaitEvent(R.copy, R.used[b], 0) != hipSuccess)
      return setError(MKF_ERR_HIP, "host stream: wait");
    if (hipMemcpyAsync(R.buf[b], src + (size_t)z0 * planeBytes, (size_t)np * planeBytes, hipMemcpyHostToDevice,
                       R.copy) != hipSuccess ||
        hipEventRecord(R.copied[b], R.copy) != hipSuccess || hipStreamWaitEvent(R.compute, R.copied[b], 0) != hipSuccess)
      return setError(MKF_ERR_HIP, "host stream: slab copy");
    const FieldView fv = planarField(R.buf[b], enc, X, Y, np);
    const hipError_t e = launchFused(fv, *pM, R.dH, chooseTiling(X, Y, np, (int)eb), R.compute);
    if (e != hipSuccess || hipEventRecord(R.used[b], R.compute) != hipSuccess)
      return setError(MKF_ERR_HIP, "host stream: slab kernel (%s)", hipGetErrorString(e));
  }
  unsigned long long h[MKF_BINS];
  if (hipMemcpyAsync(h, R.dH, sizeof(h), hipMemcpyDeviceToHost, R.compute) != hipSuccess ||
      hipStreamSynchronize(R.compute) != hipSuccess)
    return setError(MKF_ERR_HIP, "host stream: result");
  for (int i = 0; i < MKF_BINS; ++i) pBPFDH[i] += (size_t)h[i];
  return MKF_OK;
}

}  // extern "C"
