/* mkf_ext.h - additive entry points of libmkf_amd.so (no reference
 * counterparts). Included by mkf.h.
 *
 * The reference's entry points block on every call and share static buffers,
 * so they are neither re-entrant nor stream-aware. The functions here take a
 * HIP stream (hipStream_t passed as void*, NULL = the legacy default stream),
 * return an error code, and never block unless timing is requested through a
 * KernInfo pointer.
 */
#ifndef MKF_AMD_EXT_H
#define MKF_AMD_EXT_H

#ifdef __cplusplus
extern "C" {
#endif

#define MKF_OK 0
#define MKF_ERR_ARG (-1)         /* invalid argument */
#define MKF_ERR_MEM (-2)         /* device or host allocation failed */
#define MKF_ERR_HIP (-3)         /* HIP runtime or launch error */
#define MKF_ERR_UNSUPPORTED (-4) /* valid for the reference, not supported here */
#define MKF_ERR_BUDGET (-5)      /* request exceeds the memory budget */

/* Work decomposition of the histogram kernels; NULL selects the built-in
 * choice for the field's shape. A workgroup of `waves` (1..16) waves sits
 * side by side along x, each lane holding voxelsPerLane consecutive voxels
 * (1, or 4 for float / 2 for double fields; 0 = the library's choice; the
 * PBM histogram always uses 1), holds rowPairs (2, 4 or 8) row pairs and
 * marches over planePairs (>= 1) plane pairs. */
typedef struct {
  int32_t planePairs;
  int32_t waves;
  int32_t rowPairs;
  int32_t voxelsPerLane;
} MkfTiling;

/* Text of the last error raised by any entry point in this thread. */
const char *mkfLastError(void);

/* Fused threshold + histogram: ADDS the BPFD of the field (device pointers in
 * pI, as for binMapCUDA) to pBPFD (device, 256 x uint64), reading each field
 * byte from HBM once and writing no packed map. pO (optional) receives the
 * PBM organisation. With pK the call blocks and pK->dtms[1] receives the
 * kernel time (dtms[0] = 0); without it the work is only queued on stream. */
int mkfFusedBPFD(KernInfo *pK, size_t *pBPFD, BMOrg *pO, const BMFieldInfo *pI, const BinMapF64 *pM,
                 const MkfTiling *pT, void *stream);

/* Stream-ordered packed binary map into caller memory pW (device). */
int mkfPBMAsync(BMPackWord *pW, BMOrg *pO, const BMFieldInfo *pI, const BinMapF64 *pM, void *stream);

/* Stream-ordered histogram of a packed binary map; adds into pBPFD. */
int mkfBPFDAsync(size_t *pBPFD, const BMOrg *pO, const BMPackWord *pW, const MkfTiling *pT, void *stream);

/* Tiling the library would choose for a field of dims def and element size. */
MkfTiling mkfDefaultTiling(const FieldDef def[3], int elemBytes);

/* Read-only views of the measure tables used by mkfSelectMeasuresBPFD: the
 * symmetry group of each of the 256 patterns and the Euler weight of each of
 * the 22 groups (src/mkfUtil.c:33-58), for callers that verify them. */
const uint8_t *mkfPatternGroupTable(void);
const int8_t *mkfGroupWeightKTable(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* MKF_AMD_EXT_H */
