/* mkf.h - C ABI of the MI355X (gfx950) Minkowski-functional library.
 *
 * One header for C callers of DrAl-HFS/MKF's CUDA library. It declares the
 * reference's types with the same layouts and its entry points with the same
 * names, argument meaning and error behaviour, implemented by hand-written HIP
 * kernels in libmkf_amd.so. Pointers are plain C pointers; no HIP or CUDA type
 * appears in any signature (streams are passed as `void *`, a hipStream_t).
 *
 * Citations are to the reference tree (path:line).
 *
 * Pipeline: scalar field --threshold--> packed binary map (PBM)
 *           --2x2x2 cell codes--> 256-bin pattern histogram (BPFD)
 *           --weights--> measures (Kv, Vv).
 */
#ifndef MKF_AMD_H
#define MKF_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Element encoding  (src/encoding.h:15-31)
 * ------------------------------------------------------------------------- */
#define ENC_MASK_TYP 0xF0
#define ENC_MASK_NUM 0x0F
#define ENC_TYPE_FPB 0xF0
#define ENC_TYPE_INT 0xE0
#define ENC_TYPE_BTS 0x00
#define ENC_F32 (ENC_TYPE_FPB | 4)
#define ENC_F64 (ENC_TYPE_FPB | 8)
#define ENC_U8 (ENC_TYPE_BTS | 8)
#define ENC_U1 (ENC_TYPE_BTS | 1)
#define ENC_FLOAT ENC_F32
#define ENC_DOUBLE ENC_F64
typedef uint8_t NumEnc;

/* Bytes of n elements; *pElemBits (optional) receives bits per element.
 * Replaces encSizeN, src/encoding.c:9-25. */
size_t encSizeN(int *pElemBits, const size_t n, const NumEnc e);
/* Replaces encAlignment, src/encoding.c:27-40. */
int encAlignment(const NumEnc e);

/* ---------------------------------------------------------------------------
 * Threshold descriptor and PBM organisation  (src/binMapUtil.h:15-79)
 * ------------------------------------------------------------------------- */
#define BM_NUMT (1)
#define BM_TLB (0)
#define BM_TUB (0)
#define BM_FLAG_T2IVL (1 << 15)

/* One threshold t and a 3-bit relation mask m: bit LT(1), EQ(2), GT(4).
 * A value f maps to bit (m >> d) & 1 with d = 1 - (f<t) + (f>t), so NaN
 * lands on EQ.  src/binMapUtil.h:28-32, src/binMapCUDA.cuh:141-145. */
typedef struct {
  double t[BM_NUMT];
  uint32_t m;
} BinMapF64;

#define BMC_GT (0x4)
#define BMC_EQ (0x2)
#define BMC_LT (0x1)
#define BMC_NV (0x0)
#define BMC_AL (0x7)

typedef uint32_t BMStride;
typedef uint32_t BMPackWord;

/* PBM memory organisation: bit x&31 of word (x>>5) + y*rowWS + z*planeWS.
 * src/binMapUtil.h:52-57. */
typedef struct {
  uint32_t rowElem;
  uint32_t rowPairs, planePairs;
  BMStride rowWS, planeWS;
} BMOrg;

typedef union {
  const void *p;
  const float *pF32;
  const double *pF64;
  size_t w;
} ConstFieldPtr;

typedef long int FieldStride;
typedef int FieldDef;

/* Field description (LP64: 32 bytes).  src/binMapUtil.h:69-79. */
typedef struct {
  struct {
    uint32_t fieldTableMask;
    NumEnc elemID;
    uint8_t oprID, profID, flags;
  };
  const FieldDef *pD;
  const FieldStride *pS;
  ConstFieldPtr *pFieldDevPtrTable;
} BMFieldInfo;

#define BMCU_FIELD_MAX 4

/* Per-call kernel times in ms: dtms[0] PBM, dtms[1] BPFD.
 * src/binMapUtil.h:107-110. */
typedef struct {
  float dtms[2];
} KernInfo;

/* Legacy all-in-one context (deprecated in the reference).
 * src/binMapUtil.h:113-125. */
typedef struct {
  int nField, nElem;
  struct {
    NumEnc enc;
    uint8_t bytesElem, pad[2];
  };
  long bytesF;
  long nU, bytesU;
  long nZ, bytesZ;
  void *pDF, *pHF;
  BMPackWord *pDU, *pHU;
  void *pDZ, *pHZ;
  BMOrg bmo;
  KernInfo ki;
} Context;

typedef int32_t Bool32;

/* Replaces setBinMapF64, src/binMapUtil.c:37-58: "<", "=", ">" set bits,
 * a leading "!" inverts; parsing stops at the first other character. */
BinMapF64 *setBinMapF64(BinMapF64 *pBM64, const char relopChars[], const float t);
/* Replaces setBMO, src/binMapUtil.c:60-87. Returns the buffer size in words. */
size_t setBMO(BMOrg *pO, const FieldDef def[3], const char profID);
/* Replaces copyNonZeroDef, src/binMapUtil.c:89-103. */
int copyNonZeroDef(FieldDef d[], const FieldDef *pD, const int nD);
/* Replaces copyValidPtrByMask, src/binMapUtil.c:107-121. */
int copyValidPtrByMask(ConstFieldPtr r[], const int max, const ConstFieldPtr a[], const uint32_t mask);
/* Replaces countValidPtrByMask, src/binMapUtil.c:123-133. */
int countValidPtrByMask(const ConstFieldPtr a[], uint32_t mask);
/* Replaces genStride, src/binMapUtil.c:135-144. */
int genStride(FieldStride fs[], const int n, const int start, const FieldDef *pD, FieldStride stride);
/* Replaces copyOrGenStride, src/binMapUtil.c:146-155. */
int copyOrGenStride(FieldStride fs[], const int n, const int start, const FieldDef *pD, const FieldStride *pS);
/* Replaces asFieldTab, src/binMapUtil.c:158-164. */
ConstFieldPtr *asFieldTab(const void **pp, const NumEnc id);
/* Replaces setupFields, src/binMapUtil.c:166-186. */
const BMFieldInfo *setupFields(BMFieldInfo *pI, void **ppF, const int nF, const int def[3], const int elemBytes,
                               const int profile);
/* Replaces autoFieldPtr, src/binMapUtil.c:188-205. */
void **autoFieldPtr(ConstFieldPtr ptr[BMCU_FIELD_MAX], const Context *pC);

/* ---------------------------------------------------------------------------
 * Packed binary map on the GPU  (src/binMapCUDA.h:15-24)
 * ------------------------------------------------------------------------- */
/* Thresholds the field named by pF (device pointers) into the PBM at pW
 * (device; NULL selects a library-owned buffer, resized as needed), fills
 * *pO, blocks until done, stores the kernel time in pK->dtms[0] when pK is
 * given. Returns pW, or NULL on invalid input, allocation or launch failure.
 * Replaces binMapCUDA, src/binMapCUDA.cu:145-264. Tail bits x >= X of each
 * row's last word are written as zero. A partial organisation (profID &
 * 0x0F != 0; setBMO: min(2, Y) rows of min(2, Z) planes) maps exactly those
 * rows and planes of the field at the table's pointer, so stepping the
 * pointer walks the field in row-pair windows. */
BMPackWord *binMapCUDA(KernInfo *pK, BMPackWord *pW, BMOrg *pO, const BMFieldInfo *pF, const BinMapF64 *pM);
/* Frees the library-owned PBM buffer.  Replaces src/binMapCUDA.cu:267-270. */
void binMapCUDACleanup(void);

/* ---------------------------------------------------------------------------
 * Pattern histogram on the GPU  (src/mkfCUDA.h:18-31)
 * ------------------------------------------------------------------------- */
#define MKF_BINS (1 << 8)
#define MKFCU_PROFILE_FAST (3)
#define MKFCU_PROFILE_FLEX (2)

/* ADDS the BPFD of the PBM pW (device) to pBPFD (device, 256 x uint64), so
 * the caller zeroes it first. All profiles give identical counts. Blocks;
 * pK->dtms[1] receives the time. Returns pBPFD (NULL in, NULL out).
 * Replaces mkfCUDAGetBPFD, src/mkfCUDA.cu:13-61. */
size_t *mkfCUDAGetBPFD(KernInfo *pK, size_t *pBPFD, const BMOrg *pO, const BMPackWord *pW, const int profile);
/* Host-result form: zeroes a library device histogram, runs mkfCUDAGetBPFD
 * and copies the 256 counts to pBPFDH (host; NULL selects a library-owned
 * pinned buffer). Returns the host pointer or NULL.
 * Replaces mkfCUDAGetBPFDH, src/mkfCUDA.cu:73-91. */
size_t *mkfCUDAGetBPFDH(KernInfo *pK, size_t *pBPFDH, const BMOrg *pO, const BMPackWord *pW, const int profile);
/* Deprecated all-in-one path over a Context.  Replaces src/mkfCUDA.cu:102-167. */
int mkfCUDAGetBPFDautoCtx(Context *pC, const FieldDef def[3], const BinMapF64 *pMC, const int profile);
/* Frees library buffers; lvl > 0 also resets the device.
 * Replaces mkfCUDACleanup, src/mkfCUDA.cu:94-99. */
void mkfCUDACleanup(int lvl);

/* ---------------------------------------------------------------------------
 * Host-field entry points  (src/mkfACC.h:27-46)
 * ------------------------------------------------------------------------- */
/* Host double field in, host PBM (pW, planeWS*Z words) and host BPFD out;
 * computed on the GPU. Replaces mkfAccGetBPFDSimple, src/mkfACC.c:206-239. */
Bool32 mkfAccGetBPFDSimple(size_t rBPFD[MKF_BINS], BMPackWord *pW, const double *pF, const FieldDef def[3],
                           const BinMapF64 *pM);
/* Host field of encoding enc (ENC_F32 or ENC_F64) in, host BPFD out.
 * Replaces mkfAccCUDAGetBPFD, src/mkfACC.c:248-318. */
Bool32 mkfAccCUDAGetBPFD(size_t rBPFD[MKF_BINS], const void *pF, const FieldDef def[3], const NumEnc enc,
                         const BinMapF64 *pM);

/* ---------------------------------------------------------------------------
 * Measures  (src/mkfUtil.h:35-42)
 * ------------------------------------------------------------------------- */
#define MKF_REFML "Kv Mv Sv Vv"
#define CELL8_SYMM_GROUPS 22
#define CELL8_PATTERNS 256

/* The reference's book measures live in a file absent from its tree
 * (refMeasures.c); as the reference compiles (-DNO_REF_MEASURES) all four
 * are -1. Replaces mkfRefMeasureBPFD, src/mkfUtil.c:83-90. */
void mkfRefMeasureBPFD(float m[4], const size_t aBPFD[MKF_BINS], const float s);
/* m = {Kv, -1, -1, Vv}, symCh = "Kv       Vv" (12 bytes), returns 2.
 * Replaces mkfSelectMeasuresBPFD, src/mkfUtil.c:94-123. */
int mkfSelectMeasuresBPFD(float m[4], char symCh[], const size_t aBPFD[MKF_BINS], const float s, const int profID);
/* Fraction of cells whose lowest corner is set.  src/mkfUtil.c:125-135. */
float volFrac(const size_t aBPFD[MKF_BINS]);

#ifdef __cplusplus
} /* extern "C" */
#endif

#include "mkf_ext.h"

#endif /* MKF_AMD_H */
