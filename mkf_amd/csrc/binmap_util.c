/* binmap_util.c - host helpers of the C ABI: threshold descriptor parsing,
 * PBM organisation, field-table and stride helpers, element encodings.
 *
 * Each function keeps the name, argument meaning and return convention of
 * the reference helper it replaces (cited per function), because C callers
 * of the reference use them to build the arguments of the GPU entry points.
 */
#include <stdio.h>
#include <string.h>

#include "mkf.h"

/* Storage bytes of n elements of encoding e; bit encodings are packed.
 * Behaviour of encSizeN, src/encoding.c:9-25 (unknown type classes are
 * treated as bit encodings after a warning). */
size_t encSizeN(int *pElemBits, const size_t n, const NumEnc e) {
  int b = e & ENC_MASK_NUM;
  size_t bytes;
  switch (e & ENC_MASK_TYP) {
    case ENC_TYPE_FPB:
    case ENC_TYPE_INT:
      bytes = (size_t)b * n;
      b <<= 3;
      break;
    default:
      fprintf(stderr, "[mkf] encSizeN(.. 0x%02X) - assuming bit encoding\n", e);
      /* fall through */
    case ENC_TYPE_BTS:
      bytes = ((size_t)b * n) >> 3;
      break;
  }
  if (pElemBits) *pElemBits = b;
  return bytes;
}

/* Alignment in bytes of one element.  src/encoding.c:27-40. */
int encAlignment(const NumEnc e) {
  int b = e & ENC_MASK_NUM;
  switch (e & ENC_MASK_TYP) {
    case ENC_TYPE_FPB:
    case ENC_TYPE_INT:
      return b;
    default:
      fprintf(stderr, "[mkf] encAlignment(.. 0x%02X) - assuming bit encoding\n", e);
      /* fall through */
    case ENC_TYPE_BTS:
      return (b + 7) >> 3;
  }
}

/* Relation string + float threshold -> {t, m}. Characters are consumed up to
 * the first one that is not a relation symbol; '!' is accepted only as the
 * first character and inverts the finished mask.  src/binMapUtil.c:37-58.
 * (The reference dereferences both pointers unchecked: here a NULL map
 * returns NULL and a NULL string reads as "", the empty relation.) */
BinMapF64 *setBinMapF64(BinMapF64 *pM, const char relopChars[], const float t) {
  int inv = 0;
  if (!pM) return NULL;
  if (!relopChars) relopChars = "";
  pM->t[0] = t;
  pM->m = BMC_NV;
  for (int i = 0;; ++i) {
    const char c = relopChars[i];
    if (c == '<') {
      pM->m |= BMC_LT;
    } else if (c == '=') {
      pM->m |= BMC_EQ;
    } else if (c == '>') {
      pM->m |= BMC_GT;
    } else if (c == '!' && i == 0) {
      inv = 1;
    } else {
      break;
    }
  }
  if (inv) pM->m ^= BMC_AL;
  return pM;
}

/* Fills the PBM organisation for dims def[] and returns the buffer size in
 * 32-bit words. A zero low nibble of profID means a full-size buffer;
 * otherwise the buffer holds two rows by two planes (the reference's partial
 * "sliding" buffer).  src/binMapUtil.c:60-87. */
size_t setBMO(BMOrg *pO, const FieldDef def[3], const char profID) {
  if (!def) {  /* the reference reads def unchecked; here: an empty organisation */
    if (pO) memset(pO, 0, sizeof(*pO));
    return 0;
  }
  unsigned maxRow = def[1] < 2 ? (unsigned)def[1] : 2u;
  unsigned maxPlane = def[2] < 2 ? (unsigned)def[2] : 2u;
  const BMStride rowStride = (BMStride)(((unsigned)def[0] + 31u) >> 5);
  if (0 == (profID & 0x0F)) {
    maxRow = (unsigned)def[1];
    maxPlane = (unsigned)def[2];
  }
  const BMStride planeStride = rowStride * maxRow;
  if (pO) {
    pO->rowElem = (uint32_t)def[0];
    pO->rowPairs = maxRow - 1;
    pO->planePairs = maxPlane - 1;
    pO->rowWS = rowStride;
    pO->planeWS = planeStride;
  }
  return (size_t)planeStride * maxPlane;
}

/* Copies the dims > 1 to the front of d[], pads the rest with 1 (or 0 when no
 * dim exceeds 1) and returns how many were kept.  src/binMapUtil.c:89-103. */
int copyNonZeroDef(FieldDef d[], const FieldDef *pD, const int nD) {
  int rD = 0;
  if (!d) return 0;  /* (NULL output, read unchecked by the reference) */
  if (pD) {
    for (int i = 0; i < nD; i++) {
      d[rD] = pD[i];
      rD += (d[rD] > 1);
    }
  }
  const FieldDef pad = (rD > 0);
  for (int i = rD; i < nD; i++) d[i] = pad;
  return rD;
}

/* Copies the non-NULL table entries enabled by mask, in table order, at most
 * max of them (src/binMapUtil.c:107-121 selects the same entries). An entry
 * is read only when its mask bit is set: the reference's scan also reads the
 * entry one past the highest set bit, which can lie outside the caller's
 * table. */
int copyValidPtrByMask(ConstFieldPtr r[], const int max, const ConstFieldPtr a[], const uint32_t mask) {
  int n = 0;
  if (!r || !a) return 0;  /* (NULL tables, read unchecked by the reference) */
  for (int i = 0; i < 32 && n < max && (mask >> i) != 0; i++)
    if (((mask >> i) & 1u) && a[i].p != NULL) r[n++] = a[i];
  return n;
}

/* Counts the non-NULL table entries enabled by mask, reading only those
 * entries.  src/binMapUtil.c:123-133. */
int countValidPtrByMask(const ConstFieldPtr a[], uint32_t mask) {
  int n = 0;
  if (!a) return 0;
  for (int i = 0; i < 32 && (mask >> i) != 0; i++) n += ((mask >> i) & 1u) && a[i].p != NULL;
  return n;
}

/* n strides from dims, starting after `start` dims, scaled by a base stride.
 * Returns the count written (0 for NULL dims or zero base).
 * src/binMapUtil.c:135-144. Deliberate difference: each further stride is
 * scaled by the next dim; the reference scales by pD[start] every time
 * (its loop index does not advance), which gives {1, X, X*X} instead of
 * {1, X, X*Y} for n = 3 and only agrees for cubic dims. */
int genStride(FieldStride fs[], const int n, const int start, const FieldDef *pD, FieldStride stride) {
  int r = 0;
  if (pD && stride != 0) {
    int i;
    for (i = 0; i < start; i++) stride *= pD[i];
    while (r < n) {
      fs[r++] = stride;
      if (r < n) stride *= pD[i++];
    }
  }
  return r;
}

/* Copies strides pS[start..start+n) when given, else generates planar ones.
 * src/binMapUtil.c:146-155. */
int copyOrGenStride(FieldStride fs[], const int n, const int start, const FieldDef *pD, const FieldStride *pS) {
  if (pS) {
    for (int i = 0; i < n; i++) fs[i] = pS[i + start];
    return n;
  }
  return genStride(fs, n, start, pD, 1);
}

/* Reinterprets a pointer table as a field table, warning on misalignment.
 * src/binMapUtil.c:158-164. */
ConstFieldPtr *asFieldTab(const void **pp, const NumEnc id) {
  ConstFieldPtr *pFP = (ConstFieldPtr *)pp;
  if (!pFP) return NULL;
  const int a = encAlignment(id);
  if (0 != (pFP->w & (size_t)(a - 1))) fprintf(stderr, "[mkf] alignment %d %p\n", a, pFP->p);
  return pFP;
}

/* Fills a planar field description for n fields of elemBytes-sized floats.
 * Returns pI, or NULL for bad counts or element sizes.
 * src/binMapUtil.c:166-186. */
const BMFieldInfo *setupFields(BMFieldInfo *pI, void **ppF, const int n, const int def[3], const int b,
                               const int profile) {
  if (pI && n > 0 && n < 32) {
    if (b == (int)sizeof(float)) {
      pI->elemID = ENC_F32;
    } else if (b == (int)sizeof(double)) {
      pI->elemID = ENC_F64;
    } else {
      return NULL;
    }
    pI->fieldTableMask = (1u << n) - 1u;
    pI->oprID = 0;
    pI->profID = (uint8_t)profile;
    pI->flags = 0;
    pI->pD = def;
    pI->pS = NULL;
    pI->pFieldDevPtrTable = asFieldTab((const void **)ppF, pI->elemID);
    return pI;
  }
  return NULL;
}

/* Field pointer table for a Context: consecutive fields of one allocation
 * when there are several, else the address of pDF.  src/binMapUtil.c:188-205. */
void **autoFieldPtr(ConstFieldPtr ptr[BMCU_FIELD_MAX], const Context *pC) {
  if (!ptr || !pC) return NULL;  /* (read unchecked by the reference) */
  if (pC->nField > 1) {
    const FieldStride stride = (FieldStride)pC->nElem * pC->bytesElem;
    size_t w = (size_t)pC->pDF;
    const int max = pC->nField < BMCU_FIELD_MAX ? pC->nField : BMCU_FIELD_MAX;
    for (int i = 0; i < max; i++) {
      ptr[i].w = w;
      w += (size_t)stride;
    }
    return (void **)ptr;
  }
  return (void **)&(pC->pDF);
}
