/* measures.c - Minkowski measures from a 256-bin 2x2x2 pattern histogram.
 *
 * Host code, O(256). Follows the reference as it compiles (its book measures
 * live in refMeasures.c, which is absent from its tree, so the reference is
 * built with NO_REF_MEASURES): m = {Kv, -1, -1, Vv}.
 *
 * Kv (src/mkfUtil.c:64-69, 108-115): bins are summed per cube-symmetry group,
 * each group sum is truncated to int and multiplied by its integer weight in
 * 32-bit int arithmetic (wrapping), the products are summed in int64, scaled
 * by pi/6 in double and stored as float; the float is then multiplied by the
 * float 1/(s^3 * total) evaluated as the reference does. The operation order
 * and precisions are kept exactly, so results are bit-identical.
 *
 * The weight of a pattern is 24 times its share of the Euler characteristic
 * of the foreground (vertices 1/8, edges 1/4, faces 1/2, cube 1), so that
 * sum(w * n) * pi/6 = 4 pi chi for objects away from the border.
 */
#include <math.h>
#include <string.h>

#include "mkf.h"

/* Euler weights per symmetry group (src/mkfUtil.c:33-38). */
static const int8_t kGroupWeightK[CELL8_SYMM_GROUPS] = {
    0, 3, 0, 6, 6, -3, 3, 9, 0, -6, -6, 0, 12, 0, -3, -8, -3, -12, -6, 0, 3, 0,
};

/* Symmetry group of each 8-bit pattern (src/mkfUtil.c:40-58); the groups are
 * the 22 orbits of the 48-element cube symmetry group. */
static const uint8_t kPatternGroup[CELL8_PATTERNS] = {
    0,  1,  1,  2,  1,  2,  3,  5,  1,  3,  2,  5,  2,  5,  5,  8,  1,  2,  3,  5,  3,  5,  7,  9,  4,  6,
    6,  10, 6,  10, 13, 16, 1,  3,  2,  5,  4,  6,  6,  10, 3,  7,  5,  9,  6,  13, 10, 16, 2,  5,  5,  8,
    6,  10, 13, 16, 6,  13, 10, 16, 11, 15, 15, 19, 1,  3,  4,  6,  2,  5,  6,  10, 3,  7,  6,  13, 5,  9,
    10, 16, 2,  5,  6,  10, 5,  8,  13, 16, 6,  13, 11, 15, 10, 16, 15, 19, 3,  7,  6,  13, 6,  13, 11, 15,
    7,  12, 13, 14, 13, 14, 15, 18, 5,  9,  10, 16, 10, 16, 15, 19, 13, 14, 15, 18, 15, 18, 17, 20, 1,  4,
    3,  6,  3,  6,  7,  13, 2,  6,  5,  10, 5,  10, 9,  16, 3,  6,  7,  13, 7,  13, 12, 14, 6,  11, 13, 15,
    13, 15, 14, 18, 2,  6,  5,  10, 6,  11, 13, 15, 5,  13, 8,  16, 10, 15, 16, 19, 5,  10, 9,  16, 13, 15,
    14, 18, 10, 15, 16, 19, 15, 17, 18, 20, 2,  6,  6,  11, 5,  10, 13, 15, 5,  13, 10, 15, 8,  16, 16, 19,
    5,  10, 13, 15, 9,  16, 14, 18, 10, 15, 15, 17, 16, 19, 18, 20, 5,  13, 10, 15, 10, 15, 15, 17, 9,  14,
    16, 18, 16, 18, 19, 20, 8,  16, 16, 19, 16, 19, 18, 20, 16, 18, 19, 20, 19, 20, 20, 21,
};

/* Read-only views for tests that check the tables against their derivation. */
const uint8_t *mkfPatternGroupTable(void) { return kPatternGroup; }
const int8_t *mkfGroupWeightKTable(void) { return kGroupWeightK; }

static size_t sumBins(const size_t *a, int n) {
  size_t s = a[0];
  for (int i = 1; i < n; i++) s += a[i];
  return s;
}

/* Euler sum (src/mkfUtil.c:64-69): per-group products in 32-bit int. */
static double eulerSumK(const size_t g[CELL8_SYMM_GROUPS]) {
  int64_t k = 0;
  for (int i = 0; i < CELL8_SYMM_GROUPS; i++) {
    const uint32_t prod = (uint32_t)(int32_t)kGroupWeightK[i] * (uint32_t)g[i];
    k += (int64_t)(int32_t)prod;
  }
  return (double)k * M_PI / 6;
}

/* src/mkfUtil.c:125-135: odd bins (lowest corner set) over all bins. */
float volFrac(const size_t aBPFD[MKF_BINS]) {
  size_t s0 = 0, s1 = 0;
  for (int i = 0; i < MKF_BINS; i += 2) {
    s0 += aBPFD[i];
    s1 += aBPFD[i + 1];
  }
  return (float)s1 / (float)(s0 + s1);
}

/* src/mkfUtil.c:83-90 as compiled with NO_REF_MEASURES. */
void mkfRefMeasureBPFD(float m[4], const size_t aBPFD[MKF_BINS], const float s) {
  (void)aBPFD;
  (void)s;
  m[0] = m[1] = m[2] = m[3] = -1;
}

/* src/mkfUtil.c:94-123 as compiled with NO_REF_MEASURES (profID unused). */
int mkfSelectMeasuresBPFD(float m[4], char symCh[], const size_t aBPFD[MKF_BINS], const float s, const int profID) {
  (void)profID;
  size_t g[CELL8_SYMM_GROUPS] = {0};
  const size_t t = sumBins(aBPFD, MKF_BINS);
  for (int i = 0; i < MKF_BINS; i++) g[kPatternGroup[i]] += aBPFD[i];
  m[0] = (float)eulerSumK(g);
  m[3] = volFrac(aBPFD);
  m[1] = m[2] = -1;
  if (t == sumBins(g, CELL8_SYMM_GROUPS) && s > 0) {
    const float cube = s * s * s;
    const float denom = cube * (float)t;
    const float rkV = (float)(1.0 / (double)denom);
    m[0] *= rkV;
  }
  memcpy(symCh, "Kv       Vv", 12);
  return 2;
}
