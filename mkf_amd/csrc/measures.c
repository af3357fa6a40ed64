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
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
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

/* Kv terms (symmetry groups) the reference's 32-bit arithmetic does not
 * give exactly: the group count does not fit 32 bits, or its product with the
 * weight wraps. */
int mkfKvWrapCount(const size_t aBPFD[MKF_BINS]) {
  if (!aBPFD) return -1;
  size_t g[CELL8_SYMM_GROUPS] = {0};
  for (int i = 0; i < MKF_BINS; i++) g[kPatternGroup[i]] += aBPFD[i];
  int n = 0;
  for (int i = 0; i < CELL8_SYMM_GROUPS; i++) {
    const int32_t wrapped = (int32_t)((uint32_t)(int32_t)kGroupWeightK[i] * (uint32_t)g[i]);
    const int exactFits = g[i] <= (size_t)INT32_MAX &&
                          (int64_t)kGroupWeightK[i] * (int64_t)g[i] == (int64_t)wrapped;
    if (kGroupWeightK[i] != 0 && !exactFits) n++;
  }
  return n;
}

/* src/mkfUtil.c:125-135: odd bins (lowest corner set) over all bins. */
float volFrac(const size_t aBPFD[MKF_BINS]) {
  size_t s0 = 0, s1 = 0;
  if (!aBPFD) return -1;  /* no histogram (the reference reads it unchecked) */
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
  if (!m) return;
  m[0] = m[1] = m[2] = m[3] = -1;
}

/* src/mkfUtil.c:94-123 as compiled with NO_REF_MEASURES (profID unused). */
int mkfSelectMeasuresBPFD(float m[4], char symCh[], const size_t aBPFD[MKF_BINS], const float s, const int profID) {
  (void)profID;
  if (!m || !aBPFD) return 0;  /* no measures (the reference reads both unchecked) */
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
  if (symCh) memcpy(symCh, "Kv       Vv", 12);
  if (getenv("MKF_VERBOSE")) {
    const int w = mkfKvWrapCount(aBPFD);
    if (w > 0) fprintf(stderr, "[mkf] mkfSelectMeasuresBPFD: %d Kv terms wrap in 32-bit arithmetic, as in the reference\n", w);
  }
  return 2;
}

/* ---------------------------------------------------------------------------
 * Crofton-formula estimates of all four measures (no reference counterpart:
 * the reference's book measures are absent from its tree). Method after
 * Ohser & Muecklich, "Statistical Analysis of Microstructures in Materials
 * Science" (2000), ch. 4-5: intrinsic volumes of a digitised set from its
 * 2x2x2 configuration histogram, with the 13 lattice directions of the cubic
 * lattice weighted by the spherical Voronoi cells of the 26 unit directions.
 *
 * For each pattern p the library derives, once:
 *  - K: 4 pi times the pattern's share of the Euler characteristic of the
 *    6-adjacency foreground (vertex 1/8, axis edge 1/4, face 1/2, cube 1);
 *  - S: 2 sum_v c_v P_L(v), P_L the boundary crossings per unit length along
 *    direction v. A crossing is a pair of cell vertices along v that differ;
 *    a pair along an axis lies in 4 cells, along a face diagonal in 2, along
 *    a space diagonal in 1, so it counts 1/4, 1/2 or 1 of a crossing here;
 *  - M: 2 pi sum_v c_v d_v chi(v), chi(v) the Euler number of the sections by
 *    the lattice planes with normal v (spacing d_v = 1, 1/sqrt2, 1/sqrt3).
 *    Axis planes cut a cell in two square faces (each shared by two cells),
 *    face-diagonal planes in one 1 x sqrt2 rectangle, space-diagonal planes in
 *    two triangles. A square or rectangle adds (Q1 - Q3)/4 (the mean of its
 *    4- and 8-adjacency Euler numbers, Q1/Q3 = one/three corners set); a
 *    triangle adds 1/6 for one vertex set and -1/6 for two (6-adjacency).
 * c_v is the pair's share of the sphere: twice the area fraction of one
 * direction's spherical Voronoi cell among the 26 directions.
 * ------------------------------------------------------------------------- */
#define CROF_AXIS 0.04577789120476086  /* Voronoi area fraction, (1,0,0) */
#define CROF_FACE 0.03698062787607520  /* (1,1,0) */
#define CROF_SPACE 0.03519563978231666 /* (1,1,1) */

static double gCrofK[CELL8_PATTERNS], gCrofM[CELL8_PATTERNS], gCrofS[CELL8_PATTERNS];
static pthread_once_t gCrofOnce = PTHREAD_ONCE_INIT;

static int vbit(int p, int x, int y, int z) { return (p >> (x + 2 * y + 4 * z)) & 1; }

/* Mean of the 4- and 8-adjacency Euler numbers of one 2x2 plane cell. */
static double quadChi(int n) { return n == 1 ? 0.25 : (n == 3 ? -0.25 : 0.0); }
static double triChi(int n) { return n == 1 ? 1.0 / 6 : (n == 2 ? -1.0 / 6 : 0.0); }

static void crofInit(void) {
  const double cA = 2 * CROF_AXIS, cF = 2 * CROF_FACE, cS = 2 * CROF_SPACE;
  const double r2 = sqrt(2.0), r3 = sqrt(3.0);
  for (int p = 0; p < CELL8_PATTERNS; p++) {
    int nv = 0, ne = 0, nf = 0;
    int tA = 0, tF = 0, tS = 0; /* differing vertex pairs per direction class */
    for (int i = 0; i < 8; i++) {
      nv += (p >> i) & 1;
      for (int j = i + 1; j < 8; j++) {
        const int d = __builtin_popcount(i ^ j); /* 1 axis, 2 face diagonal, 3 space diagonal */
        const int diff = ((p >> i) ^ (p >> j)) & 1;
        if (d == 1) {
          tA += diff;
          ne += ((p >> i) & (p >> j)) & 1;
        } else if (d == 2) {
          tF += diff;
        } else {
          tS += diff;
        }
      }
    }
    /* faces: fix one axis at 0 or 1 */
    double mAxis = 0;
    for (int ax = 0; ax < 3; ax++)
      for (int c = 0; c < 2; c++) {
        int n = 0;
        for (int v = 0; v < 8; v++)
          if (((v >> ax) & 1) == c) n += (p >> v) & 1;
        nf += (n == 4);
        mAxis += quadChi(n) / 2;
      }
    /* face-diagonal planes: for axes a < b, a + b == 1 and a == b */
    double mFace = 0;
    for (int a = 0; a < 3; a++)
      for (int b = a + 1; b < 3; b++) {
        int nSum = 0, nEq = 0;
        for (int v = 0; v < 8; v++) {
          const int va = (v >> a) & 1, vb = (v >> b) & 1;
          if (va + vb == 1) nSum += (p >> v) & 1;
          if (va == vb) nEq += (p >> v) & 1;
        }
        mFace += quadChi(nSum) + quadChi(nEq);
      }
    /* space-diagonal planes: normals (1,1,1), (-1,1,1), (1,-1,1), (1,1,-1);
       the levels holding three vertices are the two triangles */
    double mSpace = 0;
    for (int flip = -1; flip < 3; flip++) {
      int cnt[4] = {0, 0, 0, 0}, n[4] = {0, 0, 0, 0};
      for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++)
          for (int z = 0; z < 2; z++) {
            int u[3] = {x, y, z};
            if (flip >= 0) u[flip] = 1 - u[flip]; /* reflection maps the normal to (1,1,1) */
            const int lvl = u[0] + u[1] + u[2];
            cnt[lvl]++;
            n[lvl] += vbit(p, x, y, z);
          }
      for (int l = 0; l < 4; l++)
        if (cnt[l] == 3) mSpace += triChi(n[l]);
    }
    const int cube = (p == 0xFF);
    gCrofK[p] = 4 * M_PI * (nv / 8.0 - ne / 4.0 + nf / 2.0 - cube);
    gCrofS[p] = 2 * (cA * tA / 4.0 + cF * tF / (2.0 * r2) + cS * tS / r3);
    gCrofM[p] = 2 * M_PI * (cA * mAxis + cF * mFace / r2 + cS * mSpace / r3);
  }
}

int mkfCroftonMeasuresBPFD(double m[4], const size_t aBPFD[MKF_BINS], double s) {
  if (!m || !aBPFD || !(s >= 0)) return MKF_ERR_ARG;
  pthread_once(&gCrofOnce, crofInit);
  double k = 0, mm = 0, sa = 0, t = 0, v = 0;
  for (int p = 0; p < MKF_BINS; p++) {
    const double h = (double)aBPFD[p];
    k += h * gCrofK[p];
    mm += h * gCrofM[p];
    sa += h * gCrofS[p];
    t += h;
    if (p & 1) v += h;
  }
  if (s > 0) {
    if (t == 0) return MKF_ERR_ARG;
    const double vol = t * s * s * s;
    m[0] = k / vol;
    m[1] = mm * s / vol;
    m[2] = sa * s * s / vol;
    m[3] = v / t;
  } else {
    m[0] = k;
    m[1] = mm;
    m[2] = sa;
    m[3] = v;
  }
  return 4;
}
