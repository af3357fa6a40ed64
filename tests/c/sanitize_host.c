/* sanitize_host.c - drives the library's host C (binmap_util.c, measures.c)
 * and the C oracle (oracle/mkf_oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer; tests/test_sanitize.py builds it with gcc from
 * the sources and runs it. Every host helper is called on ordinary and edge
 * inputs, and the oracle's histogram paths are cross-checked on small random
 * volumes of every relation, including rows narrower than one word. Exits
 * nonzero (or the sanitizer aborts) on any error. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mkf.h"

/* the oracle's exports (oracle/mkf_oracle.c) */
void orc_pbm_f64(uint32_t *pW, const double *pF, int X, int Y, int Z, double t, uint32_t m);
void orc_pbm_f32(uint32_t *pW, const float *pF, int X, int Y, int Z, double t, uint32_t m);
int orc_bpfd_chunk(uint64_t h[256], const uint32_t *pW, int X, int Y, int Z, int rowWS, long planeWS);
void orc_bpfd_direct(uint64_t h[256], const uint32_t *pW, int X, int Y, int Z, int rowWS, long planeWS);
void orc_bpfd_gpu_walk(uint64_t h[256], const uint32_t *pW, int X, int Y, int Z, int rowWS, long planeWS);
int orc_bpfd_field_f64(uint64_t h[256], uint32_t *pW, const double *pF, int X, int Y, int Z, double t, uint32_t m);
int orc_bpfd_field_f32(uint64_t h[256], uint32_t *pW, const float *pF, int X, int Y, int Z, double t, uint32_t m);
void orc_set_threads(int n);
int orc_measures(float m[4], const uint64_t a[256], float s, int *pOverflow);
int orc_euler_weight(unsigned code);
int orc_orbits(uint8_t label[256]);

static int gFail = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      gFail = 1;                                                   \
    }                                                              \
  } while (0)

static uint64_t gRng = 0x9E3779B97F4A7C15ull;
static uint64_t rnd(void) {
  gRng ^= gRng << 13;
  gRng ^= gRng >> 7;
  gRng ^= gRng << 17;
  return gRng;
}

static void helpers(void) {
  static const char *rel[] = {"<", "=", ">", "<=", ">=", "!=", "!<", "!>", "!", "", "=!", ">=x<", "<<<", "!>="};
  for (size_t i = 0; i < sizeof(rel) / sizeof(rel[0]); i++) {
    BinMapF64 bm;
    CHECK(setBinMapF64(&bm, rel[i], 0.25f) == &bm);
    CHECK(bm.m <= 7 && bm.t[0] == 0.25);
  }
  {  // NULL arguments are refused instead of dereferenced
    BinMapF64 bm;
    CHECK(setBinMapF64(NULL, ">=", 0.5f) == NULL);
    CHECK(setBinMapF64(&bm, NULL, 0.5f) == &bm && bm.m == 0);
    BMOrg o;
    memset(&o, 0xFF, sizeof(o));
    CHECK(setBMO(&o, NULL, 0) == 0 && o.rowWS == 0 && o.planeWS == 0);
    float m[4];
    char sym[16];
    size_t h[256] = {0};
    CHECK(mkfSelectMeasuresBPFD(NULL, sym, h, 1.0f, 2) == 0);
    CHECK(mkfSelectMeasuresBPFD(m, sym, NULL, 1.0f, 2) == 0);
    h[0] = 8;
    CHECK(mkfSelectMeasuresBPFD(m, NULL, h, 1.0f, 2) == 2);
    CHECK(volFrac(NULL) == -1.0f);
    mkfRefMeasureBPFD(NULL, h, 1.0f);
    CHECK(copyNonZeroDef(NULL, NULL, 3) == 0);
    CHECK(copyValidPtrByMask(NULL, 4, NULL, 0xFu) == 0);
    CHECK(countValidPtrByMask(NULL, 0xFu) == 0);
    CHECK(asFieldTab(NULL, ENC_F32) == NULL);
    CHECK(autoFieldPtr(NULL, NULL) == NULL);
  }
  const FieldDef defs[][3] = {{1, 1, 1}, {33, 7, 5}, {64, 1, 9}, {1, 256, 256}, {1025, 3, 2}, {2, 2, 2}};
  for (size_t i = 0; i < sizeof(defs) / sizeof(defs[0]); i++) {
    BMOrg o;
    const size_t n = setBMO(&o, defs[i], 0);
    CHECK(n == (size_t)o.planeWS * (size_t)defs[i][2] || defs[i][0] == 1);
    BMOrg p;
    (void)setBMO(&p, defs[i], 1);  // partial (2-plane) organisation
    FieldDef d[3];
    const int nd = copyNonZeroDef(d, defs[i], 3);
    CHECK(nd >= 0 && nd <= 3);
    FieldStride fs[3];
    CHECK(genStride(fs, 3, 0, defs[i], 1) >= 0);
    CHECK(copyOrGenStride(fs, 3, 0, defs[i], NULL) >= 0);
    const FieldStride given[3] = {1, 40, 400};
    CHECK(copyOrGenStride(fs, 3, 0, defs[i], given) >= 0);
  }
  float buf[4] = {0};
  ConstFieldPtr tab[8], out[BMCU_FIELD_MAX];
  for (int i = 0; i < 8; i++) tab[i].p = (i % 3 == 1) ? NULL : buf + (i & 3);
  for (uint32_t mask = 0; mask < 256; mask += 7) {
    const int n = copyValidPtrByMask(out, BMCU_FIELD_MAX, tab, mask);
    CHECK(n >= 0 && n <= BMCU_FIELD_MAX);
    CHECK(countValidPtrByMask(tab, mask) >= 0);
  }
  const void *pp[2] = {buf, buf + 1};
  CHECK(asFieldTab(pp, ENC_F32) != NULL);
  void *fields[2] = {buf, buf + 2};
  BMFieldInfo fi;
  const int d3[3] = {4, 1, 1};
  CHECK(setupFields(&fi, fields, 2, d3, 4, 0) == &fi);
  Context ctx;
  memset(&ctx, 0, sizeof(ctx));
  ctx.nField = 1;
  ctx.pDF = buf;
  ctx.bytesElem = 4;
  ctx.nElem = 4;
  ConstFieldPtr cp[BMCU_FIELD_MAX];
  (void)autoFieldPtr(cp, &ctx);
  int bits = 0;
  CHECK(encSizeN(&bits, 210, ENC_F32) == 840u && bits == 32);
  CHECK(encSizeN(&bits, 64, ENC_U1) == 8u && bits == 1);
  CHECK(encAlignment(ENC_F64) > 0);
}

static void measures(void) {
  size_t h[256];
  uint64_t hu[256];
  for (int rep = 0; rep < 200; rep++) {
    const uint64_t scale = rep < 100 ? 1000u : (rep < 150 ? (1ull << 33) : 0ull);
    for (int i = 0; i < 256; i++) hu[i] = h[i] = scale ? rnd() % scale : 0;
    if (rep == 199) hu[255] = h[255] = 1;
    float m[4], mo[4], s = (rep & 1) ? 0.0f : 3.0f / (float)(64 + rep);
    char sym[16];
    int ovf = 0;
    CHECK(mkfSelectMeasuresBPFD(m, sym, h, s, 2) == 2);
    CHECK(orc_measures(mo, hu, s, &ovf) == 2);
    CHECK(memcmp(m, mo, sizeof(m)) == 0);
    CHECK((mkfKvWrapCount(h) > 0) == (ovf != 0));
    (void)volFrac(h);
    mkfRefMeasureBPFD(m, h, s);
    double c[4];
    const int rc = mkfCroftonMeasuresBPFD(c, h, (rep & 1) ? 0.0 : 0.5);
    CHECK(rc == 4 || rc == MKF_ERR_ARG);
  }
  const uint8_t *g = mkfPatternGroupTable();
  const int8_t *w = mkfGroupWeightKTable();
  uint8_t lab[256];
  CHECK(orc_orbits(lab) == 22);
  for (int i = 0; i < 256; i++) CHECK(g[i] < 22 && w[g[i]] >= -24 && w[g[i]] <= 24 && orc_euler_weight(i) <= 24);
}

static void oracle(void) {
  const int shapes[][3] = {{2, 2, 2}, {3, 5, 4}, {31, 4, 6}, {32, 3, 3}, {33, 6, 5}, {65, 4, 7}, {70, 9, 3}};
  for (int th = 1; th <= 3; th += 2) {
    orc_set_threads(th);
    for (size_t si = 0; si < sizeof(shapes) / sizeof(shapes[0]); si++) {
      const int X = shapes[si][0], Y = shapes[si][1], Z = shapes[si][2];
      const int rowWS = (X + 31) >> 5;
      const size_t n = (size_t)X * Y * Z, nw = (size_t)rowWS * Y * Z;
      double *f64 = malloc(n * sizeof(double));
      float *f32 = malloc(n * sizeof(float));
      uint32_t *w = malloc(nw * sizeof(uint32_t)), *w2 = malloc(nw * sizeof(uint32_t));
      for (size_t i = 0; i < n; i++) {
        const uint64_t r = rnd();
        f64[i] = (double)(r % 1000) / 1000.0 - 0.5;
        if (r % 97 == 0) f64[i] = NAN;
        if (r % 89 == 0) f64[i] = -0.0;
        f32[i] = (float)f64[i];
      }
      for (uint32_t m = 0; m < 8; m++) {
        uint64_t hc[256] = {0}, hd[256] = {0}, hg[256] = {0}, hf[256];
        orc_pbm_f64(w, f64, X, Y, Z, 0.0, m);
        orc_pbm_f32(w2, f32, X, Y, Z, 0.0, m);
        CHECK(memcmp(w, w2, nw * sizeof(uint32_t)) == 0);
        orc_bpfd_direct(hd, w, X, Y, Z, rowWS, (long)rowWS * Y);
        orc_bpfd_gpu_walk(hg, w, X, Y, Z, rowWS, (long)rowWS * Y);
        uint64_t tot = 0;
        for (int i = 0; i < 256; i++) tot += hg[i];
        if (X >= 32) {
          CHECK(orc_bpfd_chunk(hc, w, X, Y, Z, rowWS, (long)rowWS * Y) == 0);
          CHECK(memcmp(hc, hd, sizeof(hc)) == 0);
          CHECK(memcmp(hg, hd, sizeof(hg)) == 0);
          CHECK(orc_bpfd_field_f64(hf, w2, f64, X, Y, Z, 0.0, m) == 0);
          CHECK(memcmp(hf, hd, sizeof(hf)) == 0);
          CHECK(orc_bpfd_field_f32(hf, w2, f32, X, Y, Z, 0.0, m) == 0);
          CHECK(memcmp(hf, hd, sizeof(hf)) == 0);
        } else {
          CHECK(tot == (uint64_t)(2 * X - 1) * (uint64_t)(Y - 1) * (uint64_t)(Z - 1));
        }
      }
      free(f64);
      free(f32);
      free(w);
      free(w2);
    }
  }
}

int main(void) {
  helpers();
  measures();
  oracle();
  if (gFail) return 1;
  printf("sanitized host code: ok\n");
  return 0;
}
