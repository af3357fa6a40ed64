/* mkf_oracle.c - CPU ORACLE for the MKF hot path. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only to check results or as the timed CPU baseline; the
 * product (libmkf_amd.so) never calls it.
 *
 * It restates the reference's CPU path in plain C:
 *   orc_relbit          src/binMapACC.c:32-40      (threshold bit, S3)
 *   orc_pbm_f64/_f32    src/binMapACC.c:43-93      (packed binary map)
 *   orc_bpfd_chunk      src/mkfACC.c:28-188        (64-bit chunk walk)
 *   orc_bpfd_field_*    src/mkfACC.c:206-239       (mkfAccGetBPFDSimple)
 *   orc_bpfd_gpu_walk   src/bpfdCUDA.cuh:92-118    (the GPU row walk, any X)
 *   orc_measures        src/mkfUtil.c:64-135       (Kv, Vv)
 * and adds second opinions that do not share code with the restatement:
 *   orc_bpfd_direct     per-cell enumeration of the cell-code definition
 *   orc_euler_weight    textbook local Euler characteristic of a 2x2x2 cell
 *   orc_orbits          orbits of the 48 cube symmetries (group structure)
 *
 * Deliberate differences from the reference, each a case where the
 * reference's result is undefined:
 *   - PBM tail bits (x >= X in a row's last word) are zero; the reference
 *     fills them from the next row and reads past the field on the last row.
 *   - orc_bpfd_chunk requires X >= 32: below that the reference walk counts
 *     2X-1 patterns per row and reads words outside the row.
 *
 * Pinning: see the header of oracle/oracle.py.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_BINS 256
#define ORC_MAX_THREADS 256

/* Worker threads (plain pthreads: the process may already hold another
 * OpenMP runtime, e.g. the one bundled with torch). 1 = serial. */
static int gThreads = 1;

typedef void (*orc_range_fn)(void *ctx, long beg, long end, int tid);

typedef struct {
  orc_range_fn fn;
  void *ctx;
  long beg, end;
  int tid;
} orc_job;

static void *orc_job_main(void *p) {
  orc_job *j = (orc_job *)p;
  j->fn(j->ctx, j->beg, j->end, j->tid);
  return NULL;
}

/* Splits [0, n) into gThreads contiguous ranges and runs fn on each. */
static void orc_parallel(long n, orc_range_fn fn, void *ctx) {
  int nt = gThreads < 1 ? 1 : gThreads;
  if (nt > n) nt = (int)(n > 0 ? n : 1);
  if (nt == 1) {
    fn(ctx, 0, n, 0);
    return;
  }
  pthread_t th[ORC_MAX_THREADS];
  orc_job jobs[ORC_MAX_THREADS];
  for (int t = 0; t < nt; t++) {
    jobs[t].fn = fn;
    jobs[t].ctx = ctx;
    jobs[t].beg = n * t / nt;
    jobs[t].end = n * (t + 1) / nt;
    jobs[t].tid = t;
    if (t > 0) pthread_create(&th[t], NULL, orc_job_main, &jobs[t]);
  }
  orc_job_main(&jobs[0]);
  for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
}

/* ---- threshold ---------------------------------------------------------- */

/* d = 1 - (f < t) + (f > t); bit = (m >> d) & 1  (src/binMapACC.c:32-40) */
static inline uint32_t orc_relbit(double f, double t, uint32_t m) {
  const int d = 1 - (f < t) + (f > t);
  return (m >> d) & 1u;
}

uint32_t orc_relbit_f64(double f, double t, uint32_t m) { return orc_relbit(f, t, m); }

/* ---- packed binary map -------------------------------------------------- */

/* Row-aligned PBM of a double field: word (x>>5) + y*rowWS + z*planeWS, bit
 * x & 31; 32 compares per word (src/binMapACC.c:43-58, 77-93). */
typedef struct {
  uint32_t *pW;
  const void *pF;
  int X, f32;
  double t;
  uint32_t m;
} orc_pbm_ctx;

static void orc_pbm_rows(void *p, long beg, long end, int tid) {
  (void)tid;
  const orc_pbm_ctx *c = (const orc_pbm_ctx *)p;
  const int X = c->X, rowWS = (X + 31) >> 5;
  for (long r = beg; r < end; r++) {
    uint32_t *pw = c->pW + r * (long)rowWS;
    for (int wi = 0; wi < rowWS; wi++) {
      uint32_t w = 0;
      const int n = (X - wi * 32) < 32 ? (X - wi * 32) : 32;
      if (c->f32) {
        const float *pr = (const float *)c->pF + r * (long)X + wi * 32;
        for (int j = 0; j < n; j++) w |= orc_relbit((double)pr[j], c->t, c->m) << j;
      } else {
        const double *pr = (const double *)c->pF + r * (long)X + wi * 32;
        for (int j = 0; j < n; j++) w |= orc_relbit(pr[j], c->t, c->m) << j;
      }
      pw[wi] = w;
    }
  }
}

void orc_pbm_f64(uint32_t *pW, const double *pF, int X, int Y, int Z, double t, uint32_t m) {
  orc_pbm_ctx c = {pW, pF, X, 0, t, m};
  orc_parallel((long)Y * Z, orc_pbm_rows, &c);
}

/* The reference CPU path is double only (src/binMapACC.h:20); a float field
 * is widened element by element and compared with the threshold rounded to
 * float, which is what the reference GPU path does for float fields
 * (src/binMapCUDA.cuh:138-139). */
void orc_pbm_f32(uint32_t *pW, const float *pF, int X, int Y, int Z, double t, uint32_t m) {
  orc_pbm_ctx c = {pW, pF, X, 1, (double)(float)t, m};
  orc_parallel((long)Y * Z, orc_pbm_rows, &c);
}

/* ---- histogram: chunk walk (restatement) -------------------------------- */

/* Four rows (y,z), (y+1,z), (y,z+1), (y+1,z+1) held as 64-bit shift
 * registers; each pattern takes the low two bits of each register
 * (src/mkfACC.c:28-91). */
static inline void orc_add_patterns(uint64_t h[ORC_BINS], uint64_t b[4], int n) {
  for (int i = 0; i < n; i++) {
    const unsigned code = (unsigned)((b[0] & 3u) | ((b[1] & 3u) << 2) | ((b[2] & 3u) << 4) | ((b[3] & 3u) << 6));
    b[0] >>= 1;
    b[1] >>= 1;
    b[2] >>= 1;
    b[3] >>= 1;
    h[code]++;
  }
}

/* One row pair of one plane pair (src/mkfACC.c:121-152): the first 32-bit
 * chunk gives 31 patterns, each later whole chunk 32 (its bits are appended
 * above the one carried-over bit), then the X & 31 residual. */
static void orc_add_row(uint64_t h[ORC_BINS], const uint32_t *r0, const uint32_t *r1, int rowWS, int X) {
  uint64_t b[4];
  b[0] = r0[0];
  b[1] = r0[rowWS];
  b[2] = r1[0];
  b[3] = r1[rowWS];
  orc_add_patterns(h, b, (X - 1) < 31 ? (X - 1) : 31);
  const int m = X >> 5;
  int i = 1;
  for (; i < m; i++) {
    b[0] |= (uint64_t)r0[i] << 1;
    b[1] |= (uint64_t)r0[i + rowWS] << 1;
    b[2] |= (uint64_t)r1[i] << 1;
    b[3] |= (uint64_t)r1[i + rowWS] << 1;
    orc_add_patterns(h, b, 32);
  }
  const int k = X & 31;
  if (k > 0) {
    b[0] |= (uint64_t)r0[i] << 1;
    b[1] |= (uint64_t)r0[i + rowWS] << 1;
    b[2] |= (uint64_t)r1[i] << 1;
    b[3] |= (uint64_t)r1[i + rowWS] << 1;
    orc_add_patterns(h, b, k);
  }
}

/* ADDS the histogram of plane pairs [zBeg, zEnd) (src/mkfACC.c:156-188).
 * Requires X >= 32. Plane pairs are split over threads, each with a private
 * histogram merged by integer addition. Returns 0, or -1 for X < 32. */
typedef struct {
  const uint32_t *pW;
  int X, Y, rowWS, zBeg;
  long planeWS;
  uint64_t (*loc)[ORC_BINS];
} orc_walk_ctx;

static void orc_walk_planes(void *p, long beg, long end, int tid) {
  const orc_walk_ctx *c = (const orc_walk_ctx *)p;
  uint64_t *loc = c->loc[tid];
  for (long zi = beg; zi < end; zi++) {
    const uint32_t *p0 = c->pW + (c->zBeg + zi) * c->planeWS;
    const uint32_t *p1 = p0 + c->planeWS;
    for (int y = 0; y < c->Y - 1; y++)
      orc_add_row(loc, p0 + (long)y * c->rowWS, p1 + (long)y * c->rowWS, c->rowWS, c->X);
  }
}

int orc_bpfd_chunk_range(uint64_t h[ORC_BINS], const uint32_t *pW, int X, int Y, int Z, int rowWS, long planeWS,
                         int zBeg, int zEnd) {
  (void)Z;
  if (X < 32) return -1;
  if (zEnd <= zBeg) return 0;
  uint64_t (*loc)[ORC_BINS] = calloc(ORC_MAX_THREADS, sizeof(*loc));
  if (!loc) return -2;
  orc_walk_ctx c = {pW, X, Y, rowWS, zBeg, planeWS, loc};
  orc_parallel(zEnd - zBeg, orc_walk_planes, &c);
  for (int t = 0; t < ORC_MAX_THREADS; t++)
    for (int i = 0; i < ORC_BINS; i++) h[i] += loc[t][i];
  free(loc);
  return 0;
}

int orc_bpfd_chunk(uint64_t h[ORC_BINS], const uint32_t *pW, int X, int Y, int Z, int rowWS, long planeWS) {
  return orc_bpfd_chunk_range(h, pW, X, Y, Z, rowWS, planeWS, 0, Z - 1);
}

/* ---- histogram: direct per-cell enumeration (second opinion) ----------- */

static inline unsigned orc_bit(const uint32_t *pW, int rowWS, long planeWS, int x, int y, int z) {
  return (pW[(x >> 5) + (long)y * rowWS + (long)z * planeWS] >> (x & 31)) & 1u;
}

/* ADDS the histogram by evaluating the cell-code definition for every cell:
 * bit i of the code of cell (x,y,z) is P(x+(i&1), y+((i>>1)&1), z+(i>>2)).
 * Valid for every X >= 2. */
void orc_bpfd_direct(uint64_t h[ORC_BINS], const uint32_t *pW, int X, int Y, int Z, int rowWS, long planeWS) {
  for (int z = 0; z + 1 < Z; z++)
    for (int y = 0; y + 1 < Y; y++)
      for (int x = 0; x + 1 < X; x++) {
        unsigned code = 0;
        for (int i = 0; i < 8; i++) code |= orc_bit(pW, rowWS, planeWS, x + (i & 1), y + ((i >> 1) & 1), z + (i >> 2)) << i;
        h[code]++;
      }
}

/* ---- histogram: the reference GPU row walk ------------------------------ */

/* addRowBPFD of the GPU path (src/bpfdCUDA.cuh:41-118): four 64-bit
 * registers hold rows (y,z), (y+1,z), (y,z+1), (y+1,z+1) of a row pair; each
 * pattern takes bits 0..1 of every register (in that order, 2 bits each) and
 * then shifts all four right by one. The first word gives min(32, n) - 1
 * patterns; each later whole word is ORed in shifted up by one and gives 32;
 * a residual of n & 31 bits ORs in word n >> 5 the same way and gives that
 * many. For n < 32 the residual step reads word 0 again. ADDS to h. */
static void orc_gpu_emit(uint64_t h[ORC_BINS], uint64_t u[4], int cnt) {
  for (int j = 0; j < cnt; j++) {
    h[(u[0] & 3u) | ((u[1] & 3u) << 2) | ((u[2] & 3u) << 4) | ((u[3] & 3u) << 6)]++;
    for (int q = 0; q < 4; q++) u[q] >>= 1;
  }
}

static void orc_gpu_row(uint64_t h[ORC_BINS], const uint32_t *r00, const uint32_t *r01, const uint32_t *r10,
                        const uint32_t *r11, int n) {
  const uint32_t *r[4] = {r00, r01, r10, r11};
  uint64_t u[4] = {r00[0], r01[0], r10[0], r11[0]};
  const int m = n >> 5, k = n & 31;
  orc_gpu_emit(h, u, (n < 32 ? n : 32) - 1);
  for (int i = 1; i < m; i++) {
    for (int q = 0; q < 4; q++) u[q] |= ((uint64_t)r[q][i]) << 1;
    orc_gpu_emit(h, u, 32);
  }
  if (k > 0) {
    for (int q = 0; q < 4; q++) u[q] |= ((uint64_t)r[q][m]) << 1;
    orc_gpu_emit(h, u, k);
  }
}

/* The reference GPU walk (src/bpfdCUDA.cuh:158-206 over src/bpfdCUDA.cuh:92-118)
 * over every row pair and plane pair of the map: ADDS to h. For X >= 32 it
 * equals the per-cell enumeration; for X < 32 it emits 2X - 1 patterns per
 * row pair. */
void orc_bpfd_gpu_walk(uint64_t h[ORC_BINS], const uint32_t *pW, int X, int Y, int Z, int rowWS, long planeWS) {
  for (int z = 0; z + 1 < Z; z++)
    for (int y = 0; y + 1 < Y; y++) {
      const uint32_t *p0 = pW + (long)z * planeWS + (long)y * rowWS, *p1 = p0 + planeWS;
      orc_gpu_row(h, p0, p0 + rowWS, p1, p1 + rowWS, X);
    }
}

/* ---- whole CPU path (mkfAccGetBPFDSimple) ------------------------------- */

/* Zeroes h, thresholds the field into pW (rowWS*Y*Z words) and walks it
 * (src/mkfACC.c:206-239). Returns 0, or -1 when X < 32. */
int orc_bpfd_field_f64(uint64_t h[ORC_BINS], uint32_t *pW, const double *pF, int X, int Y, int Z, double t,
                       uint32_t m) {
  const int rowWS = (X + 31) >> 5;
  memset(h, 0, sizeof(uint64_t) * ORC_BINS);
  orc_pbm_f64(pW, pF, X, Y, Z, t, m);
  return orc_bpfd_chunk(h, pW, X, Y, Z, rowWS, (long)rowWS * Y);
}

int orc_bpfd_field_f32(uint64_t h[ORC_BINS], uint32_t *pW, const float *pF, int X, int Y, int Z, double t,
                       uint32_t m) {
  const int rowWS = (X + 31) >> 5;
  memset(h, 0, sizeof(uint64_t) * ORC_BINS);
  orc_pbm_f32(pW, pF, X, Y, Z, t, m);
  return orc_bpfd_chunk(h, pW, X, Y, Z, rowWS, (long)rowWS * Y);
}

int orc_num_threads(void) { return gThreads; }

void orc_set_threads(int n) {
  if (n < 1) n = 1;
  if (n > ORC_MAX_THREADS) n = ORC_MAX_THREADS;
  gThreads = n;
}

/* ---- cell geometry: Euler weights and symmetry orbits ------------------- */

/* 24 x the Euler-characteristic share of one 2x2x2 cell of the cubical
 * complex spanned by the set voxels (vertices shared by 8 cells, edges by 4,
 * faces by 2): 24 * (V/8 - E/4 + F/2 - C) = 3V - 6E + 12F - 24C. Corner i is
 * (i & 1, (i >> 1) & 1, i >> 2). */
int orc_euler_weight(unsigned code) {
  int V = 0, E = 0, F = 0, C = 0;
  for (int i = 0; i < 8; i++) V += (code >> i) & 1;
  for (int i = 0; i < 8; i++)
    for (int a = 0; a < 3; a++) {
      const int j = i | (1 << a);
      if (j != i && ((code >> i) & 1) && ((code >> j) & 1)) E++;
    }
  /* faces: fix one axis a at value v, the other two free */
  for (int a = 0; a < 3; a++)
    for (int v = 0; v < 2; v++) {
      int all = 1;
      for (int i = 0; i < 8; i++)
        if (((i >> a) & 1) == v && !((code >> i) & 1)) all = 0;
      F += all;
    }
  C = (code == 255u);
  return 3 * V - 6 * E + 12 * F - 24 * C;
}

/* Orbit label of every pattern under the 48 symmetries of the cube (axis
 * permutations x reflections); labels are the smallest pattern of the orbit.
 * Returns the number of orbits. */
int orc_orbits(uint8_t label[ORC_BINS]) {
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int n = 0;
  for (unsigned p = 0; p < ORC_BINS; p++) {
    unsigned best = p;
    for (int s = 0; s < 6; s++)
      for (int flip = 0; flip < 8; flip++) {
        unsigned q = 0;
        for (int i = 0; i < 8; i++) {
          if (!((p >> i) & 1)) continue;
          const int c[3] = {i & 1, (i >> 1) & 1, i >> 2};
          int d[3];
          for (int a = 0; a < 3; a++) d[a] = c[perm[s][a]] ^ ((flip >> a) & 1);
          q |= 1u << (d[0] | (d[1] << 1) | (d[2] << 2));
        }
        if (q < best) best = q;
      }
    label[p] = (uint8_t)best;
    n += (best == p);
  }
  return n;
}

/* ---- measures ----------------------------------------------------------- */

/* The reference's per-group Euler weights (src/mkfUtil.c:33-38) equal
 * orc_euler_weight for 20 of the 22 orbits. For two orbits its table holds
 * other values: the orbit of pattern 0x3d (5 voxels) has -8 where the
 * cubical-complex count gives -9, and the orbit of 0x7e (6 voxels, two
 * opposite corners clear) has -12 where the count gives -18. The oracle
 * restates the reference, so it applies these two entries; every other
 * weight comes from the derivation. (0x3d and 0x7e are the smallest
 * patterns of their orbits.) */
int orc_reference_weight(unsigned orbitRep) {
  if (orbitRep == 0x3du) return -8;
  if (orbitRep == 0x7eu) return -12;
  return orc_euler_weight(orbitRep);
}

/* Kv and Vv as the reference computes them (src/mkfUtil.c:64-69, 94-135),
 * built from the derivations above instead of the reference's tables (with
 * the two reference weights noted at orc_reference_weight): bins are summed
 * per symmetry orbit, each orbit sum truncated to 32-bit int and
 * multiplied by the orbit's Euler weight in wrapping 32-bit arithmetic,
 * summed in int64, scaled by pi/6 in double, stored as float, then scaled by
 * the float 1 / (s^3 * total). m = {Kv, -1, -1, Vv}; returns 2.
 * *pOverflow (optional) is set when some 32-bit product wrapped. */
int orc_measures(float m[4], const uint64_t a[ORC_BINS], float s, int *pOverflow) {
  uint8_t label[ORC_BINS];
  orc_orbits(label);
  uint64_t g[ORC_BINS];
  memset(g, 0, sizeof(g));
  uint64_t t = 0;
  for (int i = 0; i < ORC_BINS; i++) {
    g[label[i]] += a[i];
    t += a[i];
  }
  int64_t k = 0;
  int ovf = 0;
  for (int i = 0; i < ORC_BINS; i++) {
    if (label[i] != i) continue;
    const int w = orc_reference_weight((unsigned)i);
    const int64_t exact = (int64_t)w * (int64_t)(int32_t)(uint32_t)g[i];
    const int32_t wrapped = (int32_t)((uint32_t)w * (uint32_t)g[i]);
    if (exact != wrapped || (uint64_t)(uint32_t)g[i] != g[i]) ovf = 1;
    k += wrapped;
  }
  m[0] = (float)((double)k * M_PI / 6);
  uint64_t s0 = 0, s1 = 0;
  for (int i = 0; i < ORC_BINS; i += 2) {
    s0 += a[i];
    s1 += a[i + 1];
  }
  m[3] = (float)s1 / (float)(s0 + s1);
  m[1] = m[2] = -1;
  if (s > 0) {
    const float cube = s * s * s;
    const float rkV = (float)(1.0 / (double)(cube * (float)t));
    m[0] *= rkV;
  }
  if (pOverflow) *pOverflow = ovf;
  return 2;
}
