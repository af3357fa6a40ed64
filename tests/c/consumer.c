/* consumer.c - a C caller of libmkf_amd.so written against the reference's
 * API the way its external consumers use it (a field in device memory,
 * binMapCUDA then mkfCUDAGetBPFD; the host-field entry point; the measures).
 * Test infrastructure: tests/test_c_consumer.py compiles it with gcc and
 * compares its output with the oracle.
 *
 * usage: consumer X Y Z field.bin hist.txt
 * Writes the float32 field it generated to field.bin and, to hist.txt, four
 * lines: the device-path histogram (256 counts), the host-result histogram,
 * the host-field histogram, and the measures "Kv Vv" with their return code. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "mkf.h"

static void putHist(FILE *f, const size_t *h) {
  for (int i = 0; i < MKF_BINS; i++) fprintf(f, "%zu%c", h[i], i + 1 < MKF_BINS ? ' ' : '\n');
}

int main(int argc, char **argv) {
  if (argc != 6) return 2;
  const int X = atoi(argv[1]), Y = atoi(argv[2]), Z = atoi(argv[3]);
  const size_t n = (size_t)X * Y * Z;
  float *hF = (float *)malloc(n * sizeof(float));
  if (!hF) return 3;
  for (int z = 0; z < Z; z++)
    for (int y = 0; y < Y; y++)
      for (int x = 0; x < X; x++)
        hF[((size_t)z * Y + y) * X + x] = sinf(0.21f * x + 0.3f) * cosf(0.17f * y) + 0.5f * sinf(0.13f * z + 0.1f * x);
  FILE *fb = fopen(argv[4], "wb");
  if (!fb || fwrite(hF, sizeof(float), n, fb) != n) return 4;
  fclose(fb);

  void *dF = NULL;
  size_t *dH = NULL;
  if (hipMalloc(&dF, n * sizeof(float)) != hipSuccess || hipMalloc((void **)&dH, MKF_BINS * sizeof(size_t)) != hipSuccess)
    return 5;
  if (hipMemcpy(dF, hF, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return 6;
  if (hipMemset(dH, 0, MKF_BINS * sizeof(size_t)) != hipSuccess) return 7;

  BinMapF64 bm;
  setBinMapF64(&bm, ">=", 0.1f);
  int def[3] = {X, Y, Z};
  ConstFieldPtr tab[1];
  tab[0].p = dF;
  BMFieldInfo fi;
  setupFields(&fi, (void **)tab, 1, def, sizeof(float), 0);
  BMOrg bmo;
  KernInfo ki;
  BMPackWord *pW = binMapCUDA(&ki, NULL, &bmo, &fi, &bm); /* library PBM buffer */
  if (!pW) return 8;
  if (!mkfCUDAGetBPFD(&ki, dH, &bmo, pW, MKFCU_PROFILE_FAST)) return 9;
  size_t h1[MKF_BINS], h2[MKF_BINS], h3[MKF_BINS];
  if (hipMemcpy(h1, dH, sizeof(h1), hipMemcpyDeviceToHost) != hipSuccess) return 10;
  if (!mkfCUDAGetBPFDH(&ki, h2, &bmo, pW, MKFCU_PROFILE_FLEX)) return 11;
  memset(h3, 0, sizeof(h3));
  if (!mkfAccCUDAGetBPFD(h3, hF, def, ENC_F32, &bm)) return 12;

  float m[4];
  char sym[12];
  const int nm = mkfSelectMeasuresBPFD(m, sym, h1, 3.0f / (X + Y + Z), 2);

  FILE *fh = fopen(argv[5], "w");
  if (!fh) return 13;
  putHist(fh, h1);
  putHist(fh, h2);
  putHist(fh, h3);
  fprintf(fh, "%.9g %.9g %d\n", (double)m[0], (double)m[3], nm);
  fclose(fh);

  binMapCUDACleanup();
  mkfCUDACleanup(0);
  (void)hipFree(dF);
  (void)hipFree(dH);
  free(hF);
  return 0;
}
