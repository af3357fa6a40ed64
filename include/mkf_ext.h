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
 * (1, or 4 for float / 2 for double fields, 0 = the library's choice; the
 * histogram of a PBM uses 4 unless 1 is asked for; several fields or an x
 * stride use 1), holds rowPairs (2, 4 or 8) row pairs and marches over
 * planePairs (1..65536) plane pairs. voxelsPerLane 16 or 8 (float fields,
 * 2 or 4 row pairs; the fused histogram only) selects a persistent kernel of
 * one-wave workgroups, each lane holding 4 or 2 groups of 4 voxels 256 apart
 * (a wave covers 1024 or 512 voxels of a row); `waves` is then ignored. The
 * histogram of a PBM also takes 16 (half PBM words per lane); it reads 8 as
 * 4. */
typedef struct {
  int32_t planePairs;
  int32_t waves;
  int32_t rowPairs;
  int32_t voxelsPerLane;
} MkfTiling;

/* Text of the last error raised by any entry point in this thread. */
const char *mkfLastError(void);

/* Rows narrower than one PBM word (X < 32). The reference's row walk
 * (src/bpfdCUDA.cuh:92-118, src/mkfACC.c:121-152) emits 2X - 1 patterns per
 * row pair for them: its GPU walk reads the row's only word twice, which
 * with zero tail bits counts every cell of the row pair twice plus the
 * wrap-around cell (x = X-1 beside x = 0); its CPU walk reads the next row's
 * word instead, and past the map on the last row. This process-wide mode
 * (default MKF_NARROW_CELLS) sets what the drop-in histogram entry points
 * (mkfCUDAGetBPFD[H], mkfCUDAGetBPFDautoCtx, mkfBPFDAsync, mkfAccGetBPFDSimple,
 * mkfAccCUDAGetBPFD) do with such fields:
 *   MKF_NARROW_CELLS      count the X-1 cells of each row pair (the histogram
 *                         the fused, batched and serving forms always give);
 *   MKF_NARROW_REFERENCE  the GPU walk's counts (mkfAccGetBPFDSimple, whose
 *                         reference is the CPU walk, refuses);
 *   MKF_NARROW_REJECT     refuse: NULL from mkfCUDAGetBPFD[H], 0 from the
 *                         mkfAcc* calls, MKF_ERR_UNSUPPORTED from mkfBPFDAsync,
 *                         the histogram untouched.
 * With MKF_VERBOSE set, MKF_NARROW_CELLS logs each such call. Returns the
 * previous mode, or MKF_ERR_ARG. No counterpart in the reference. */
#define MKF_NARROW_CELLS 0
#define MKF_NARROW_REFERENCE 1
#define MKF_NARROW_REJECT 2
int mkfSetNarrowRowMode(int mode);

/* Test hook for error paths (no counterpart in the reference): the nth call
 * from now (1 = the next) of the named internal step fails as an allocation
 * or runtime error would, once; nth 0 disarms the point. Points:
 * "batch.upload" (a batch's work-queue upload) and "batch.capture" (its
 * graph capture). Returns MKF_OK, or MKF_ERR_ARG for an unknown point. */
int mkfDebugInjectFault(const char *point, int nth);

/* Fused threshold + histogram: ADDS the BPFD of the field (device pointers in
 * pI, as for binMapCUDA) to pBPFD (device, 256 x uint64), reading each field
 * byte from HBM once and writing no packed map. pO (optional) receives the
 * PBM organisation. With pK the call blocks and pK->dtms[1] receives the
 * kernel time (dtms[0] = 0); without it the work is only queued on stream. */
int mkfFusedBPFD(KernInfo *pK, size_t *pBPFD, BMOrg *pO, const BMFieldInfo *pI, const BinMapF64 *pM,
                 const MkfTiling *pT, void *stream);

/* Stream-ordered packed binary map into caller memory pW (device). */
int mkfPBMAsync(BMPackWord *pW, BMOrg *pO, const BMFieldInfo *pI, const BinMapF64 *pM, void *stream);

/* Phase moments fused into the packed-binary-map pass: as mkfPBMAsync (pW
 * may be NULL to write no map), and ADDS to pMom (device, MKF_MOM_N doubles)
 * the sum and sum of squares of the field values per phase (map bit 0, 1)
 * and the phase counts:
 *   {sum0, sumsq0, sum1, sumsq1, n0, n1}
 * The first four follow the categorical layout of the reference's
 * CUDAImgMom::addCat (src/binMapCUDA.cuh:89-121: category c at 2c, its
 * square at 2c + 1), which the reference declares but never instantiates.
 * Several fields contribute their sum (the thresholded value). Values are
 * converted to double and summed in double; per-workgroup partials are added
 * in a fixed order, so repeated calls give identical results. A NaN voxel
 * makes its phase's sums NaN. */
#define MKF_MOM_N 6
int mkfPBMMomentsAsync(BMPackWord *pW, double *pMom, BMOrg *pO, const BMFieldInfo *pI, const BinMapF64 *pM,
                       void *stream);

/* Histogram and phase moments in one read of the field: ADDS the BPFD to
 * pBPFD as mkfFusedBPFD does and the moments to pMom as mkfPBMMomentsAsync
 * does (same layout, same double sums in a fixed order). pT NULL selects a
 * tiling for this kernel (large float fields: 16 voxels per lane for rows of
 * whole 1024-voxel segments, else 8; small float fields 4, doubles 2); a
 * given tiling is used with 4 row pairs. A field with a dimension below 2
 * has no cells: the histogram is unchanged and the moments are still added.
 * At 1024^3 f32 one call takes about 45 % less time than mkfFusedBPFD
 * followed by mkfPBMMomentsAsync without a map. No reference counterpart
 * (see mkfPBMMomentsAsync). */
int mkfFusedBPFDMoments(size_t *pBPFD, double *pMom, BMOrg *pO, const BMFieldInfo *pI, const BinMapF64 *pM,
                        const MkfTiling *pT, void *stream);

/* As mkfFusedBPFDMoments, but allocating nothing: no hipMalloc, hipFree or
 * allocating library call. No counterpart in the reference; it is for callers
 * that cannot allocate (code under HIP graph capture, serving loops and
 * harnesses that forbid device allocation). The per-workgroup moment
 * partials live in a static device array of MKF_NOALLOC_SCRATCH_BYTES that
 * each device holds until the process ends; a launch needing more partials
 * than fit (MKF_NOALLOC_SCRATCH_BYTES / 48, about 87,000 workgroups; a
 * 1024^3 float field takes 8,192 with the built-in tiling) fails with
 * MKF_ERR_BUDGET before any work is queued. Every call on one device shares
 * that device's array: calls that overlap on one device (other streams or
 * threads) race, so the caller serialises them. */
#define MKF_NOALLOC_SCRATCH_BYTES (4 << 20)
int mkfFusedBPFDMomentsNoAlloc(size_t *pBPFD, double *pMom, BMOrg *pO, const BMFieldInfo *pI, const BinMapF64 *pM,
                               const MkfTiling *pT, void *stream);

/* Stream-ordered histogram of a packed binary map; adds into pBPFD. */
int mkfBPFDAsync(size_t *pBPFD, const BMOrg *pO, const BMPackWord *pW, const MkfTiling *pT, void *stream);

/* Tiling the library chooses for the fused kernel on a field of dims def
 * and element size (the histogram of a PBM has its own default). The choice
 * depends on the current device: its CU count and how many workgroups of the
 * persistent kernels it holds at once. def NULL gives the all-zero tiling,
 * which every entry point refuses, and sets mkfLastError. */
MkfTiling mkfDefaultTiling(const FieldDef def[3], int elemBytes);

/* Geometry of a histogram launch: its tiles, the workgroups launched, the
 * waves of each, and the tiling the kernel runs with (after the adjustments
 * the library makes, e.g. one voxel per lane for several fields). The
 * persistent kernels (voxelsPerLane 16 and 8) launch at most one workgroup
 * per resident slot, and workgroup b walks tiles b, b + workgroups, ... */
typedef struct {
  int64_t tiles;
  int64_t workgroups;
  int32_t wavesPerGroup;
  MkfTiling used;
} MkfLaunchInfo;

/* What mkfFusedBPFD would launch for this field, relation (pM NULL: ">=")
 * and tiling (pT NULL: the tiling a call without one uses). Launches
 * nothing; needs the current device for the resident-slot count. */
int mkfFusedLaunchInfo(const BMFieldInfo *pI, const BinMapF64 *pM, const MkfTiling *pT, MkfLaunchInfo *pInfo);
/* The same for mkfBPFDAsync on a PBM of organisation pO. */
int mkfBPFDLaunchInfo(const BMOrg *pO, const MkfTiling *pT, MkfLaunchInfo *pInfo);

/* Measurement aid, no counterpart in the reference: reads `bytes` (a
 * multiple of 16, 16-byte aligned) of device memory p once with 16-byte
 * loads from blocksPerCU (1..64) workgroups of 4 waves per CU, and XORs
 * every word into *pOut (device). Each wave streams blocks of loadsInFlight
 * (4, 8 or 16) KiB of consecutive bytes, all loads of a block in flight at
 * once. flags: MKF_ROOF_NONTEMPORAL (non-temporal loads), MKF_ROOF_CHUNKED
 * (each wave reads one contiguous run of blocks; otherwise waves take
 * blocks round robin). Timed around the launch, it gives the practical HBM
 * read roof the histogram kernels are compared with. */
#define MKF_ROOF_NONTEMPORAL 1
#define MKF_ROOF_CHUNKED 2
int mkfReadRoof(const void *p, size_t bytes, int flags, int loadsInFlight, int blocksPerCU, unsigned *pOut,
                void *stream);

/* Times the fused kernel on this field (the caller's data, a scratch
 * histogram) under a few tilings, keeps the fastest for fields of this shape
 * and encoding (process-wide), and stores it in *pChosen (may be NULL).
 * Later mkfFusedBPFD calls without a tiling use it. Blocking. */
int mkfFusedAutotune(const BMFieldInfo *pI, const BinMapF64 *pM, MkfTiling *pChosen, void *stream);

/* ---------------------------------------------------------------------------
 * Batched form: many fields x many threshold levels in one launch
 * ------------------------------------------------------------------------- */
#define MKF_BATCH_MAX_LEVELS 16
#define MKF_BATCH_ZERO_OUTPUT 1 /* flag: clear every output before each run */

/* One field of a batch: a planar device field of dims def, evaluated at
 * nLevels threshold maps pM[] (host array). Its nLevels consecutive 256-bin
 * histograms at pBPFD (device, uint64) are ADDED to by each run (or cleared
 * first with MKF_BATCH_ZERO_OUTPUT). */
typedef struct {
  const void *pF;
  NumEnc enc;
  FieldDef def[3];
  int32_t nLevels;
  const BinMapF64 *pM;
  size_t *pBPFD;
} MkfBatchField;

/* Work-item (slab) shape of a batch. A work item is a slab of planePairs
 * plane pairs by 8 row-stacks-worth of row pairs by one x-tile of
 * `segments` segments of 64 cells (63 for fields in the Gray-plane form:
 * 3 or more levels, all ">=" / "<" or all ">" / "<=", "all" and "none"
 * joining either, no NaN threshold); the 8 waves of a workgroup cover
 * 8 / segments row stacks of 4 row pairs each. */
typedef struct {
  int32_t planePairs; /* slab depth, 1..4096; 0 selects 32 */
  int32_t segments;   /* largest x-tile: 1, 2, 4 or 8 segments; 0 selects 8. Each field
                         uses the largest power of two up to this that wastes the fewest
                         segments at the end of its rows */
} MkfBatchShape;

typedef struct MkfBatch MkfBatch;

/* Plans a batch: a device work queue of (field, x-tile, y-block, slab) items,
 * each field read once for all of its levels by a persistent kernel,
 * captured with the queue reset into a HIP graph. shape NULL selects the
 * defaults; see mkfBatchAutotune. NULL on error (mkfLastError). */
MkfBatch *mkfBatchCreate(const MkfBatchField *fields, int nFields, const MkfBatchShape *shape, int flags,
                         void *stream);
/* Replays the captured graph on stream (asynchronous). Runs on one device
 * overlap only if their outputs are distinct. */
int mkfBatchRun(MkfBatch *b, void *stream);
/* Launches the same kernels without the graph (for comparison). */
int mkfBatchRunDirect(MkfBatch *b, void *stream);
/* Times each candidate shape (one warm-up and reps runs each, into scratch
 * histograms: the outputs are not touched), keeps the fastest, re-captures
 * and stores it in *chosen (may be NULL). Returns MKF_OK or an error. */
int mkfBatchAutotune(MkfBatch *b, const MkfBatchShape *candidates, int nCandidates, int reps,
                     MkfBatchShape *chosen, void *stream);
/* Work items, shape and persistent grid size of a plan (pointers may be NULL). */
int mkfBatchInfo(const MkfBatch *b, int *nItems, MkfBatchShape *shape, int *grid);
void mkfBatchDestroy(MkfBatch *b);

/* ---------------------------------------------------------------------------
 * Host fields streamed through the device in z-slabs
 * ------------------------------------------------------------------------- */
/* Adds the histogram of a HOST field (planar, f32 or f64) to pBPFDH (host,
 * 256 counts). The field goes through two device slab buffers of
 * slabPlanes + 1 planes (slabPlanes <= 0: about 256 MiB per slab):
 * the copy of slab k+1 overlaps the fused kernel on slab k, and consecutive
 * slabs share one plane, so device memory stays bounded for any field size.
 * Blocking. Pinned host memory lets the copies run at the full link rate.
 * No counterpart in the reference, whose host-field path
 * (mkfAccCUDAGetBPFD, src/mkfACC.c:248-318) copies the whole field; that
 * entry point uses this one here. */
int mkfHostStreamBPFD(size_t *pBPFDH, const void *pF, NumEnc enc, const FieldDef def[3], const BinMapF64 *pM,
                      int slabPlanes);

/* ---------------------------------------------------------------------------
 * Serving form: a request loop with continuous batching
 *
 * A worker thread admits queued requests while fewer than maxInFlight are
 * admitted: host requests in arrival order, while few copies are queued and
 * the device bytes held for copied-in fields stay within hbmBudget;
 * device-resident requests (no copy, no budget) ahead of host requests that
 * wait. It copies host fields in on a copy stream and launches the batched
 * kernel over every request whose field is on the device as soon as one of
 * its two launch slots is free, so requests that arrive while others run
 * join the next launch. Results
 * (histograms written to the request's host buffer) are reported through
 * mkfServePoll. No counterpart in the reference.
 * ------------------------------------------------------------------------- */
typedef struct {
  size_t hbmBudget;     /* device bytes for copied-in fields; 0 = half the free memory at creation */
  int32_t maxInFlight;  /* admitted, not yet reported requests; 0 = 64 */
  int32_t maxBatch;     /* requests per launch; 0 = 32 */
  MkfBatchShape shape;  /* work-item shape of every launch; zeros select the defaults */
} MkfServeConfig;

#define MKF_SERVE_ON_DEVICE 1u /* request flag: pF is device memory, read in place */
#define MKF_SERVE_AFTER_STREAM 2u /* request flag: the field is read only after the work queued on
                                     `stream` at submission (the producer's stream; NULL is the null stream) */

typedef struct {
  const void *pF;      /* host field (pinned memory overlaps its copy) or, with MKF_SERVE_ON_DEVICE, device field */
  uint32_t flags;
  void *stream;        /* with MKF_SERVE_AFTER_STREAM: the stream that produces the field */
  NumEnc enc;          /* ENC_F32 or ENC_F64, planar (x fastest) */
  FieldDef def[3];
  int32_t nLevels;     /* 1 .. MKF_BATCH_MAX_LEVELS */
  const BinMapF64 *pM; /* nLevels maps, copied at submission */
  size_t *pBPFD;       /* host output: nLevels x 256 counts, written before the result is reported */
  uint64_t tag;        /* returned in the result */
} MkfServeRequest;

typedef struct {
  uint64_t tag;
  int32_t status;      /* MKF_OK or an error code */
  int32_t batchSize;   /* requests in the launch that served it */
  double tSubmit, tAdmit, tLaunch, tDone; /* seconds on the server clock (mkfServeClock) */
} MkfServeResult;

typedef struct {
  uint64_t submitted, completed, launches;
  uint64_t bytesIn;    /* field bytes copied host to device */
  uint64_t heldMax;    /* high-water mark of device bytes held for fields */
  double gpuBusy;      /* summed launch-to-completion seconds */
} MkfServeStats;

typedef struct MkfServer MkfServer;

/* Starts a server on the current device (cfg may be NULL). NULL on error. */
MkfServer *mkfServeCreate(const MkfServeConfig *cfg);
/* Queues a request; thread-safe and non-blocking. A host field larger than
 * the budget is served as z-slabs that share one plane, of at most a quarter
 * of the budget (or, when two planes exceed a quarter, of as many planes as
 * the budget holds); each slab is an internal request admitted, copied and
 * batched like any other, and the result is reported once, after the last
 * slab, with the summed histograms. MKF_ERR_BUDGET only if two planes exceed
 * the budget.
 * The field, and for a device field its memory, must stay valid and
 * unchanged until the result is reported. */
int mkfServeSubmit(MkfServer *s, const MkfServeRequest *r);
/* Moves up to maxOut reported results to out, waiting up to waitMicros for
 * the first. Returns the number moved, or < 0 on error. */
int mkfServePoll(MkfServer *s, MkfServeResult *out, int maxOut, int waitMicros);
double mkfServeClock(const MkfServer *s);
int mkfServeGetStats(MkfServer *s, MkfServeStats *st);
/* Finishes every queued request, stops the worker and frees the server. */
void mkfServeDestroy(MkfServer *s);

/* Crofton-formula estimates of all four Minkowski measures from a BPFD, after
 * Ohser & Muecklich (2000): m = {Kv, Mv, Sv, Vv}, in the order of MKF_REFML.
 * No counterpart in the reference: its book measures (mkfRefMeasureBPFD,
 * src/mkfUtil.c:83-90) need refMeasures.c, which is absent from its tree, and
 * the per-group S/M weights in src/mkfHacks.c:9-21 are experiments that do not
 * converge on digitised balls. This function is NOT the reference's Mv/Sv.
 *  - Kv: 4 pi chi of the 6-adjacency foreground (the textbook Euler weights;
 *    mkfSelectMeasuresBPFD keeps the reference's table, which differs on two
 *    pattern groups);
 *  - Sv: 2 x mean boundary crossings per unit length over the 13 lattice
 *    directions (Crofton), weighted by spherical Voronoi cells;
 *  - Mv: 2 pi x mean Euler number per unit area of the sections by the 13
 *    lattice-plane orientations (integral of mean curvature (k1+k2)/2);
 *  - Vv: fraction of cells whose lowest corner is set (as volFrac).
 * s > 0: densities per unit volume of a lattice of spacing s (the volume is
 * the cell count times s^3). s == 0: totals in lattice units {K, M, S, cells
 * with the lowest corner set}. Returns 4, or MKF_ERR_ARG for NULL, s < 0, or
 * an empty histogram with s > 0. Thread-safe. */
int mkfCroftonMeasuresBPFD(double m[4], const size_t aBPFD[MKF_BINS], double s);

/* The number of Kv terms (of 22 symmetry groups) that mkfSelectMeasuresBPFD
 * computes, as the reference does (src/mkfUtil.c:64-69), in wrapping 32-bit
 * arithmetic and that therefore differ from the exact Euler sum: the group's
 * cell count does not fit 32 bits, or its product with the group weight
 * wraps. 0: Kv is exact. -1 for NULL. With MKF_VERBOSE set,
 * mkfSelectMeasuresBPFD reports a nonzero count on stderr. No counterpart
 * in the reference. */
int mkfKvWrapCount(const size_t aBPFD[MKF_BINS]);

/* Read-only views of the measure tables used by mkfSelectMeasuresBPFD: the
 * symmetry group of each of the 256 patterns and the Euler weight of each of
 * the 22 groups (src/mkfUtil.c:33-58), for callers that verify them. */
const uint8_t *mkfPatternGroupTable(void);
const int8_t *mkfGroupWeightKTable(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* MKF_AMD_EXT_H */
