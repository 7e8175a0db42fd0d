/*
 * gdsjoin.h -- C ABI of the MI355X exact Euclidean epsilon-grid self-join
 * (libgdsjoin.so).  Plain pointers and sizes only.
 *
 * Legacy drop-in symbols replace the reference's Python-facing C ABI:
 *   GDSJoinPy                  main.cu:360   (6-argument layout kept)
 *   copyResultIntoPythonArray  main.cu:507
 * The handle API (gdsj_*) replaces the reference's internal host driver for
 * the same path: ReorderByDimension (main.cu:981), generateNDGridDimensions
 * (main.cu:891), populateNDGridIndexAndLookupArray (main.cu:690) and
 * distanceTableNDGridBatches (GPU.cu:270), which the CLI main mode 3 calls
 * (main.cu:246-254).
 *
 * Result contract (identical to the reference on the same inputs): for each
 * point i, the ids j of every point (i itself included) in a non-empty grid
 * cell adjacent to i's cell in all k indexed dimensions with
 * sqrt(acc(i,j)) <= epsilon, where acc is the DTYPE sum of squared
 * differences accumulated in the reference's FMA order for the chosen ILP
 * (kernel.cu:307-365).  The table is CSR in point-id order: counts[i] and
 * the concatenation of every point's neighbor ids.
 */
#ifndef GDSJOIN_H
#define GDSJOIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference exits the process instead) ---------- */
#define GDSJ_OK 0
#define GDSJ_EINVAL (-1)      /* bad argument (dims, k, epsilon <= 0, non-finite data) */
#define GDSJ_ERANGE (-2)      /* product of cells per dimension overflows uint64 */
#define GDSJ_EOOM (-3)        /* device or host allocation failed */
#define GDSJ_EDEVICE (-4)     /* HIP runtime error */
#define GDSJ_EOVERFLOW (-5)   /* caller's output buffer smaller than the result */
#define GDSJ_ESTATE (-6)      /* call order violated (e.g. fill before join) */
#define GDSJ_EUNSUPPORTED (-7)

#define GDSJ_DTYPE_F32 0
#define GDSJ_DTYPE_F64 1

/* filter modes */
#define GDSJ_FILTER_PLAIN 0   /* direct differencing only */
#define GDSJ_FILTER_GRAM 1    /* MFMA Gram-form filter + exact recheck */
#define GDSJ_FILTER_AUTO 2    /* GRAM when dims >= 16, else PLAIN */
#define GDSJ_FILTER_GRAM_BF16 3 /* GRAM with bf16 hi + lo operands (three MFMAs per k-step)
                                   for FP32 data; GRAM, AUTO and FP64 data use fp16 */

/* result-size estimators (sizing only; never change the result) */
#define GDSJ_EST_EXACT 0      /* exact counting pass, then fill */
#define GDSJ_EST_LEARNED 1    /* MLP-predicted per-tile slots; exact pass on overflow */
#define GDSJ_EST_SAMPLE 2     /* reference sampling estimator (every 66th query) */

/* Runtime form of the reference's compile-time params.h (params.h:17-93).
 * gdsj_params_default() gives params.h as shipped except dims/k, which
 * follow the data: k = min(dims, 6). */
typedef struct gdsj_params {
    uint32_t k;             /* NUMINDEXEDDIM, 1..12 and <= dims (0 = min(dims,6)) */
    uint32_t ilp;           /* ILP: 0 = sequential FMA accumulator; else partials */
    uint32_t shortcircuit;  /* SHORTCIRCUIT (speed only) */
    uint32_t reorder;       /* REORDER: permute dims by descending variance */
    uint32_t stamp;         /* unicomp, each adjacent cell pair evaluated once and both directions
                               emitted (speed only): 0 off, 1 STAMP (stamploops.h ownership),
                               2 LINEARSTAMP (kernel.cu:272-289, the lower linear id owns) */
    uint32_t queryreorder;  /* QUERYREORDER: 1 heaviest tiles first; 2 by occupancy class
                               (powers of two), cell order within a class (speed only) */
    uint32_t filter;        /* GDSJ_FILTER_* */
    uint32_t estimator;     /* GDSJ_EST_* */
    uint32_t sort_neighbors;/* 1: each point's neighbor ids ascending */
    uint32_t verbose;       /* 1: print progress to stdout */
    uint32_t tile_queries;  /* queries per work tile, 0 = autotuned */
    uint32_t schedule;      /* tile->workgroup order: 0 auto, 1 round-robin, 2 XCD-contiguous */
    uint32_t prune;         /* skip neighbor cells farther than eps from a tile's query box:
                               0 auto, 1 on, 2 off (speed only) */
    uint32_t train;         /* learned estimator: 0 train until first trained, 1 after every
                               join, 2 never */
    uint32_t filter_stats;  /* 1: the Gram filter also measures its error on the rechecked
                               pairs (gdsj_stats.filter_max_err; slower) */
    uint32_t hitmask;       /* exact passes without the Gram filter: the counting pass keeps
                               one bit per candidate pair and the fill pass reads them back
                               instead of recomputing distances: 0 auto (on from 4 dims
                               when they take at most 8 GiB), 1 on (error if they do not
                               fit), 2 off (speed only) */
    uint32_t ilp_fold;      /* ILP > 0 with SHORTCIRCUIT = 1: which contraction of the fold
                               "runningDist[0] += runningDist[j]" (kernel.cu:320-334) the sums
                               follow.  0: each group's first dim fused into the running sum,
                               the others rounded (t*t) and added (default); 1: every dim fused
                               into the running sum, which is the arithmetic of ILP = 0.  The
                               two readings differ only on pairs whose sums straddle the
                               threshold by a few ulps */
    uint32_t vfilter;       /* Gram-form prefilter of the plain FP32 kernels when every dim is
                               indexed (3 to 8 dims): pairs are decided from |c'|^2 - 2 q'.c'
                               in tile-local coordinates, and those within a proven error
                               band of the threshold by the reference's arithmetic: 0 auto
                               (off), 1 on where it applies, 2 off (speed only) */
    uint32_t sort;          /* SORT (params.h:46): each cell's points ordered by their value at
                               dim 0 when every dim is indexed, else at dim k after REORDER
                               ("sortidu"), and each query walks only the part of a neighbor
                               cell the reference walks (kernel.cu:400-467).  Changes results
                               at the epsilon boundary of the sort dim as the reference does.
                               Runs with the plain kernels, without unicomp. 0 off, 1 on */
    uint32_t cpu_sum;       /* all-pairs join only: 1 sums the squared differences as the
                               reference's CPU paths do, sequentially and unfused
                               (runningDist += t*t, main.cu:963-967; the CLI's mode 0), in
                               place of the kernels' reading of ILP / SHORTCIRCUIT */
    uint32_t mfma;          /* data whose dims are all indexed (3 to 7 dims): the pair tests of the
                               tiles of well-occupied cells run on the matrix cores (FP32 MFMA, the
                               Gram form in tile-local coordinates, concurrently with the plain
                               kernels on the other tiles); pairs within a proven error band of the
                               threshold are decided by the reference's arithmetic, so the table is
                               unchanged (speed only).  0 auto (off: where the pass is bound by
                               candidate fetches, as on 6-D mixtures, it gains nothing), 1 every
                               tile, 2 off, n >= 3: the tiles of cells of at least n points */
    uint32_t ffilter;       /* ILP > 0, counting passes of the plain kernels (register-resident
                               pair tests): each pair's squared differences are first summed by
                               one FMA per dim and compared with a proven error band around the
                               threshold; a 64-pair chunk with a pair inside the band is
                               evaluated in the reference's order, so the counts are unchanged
                               (speed only).  0 auto (on), 1 on, 2 every pair in the band (the
                               reference's arithmetic for all) */
    uint32_t index;         /* the dims the grid indexes (speed only: the pair test is unchanged).
                               1 the reference's k dims of largest variance (REORDER); 2 the top
                               k principal axes of the REORDER sample (y = R x, cells slightly
                               wider than eps so that every pair the reference's arithmetic
                               accepts lies in adjacent cells), used only when no point lies
                               within rounding of a face of the reference's own grid (so that
                               the reference finds every such pair too), with more dims than k,
                               up to 256 dims, no unicomp and no SORT; else 1.  0 auto: 2 where
                               the principal axes leave at most a twentieth of the variance the
                               reference's dims leave (data near a rotated subspace) */
    uint32_t index_dims;    /* principal axes the grid indexes when index 2 is used (at most k;
                               0: k - 2 from k = 6, else k).  Fewer axes give larger cells: more
                               candidates per query, but more queries per tile sharing them */
    uint32_t gram_axes;     /* the Gram filter's operands (speed only: every pair within eps' is
                               kept and rechecked by the reference's arithmetic).  1 the points'
                               own dims; 16 or 32: the point's coordinates on that many principal
                               axes of the REORDER sample (32 to 128 dims, fewer axes than the
                               padded dims, no SORT; eps' widened by the basis' spectral norm and
                               the bound by the projection's rounding), m / d of the operand
                               bytes and MFMAs; 0 auto: 16 axes where twice the variance they
                               leave out is under a hundredth of eps^2, else the own dims */
} gdsj_params;

/* Timing and diagnostics of the last join on a handle. */
typedef struct gdsj_stats {
    uint64_t total_pairs;        /* sum of counts, self-pairs included */
    uint64_t candidates;         /* point comparisons (COUNTMETRICS analogue) */
    uint64_t n_nonempty_cells;
    uint64_t total_cells;        /* product of cells per indexed dim */
    uint64_t n_tiles;
    uint64_t n_runs;
    uint64_t filter_rechecks;    /* Gram filter: pairs sent to the exact recheck */
    uint64_t filter_hits;        /* Gram filter: rechecked pairs that are neighbors */
    uint64_t est_predicted;      /* estimator: predicted total pairs */
    uint64_t est_capacity;       /* estimator: result slots allocated (pairs) */
    uint64_t est_overflow_tiles; /* tiles whose slot overflowed (completed exactly) */
    uint64_t est_tiles;          /* tiles sized by the estimator */
    double est_mean_rel_err;     /* mean |predicted - actual| / actual over tiles */
    double est_max_rel_err;
    double filter_max_err;       /* Gram filter, with params.filter_stats: max over rechecked pairs
                                    of |Gram-form squared distance - FP64 squared distance| / eps^2 */
    double ms_grid;              /* reorder + grid + neighbor runs (device) */
    double ms_estimate;          /* estimator inference + slot sizing */
    double ms_count;             /* counting (or slotted) join pass */
    double ms_scan;              /* exclusive scan of counts -> CSR offsets */
    double ms_compact;           /* slot compaction into the CSR */
    double ms_fill;              /* fill pass (all tiles, or overflowed tiles) */
    double ms_sort;              /* per-point row sort (sort_neighbors) */
    double ms_train;             /* estimator training on this join's counts */
    double ms_total;             /* sum of the phases above except training */
    uint32_t perm[64];           /* dimension order after REORDER (first 64) */
    uint64_t mfma_tiles;         /* tiles whose pair tests ran on the matrix cores (gdsj_params.mfma) */
    uint64_t index_used;         /* 1 the reference's dims, 2 the principal axes (gdsj_params.index) */
    uint64_t index_risky;        /* principal axes asked for: points within rounding of a face of the
                                    reference's grid (any: the reference's dims were used) */
    uint64_t gram_axes;          /* principal axes the Gram filter's operands were taken on (0: own dims) */
} gdsj_stats;

typedef struct gdsj_handle gdsj_handle;

/* Fills *p with the defaults (params.h as shipped, k follows the data). */
int gdsj_params_default(gdsj_params *p);

/* One handle per device; the handle owns its stream and scratch buffers. */
int gdsj_create(int device, gdsj_handle **out);
int gdsj_destroy(gdsj_handle *h);

/* Device-resident join.  d_data: n x dims row-major DTYPE on the handle's
 * device.  d_counts: caller's uint32[n].  d_ids: caller's uint32[ids_cap]
 * (may be NULL with ids_cap 0 to only count).  d_offsets: optional caller
 * uint64[n+1] exclusive prefix of counts (NULL = internal).
 * On return *total holds the number of pairs; GDSJ_EOVERFLOW when
 * total > ids_cap (counts and offsets are valid, ids untouched), after which
 * gdsj_fill_device() with a large enough buffer completes the table without
 * recomputing the grid or the counts.  stream: hipStream_t or NULL for the
 * handle's own stream; the call returns after the stream has drained. */
int gdsj_join_device(gdsj_handle *h, const void *d_data, uint32_t n, uint32_t dims,
                     int dtype, double epsilon, const gdsj_params *p,
                     uint32_t *d_counts, uint64_t *d_offsets, uint32_t *d_ids,
                     uint64_t ids_cap, uint64_t *total, void *stream);

/* Completes the table of the last join on h into d_ids (cap >= total). */
int gdsj_fill_device(gdsj_handle *h, uint32_t *d_ids, uint64_t ids_cap);

/* Runs the last gdsj_join_device of h again over the grid, neighbor runs and
 * tiles it built, on the same points (which must be unchanged since), with
 * the same epsilon and parameters: the counting (or slotted) pass, the scan
 * and the fill.  Replaces nothing in the reference, which rebuilds the index
 * on every call (GDSJoinPy main.cu:428-450); for repeated joins of one
 * dataset.  Outputs as gdsj_join_device; GDSJ_ESTATE without a preceding
 * successful join. */
int gdsj_rejoin_device(gdsj_handle *h, uint32_t *d_counts, uint64_t *d_offsets, uint32_t *d_ids,
                       uint64_t ids_cap, uint64_t *total, void *stream);

/* Host-buffer join (the path GDSJoinPy takes): copies data in, runs the
 * join, writes counts[n] to host, keeps the table on the handle for
 * gdsj_result_copy(). */
int gdsj_join_host(gdsj_handle *h, const void *data, uint32_t n, uint32_t dims,
                   int dtype, double epsilon, const gdsj_params *p,
                   uint32_t *counts, uint64_t *total);
int gdsj_result_copy(gdsj_handle *h, uint32_t *ids, uint64_t cap);

int gdsj_get_stats(gdsj_handle *h, gdsj_stats *out);

/* All-pairs join: replaces the reference's search mode 1
 * (makeDistanceTableGPUBruteForce GPU.cu:1285-1480, kernelBruteForce
 * kernel.cu:486-515), which tests every ordered pair (i, j), i included,
 * and only counts them.  Every pair is tested with the join's distance
 * arithmetic for p->ilp / p->shortcircuit (ILP = 0 is the reference
 * kernel's sequential sum) and the predicate sqrt(acc) <= epsilon, in the
 * input's dimension order (REORDER does not apply, as in mode 1).  Writes
 * counts[n], offsets[n+1] (NULL = internal) and *total; with d_ids, the
 * rows, each ascending (GDSJ_EOVERFLOW when total > ids_cap).  The
 * all-pairs result differs from the grid join's only by pairs within
 * epsilon whose cells are not adjacent (rounding of the cell coordinates). */
int gdsj_bruteforce_device(gdsj_handle *h, const void *d_data, uint32_t n, uint32_t dims,
                           int dtype, double epsilon, const gdsj_params *p,
                           uint32_t *d_counts, uint64_t *d_offsets, uint32_t *d_ids,
                           uint64_t ids_cap, uint64_t *total, void *stream);
/* Host-buffer form: counts[n] on the host and the total (the CLI's mode 1). */
int gdsj_bruteforce_host(gdsj_handle *h, const void *data, uint32_t n, uint32_t dims,
                         int dtype, double epsilon, const gdsj_params *p,
                         uint32_t *counts, uint64_t *total);
/* The rows of the last gdsj_bruteforce_host on h (each ascending, in point
 * order) into the host buffer ids[cap], cap >= its total: the neighbor table
 * of the reference's CPU brute force (CPUBruteForceTable main.cu:953-978,
 * printed by search mode 0, main.cu:121-141). */
int gdsj_bruteforce_rows_host(gdsj_handle *h, uint32_t *ids, uint64_t cap);

/* Allocation-free all-pairs join.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It runs
 * gdsj_bruteforce_device's kernels with every scratch buffer (the SoA copy of
 * the points, the scan's partial sums) in one fixed device array of
 * GDSJ_NOALLOC_SCRATCH_BYTES that each device holds from the library's load
 * to the end of the process; it calls no allocating routine.  The caller
 * passes the points, counts[n], offsets[n+1] (required) and, to get the rows,
 * ids[ids_cap].  Inputs needing more scratch than the array holds
 * (about n * dims * sizeof(DTYPE) bytes) are rejected with GDSJ_EINVAL.
 * Runs on the current device and the given stream (NULL: the null stream)
 * and returns after the stream has drained.  Calls that overlap on one
 * device race on the shared array. */
#define GDSJ_NOALLOC_SCRATCH_BYTES (32u << 20)
int gdsj_bruteforce_device_noalloc(const void *d_data, uint32_t n, uint32_t dims,
                                   int dtype, double epsilon, const gdsj_params *p,
                                   uint32_t *d_counts, uint64_t *d_offsets, uint32_t *d_ids,
                                   uint64_t ids_cap, uint64_t *total, void *stream);

/* Largest x with sqrt_rn(x) <= eps, the threshold that replaces the
 * reference's per-pair square root (kernel.cu:361,364). */
float gdsj_eps_threshold_f32(float eps);
double gdsj_eps_threshold_f64(double eps);

/* ---- learned estimator ------------------------------------------------- */
/* The 16-32-32-1 MLP that sizes per-tile result slots (no counterpart in the
 * reference, which samples every 66th query: GPU.cu:57-268).  Its weights
 * can be saved and restored; the two host-buffer calls run its forward and
 * gradient kernels on the current device. */
uint32_t gdsj_mlp_param_count(void);
uint32_t gdsj_mlp_feature_count(void);
int gdsj_mlp_forward(const float *params, const float *x, uint32_t n, float *y);
/* grad = d/dparams mean_i (MLP(x_i) - t_i)^2, *loss = that mean */
int gdsj_mlp_grad(const float *params, const float *x, const float *t, uint32_t n, float *grad, float *loss);
int gdsj_model_get(gdsj_handle *h, float *params, uint32_t cap, double *alpha);
int gdsj_model_set(gdsj_handle *h, const float *params, uint32_t count, double alpha);

/* Thread-local text of the last error. */
const char *gdsj_last_error(void);

/* ---- dataset import ------------------------------------------------------ */
/* Replaces importNDDataset (import_dataset.cpp:49-95): the text is cut at
 * every ',' and each piece read with C++ formatted extraction into DTYPE
 * (decimal -> DTYPE directly) until an extraction fails.  *count = number of
 * values read; out (may be NULL with cap 0) receives the first cap of them.
 * Host only: needs no GPU. */
int gdsj_csv_load(const char *path, int dtype, void *out, uint64_t cap, uint64_t *count);

/* ---- batched joins ------------------------------------------------------ */
/* Many independent self-joins of uneven n, dims and epsilon run as one job:
 * each pass is one persistent launch over a device work queue of
 * (instance, tile) items, and a run is one hipGraph, captured on the first
 * run for a given output and replayed afterwards.  No counterpart in the
 * reference (one join per GDSJoinPy call, main.cu:360; the CLI joins one
 * file, main.cu:76-254): each instance's rows are exactly what
 * gdsj_join_device returns for that instance alone with the same params. */
#define GDSJ_ILP_AUTO 0xFFFFFFFFu  /* params.ilp for batches: the autotuner picks ILP 0 or 8;
                                      the choice is part of the result contract and is
                                      reported in gdsj_batch_tuning.ilp */

typedef struct gdsj_instance {
    const void *d_data;  /* n x dims row-major DTYPE, device memory (kept valid and unchanged) */
    uint32_t n;
    uint32_t dims;
    double epsilon;
} gdsj_instance;

typedef struct gdsj_batch_tuning {
    uint32_t tile_queries;    /* queries per work tile (16, 32, 64) */
    uint32_t waves_per_block; /* 1, 2 or 4 wave64s per workgroup */
    uint32_t ilp;             /* ILP in force */
    uint32_t lpt;             /* 1: queue hands out the heaviest tiles first */
    uint32_t blocks;          /* persistent grid size of the last counting pass */
    uint32_t reg_budget;      /* 24 or 48 coordinate words per lane of candidates in registers */
    double ms;                /* counting-pass time the autotuner measured for this choice */
} gdsj_batch_tuning;

typedef struct gdsj_batch_stats {
    uint64_t total_pairs;        /* all instances, self-pairs included */
    uint64_t candidates;         /* point comparisons */
    uint64_t n_points;           /* sum of n */
    uint64_t n_tiles;
    uint64_t est_predicted;      /* learned: predicted pairs */
    uint64_t est_capacity;       /* learned: result slots (pairs) */
    uint64_t est_overflow_tiles; /* learned: tiles completed by the exact fill */
    uint32_t slotted;            /* 1: the last run used learned slots */
    uint32_t graph_replays;      /* runs so far */
    uint32_t graph_captures;     /* graphs captured so far (one per output buffers and mode) */
    uint32_t reserved;
    double ms_build;             /* grids of every instance (set-up) */
    double ms_tune;              /* autotuner */
    double ms_train;             /* estimator training */
    double ms_last_run;          /* the last run's graph (device time) */
    gdsj_batch_tuning tuning;
} gdsj_batch_stats;

typedef struct gdsj_batch gdsj_batch;

/* Builds every instance's grid, runs and tiles on `device` (index
 * construction).  p is shared: k = 0 follows each instance's dims; ilp may
 * be GDSJ_ILP_AUTO; filter PLAIN or AUTO (batches use the plain kernels);
 * estimator EXACT, or LEARNED (the first run is exact and trains the MLP,
 * later runs are single slotted passes with the exact fill for overflowed
 * tiles). */
int gdsj_batch_create(int device, const gdsj_instance *inst, uint32_t ninst, int dtype,
                      const gdsj_params *p, gdsj_batch **out);
int gdsj_batch_destroy(gdsj_batch *b);
/* Times the counting pass for every tile size, block size, register budget
 * and queue order (and ILP 0/8 under GDSJ_ILP_AUTO) and keeps the fastest;
 * rebuilds the tiles when the tile size changes.  chosen may be NULL. */
int gdsj_batch_autotune(gdsj_batch *b, gdsj_batch_tuning *chosen);
/* Runs every join.  Row r of instance i is row rowBase_i + r with
 * rowBase_i = n_0 + ... + n_{i-1}: d_counts uint32[sum n], d_offsets
 * uint64[sum n + 1] (positions in d_ids), d_ids uint32[ids_cap] holding
 * instance-local point ids.  *total = all pairs; GDSJ_EOVERFLOW when
 * total > ids_cap (counts and offsets valid).  stream: hipStream_t or NULL;
 * returns after the stream drained. */
int gdsj_batch_run(gdsj_batch *b, uint32_t *d_counts, uint64_t *d_offsets, uint32_t *d_ids,
                   uint64_t ids_cap, uint64_t *total, void *stream);
int gdsj_batch_get_stats(gdsj_batch *b, gdsj_batch_stats *out);

/* ---- legacy drop-in symbols -------------------------------------------- */
/* main.cu:360.  DTYPE float in this symbol; GDSJoinPyDouble for double.
 * Dims and k come from the arguments and gdsj_set_legacy_params() instead
 * of compile-time GPUNUMDIM/NUMINDEXEDDIM.  On error counts are left 0 and
 * gdsj_last_error() says why (the reference prints and returns). */
void GDSJoinPy(float *dataset, unsigned int NUMPOINTS, float epsilon, unsigned int NDIM,
               int gpuContext, unsigned int *outNumNeighborsWithinEps);
void GDSJoinPyDouble(double *dataset, unsigned int NUMPOINTS, double epsilon,
                     unsigned int NDIM, int gpuContext, unsigned int *outNumNeighborsWithinEps);
/* main.cu:507: copies the pending table and frees it (a second call is a no-op). */
void copyResultIntoPythonArray(unsigned int *outPythonNeighborTable, unsigned int numResults);
/* Parameters the legacy symbols use (defaults: gdsj_params_default). */
int gdsj_set_legacy_params(const gdsj_params *p);
/* Status of the last legacy call (GDSJ_OK or an error code). */
int gdsj_legacy_status(void);

#ifdef __cplusplus
}
#endif

#endif /* GDSJOIN_H */
