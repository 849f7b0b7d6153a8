/*
 * sgtsne.h -- C-ABI of the MI355X (gfx950) SG-t-SNE engine.
 *
 * Plain C: int status codes, plain pointers and sizes, per-context state, one
 * HIP stream per context, no globals.  Host pointers unless a name says
 * "_dev".  Every entry point names the reference interface it replaces
 * (paths relative to the reference root: CU = code/cuda/SG-tSNE-CUDA/src,
 * HY = code/cuda/SG-tSNE-HYB/sgtsnepi/src).
 *
 * Errors: the reference aborts the process (gpuErrchk -> exit, CU/common.cuh:15-22;
 * exit(1) on an out-of-grid point, HY/nuconv.cpp:43-44).  Here every call
 * returns SGTSNE_OK or a negative code and sgtsne_last_error() describes it.
 */
#ifndef SGTSNE_H
#define SGTSNE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define SGTSNE_OK 0
#define SGTSNE_ERR_ARG (-1)       /* invalid argument / size              */
#define SGTSNE_ERR_HIP (-2)       /* HIP runtime error                     */
#define SGTSNE_ERR_NOMEM (-3)     /* device or host allocation failed      */
#define SGTSNE_ERR_NUMERIC (-4)   /* non-finite embedding (out of grid)    */
#define SGTSNE_ERR_STATE (-5)     /* call order (e.g. step before upload)  */
#define SGTSNE_ERR_NODEV (-6)     /* no gfx950 device visible              */
#define SGTSNE_ERR_FULL (-8)      /* no room: job slots, arena or grid budget
                                     (retry after a job leaves), or a job's
                                     grid outgrew its capacity               */

/* job states reported by sgtsne_batch_job_status (errors are negative) */
#define SGTSNE_JOB_RUNNING 0
#define SGTSNE_JOB_DONE 1

/* ---- sparse formats: the reference's numbering (CU/types.hpp:66,
 *      CU/pq.cu:25-57): 0 CSR, 1 BSR (stub in the reference, rejected
 *      here), 2 COO, 3 hybrid ELL+COO.  4 is this engine's own
 *      wavefront-segmented CSR (the default).                             */
#define SGTSNE_FORMAT_CSR 0
#define SGTSNE_FORMAT_COO 2
#define SGTSNE_FORMAT_HYBRID 3
#define SGTSNE_FORMAT_CSR_SEG 4

/* grid-size policy: h-driven (HY/qq.cpp:165-166, the varGrid semantics) or
 * the fixed getBestGridSize(ng) of the plain CUDA path (CU/gradient_descend.cu:139) */
#define SGTSNE_GRID_H 0
#define SGTSNE_GRID_FIXED 1

/* normalisation term Z: DIRECT = sum over points of the grid convolution of
 * the unit charges with 1/(1+r^2), minus N (no cancellation in FP32);
 * ALGEBRAIC = sum_i (1+2|y_i|^2) Phi0_i - 2 y_i . Phi_i - N from the squared
 * kernel alone, the reference's formula (HY/qq.cpp:75-106, CU/Frep.cu:43-80). */
#define SGTSNE_ZETA_DIRECT 0
#define SGTSNE_ZETA_ALGEBRAIC 1

/* Y layouts: SoA [d][N] (CU path) or AoS [N][d] (HYB path) */
#define SGTSNE_SPREAD_AUTO 0   /* matrix cores where they apply, else LDS-privatised
                                  grids, else fixed-point LDS windows (exact)    */
#define SGTSNE_SPREAD_ATOMIC 1 /* LDS-privatised grids, or global float atomics */
#define SGTSNE_SPREAD_MFMA 2   /* matrix cores (2-D side <= 32, 3-D side <= 16; else AUTO) */
#define SGTSNE_LAYOUT_SOA 0
#define SGTSNE_LAYOUT_AOS 1

/* Parameter set: every field of tsneparams (CU/types.hpp:55-72; HYB subset
 * HY/sgtsne.hpp:25-37) plus the engine's layout and grid-policy flags. */
typedef struct sgtsne_params {
  int d;              /* embedding dimensions 1..3                      (2)   */
  double lambda;      /* lambda rescaling; 1 skips it                   (1)   */
  double alpha;       /* early exaggeration multiplier                  (12)  */
  int max_iter;       /* gradient iterations                            (1000)*/
  int early_iter;     /* iterations 0..early_iter use alpha             (250) */
  int n;              /* number of points (filled from P when 0)              */
  double h;           /* grid side length; <= 0: 0.5/0.7/1.2 for d=1/2/3      */
  int drop_leaf;      /* drop edges of leaf rows during lambda rescaling      */
  int np;             /* CPU workers in the reference; unused on the GPU      */
  int format;         /* SGTSNE_FORMAT_*                        (CSR_SEG)     */
  int bs;             /* BSR block size in the reference; must be 0           */
  int compute_error;  /* > 0: record the relative L2 error of the interpolated
                         repulsion against the exact O(N^2) one every 10
                         iterations (CU/gradient_descend.cu:206-215)          */
  int ng;             /* grid request for SGTSNE_GRID_FIXED             (50)  */
  int grid_mode;      /* SGTSNE_GRID_*                          (H)           */
  int layout;         /* SGTSNE_LAYOUT_* of y_in / y_out        (SOA)         */
  uint64_t seed;      /* start when y_in is NULL: 0 = the reference sequence, else a seeded N(0, 1e-4^2) */
  int ell_width;      /* hybrid ELL width; <= 0 = automatic                   */
  int zeta_mode;      /* SGTSNE_ZETA_*                          (DIRECT)      */
  int reorder;        /* device order of the points (the API always reads and
                         writes the caller's order): 0 as given; 1 reverse
                         Cuthill-McKee order of P at upload; 2 also tries a
                         Morton order of Y at iterations 50, 150, 300, 600
                         and keeps it when the attractive kernel measures
                         faster (CSR formats)                            (2)   */
  int spread;         /* SGTSNE_SPREAD_*: how points are spread onto the grid
                         (AUTO: the matrix-core form for 2-D grids of side
                         <= 32 and 3-D grids of side <= 16, else
                         LDS-privatised, else fixed-point windows)  (AUTO)  */
} sgtsne_params;

/* Fills the defaults listed above.  Replaces the in-class defaults of
 * tsneparams (CU/types.hpp:55-72). */
int sgtsne_params_default(sgtsne_params *p);

/* Description of the last error raised on this thread (never NULL). */
const char *sgtsne_last_error(void);

/* Library build string (version, target). */
const char *sgtsne_version(void);

/* ---- P construction front end --------------------------------------- */

/* Exact all-kNN on the GPU: k+1 neighbours per point, self first, squared
 * FP64 distances summed in dimension order with fma, ties broken by the lower
 * index (bit-exact against a CPU computation of the same order).
 * X row-major [n][dim] float; I [n][k+1] int32; D [n][k+1] double.
 * Replaces allKNNsearchflann (CU/allknn.cuh:75-92), the approximate
 * FLANN search the reference calls from CU/testing_modules/demo_perplexity.cu:131. */
int sgtsne_knn_exact(const float *X, int n, int dim, int k, int32_t *I, double *D);

/* kNN methods.  Every method returns the rows of sgtsne_knn_exact bit for bit.
 * EXACT_FP64: FP64 distances for every pair.  FILTERED: FP32 matrix-core
 * distances of the mean-centred points select k + 16 candidates per point; a
 * rigorous bound on their error certifies that the exact k nearest are among
 * them, and the candidates' FP64 distances are recomputed exactly as
 * EXACT_FP64 does; rows that cannot be certified (ties or near-ties wider than
 * the 16 spare candidates) are recomputed with EXACT_FP64.  FILTERED needs
 * k <= 240, dim <= 128 and n >= 64; AUTO takes it when it applies. */
#define SGTSNE_KNN_AUTO 0
#define SGTSNE_KNN_EXACT_FP64 1
#define SGTSNE_KNN_FILTERED 2

/* sgtsne_knn_exact with a chosen method; *fallback_rows (may be NULL)
 * receives the rows FILTERED recomputed exactly (-1 when it did not run).
 * No counterpart in the reference; sgtsne_knn_exact uses AUTO. */
int sgtsne_knn(const float *X, int n, int dim, int k, int method, int32_t *I, double *D, int *fallback_rows);

/* Perplexity equalisation on the GPU (FP64 bisection, |H - log u| < 1e-5,
 * <= 200 steps, self column dropped, rows normalised).  I, D as above with
 * k = nn.  Output CSR: rowptr [n+1] (= k*i), col [n*k], val [n*k] float.
 * Replaces perplexityEqualization (CU/perplexity_search.cuh:18-20,
 * CU/perplexity_search.cu:114-141; FP64 twin HY/sgtsne.cpp:257-297). */
int sgtsne_perplexity_equalize(const int32_t *I, const double *D, int n, int k, double u,
                               int32_t *rowptr, int32_t *col, float *val);

/* Non-periodic convolution of an ng^d grid with d+1 channels by the kernel
 * (1 + h^2 r^2)^-2, r in grid steps, on the GPU: the G2G step of the
 * repulsion (HY/non_periodic_conv.cpp:26-708, CU/non_periodic_convF.cu:162-517).
 * V is a host array [d+1][ng^d], x fastest; Phi receives [d+1][ng^d] and,
 * when zdirect != 0, a channel d+1 with channel 0 convolved by
 * (1 + h^2 r^2)^-1.  4 <= ng <= 512. */
int sgtsne_grid_conv(int d, int ng, double h, int zdirect, const float *V, float *Phi);

/* sgtsne_grid_conv without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  V and Phi
 * are DEVICE arrays of the caller (same layouts as sgtsne_grid_conv);
 * `stream` is a hipStream_t (NULL: the null stream) and the call returns once
 * the work is queued on it.  The FFT workspace comes from one 32 MiB device
 * array of the library per GPU, held until the process ends, which fits every
 * 1-D and 2-D grid (ng <= 512) and 3-D grids of side ng <= 64; larger grids
 * return SGTSNE_ERR_ARG.  Every call on a GPU shares that array, so calls
 * that overlap on one GPU (other streams or threads) race: order them. */
int sgtsne_grid_conv_noalloc(int d, int ng, double h, int zdirect, const float *V, float *Phi, void *stream);

/* ---- drop-in entry -------------------------------------------------- */

/* Runs the whole method: lambda rescaling (if lambda != 1), P + P^T and
 * normalisation to sum 1, the sparse format build, initialisation (y_in or a
 * seeded N(0, 1e-4^2) start), params.max_iter iterations, mean-centring.
 * P is the row-stochastic kNN CSR (rowptr [n+1], col/val [nnz]).
 * y_in / y_out use params.layout; y_in may be NULL; time_info may be NULL,
 * else receives 7 doubles per iteration (ms: attractive, S2G, G2G, G2S,
 * forces, repulsion total, preparation) followed by the total loop time
 * (ms) at index 7*max_iter.
 * Replaces sgtsneCUDA (CU/sgtsne.cuh:13-18, CU/sgtsne.cu:26-128) and
 * sgtsne (HY/sgtsne.hpp:68-72, HY/sgtsne.cpp:37-197). */
int sgtsne_run(int n, int64_t nnz, const int32_t *rowptr, const int32_t *col, const float *val,
               const sgtsne_params *params, const float *y_in, float *y_out, double *time_info);

/* ---- stepwise context API (parity, batching, serving) ----------------- */

typedef struct sgtsne_ctx sgtsne_ctx;

/* Creates a context on the current device (own stream, own buffers).
 * Replaces the allocation block of kl_minimization (CU/gradient_descend.cu:125-188). */
int sgtsne_ctx_create(sgtsne_ctx **ctx, const sgtsne_params *params);

/* Uploads P (CSR, host memory).  symmetrize != 0: lambda rescaling (if
 * params.lambda != 1), P_sym = (P + P^T) / sum and the format build run on the
 * GPU (CU/sgtsne.cu:50-93, CU/graph_rescaling.cu:266-398).  symmetrize == 0:
 * P is taken as the final symmetric, normalised matrix. */
int sgtsne_ctx_upload_P(sgtsne_ctx *ctx, int n, int64_t nnz, const int32_t *rowptr,
                        const int32_t *col, const float *val, int symmetrize);

/* Downloads the device P in CSR (after symmetrisation).  Call with col ==
 * NULL to query nnz into *nnz. */
int sgtsne_ctx_get_P(sgtsne_ctx *ctx, int64_t *nnz, int32_t *rowptr, int32_t *col, float *val);

/* Sets / reads Y (layout per argument), and resets the optimiser state
 * (uy = 0, gains = 1, iteration counter = 0) when reset_state != 0. */
int sgtsne_ctx_set_Y(sgtsne_ctx *ctx, const float *y, int layout, int reset_state);
int sgtsne_ctx_get_Y(sgtsne_ctx *ctx, float *y, int layout);

/* Runs `iters` iterations of the schedule from the context's iteration
 * counter (CU/gradient_descend.cu:193-234; HY/gradient_descendHYB.cpp:203-244). */
int sgtsne_ctx_step(sgtsne_ctx *ctx, int iters);

/* Evaluates, for the current Y and without updating it, the per-iteration
 * terms the oracle compares: Fattr (attractive, without alpha), Frep (repulsive,
 * normalised by Z), Z, and the grid side ng (= nGrid + 2).  Fattr/Frep in
 * `layout`; any output pointer may be NULL.
 * Replaces compute_gradient (CU/gradient_descend.cu:53-101). */
int sgtsne_ctx_terms(sgtsne_ctx *ctx, float *Fattr, float *Frep, double *Z, int *ng, int layout);

/* KL divergence of the current Y with exaggeration `alpha`
 * (CU/pq.cu:61-125, tsneCost). */
int sgtsne_ctx_kl(sgtsne_ctx *ctx, double alpha, double *kl);

/* Exact O(N^2) repulsion on the GPU (validation; CU/compute_error.cu:8-53). */
int sgtsne_ctx_repulsive_exact(sgtsne_ctx *ctx, float *Frep, double *Z, int layout);

/* Per-phase device time of the last sgtsne_ctx_step call, ms summed over
 * its iterations: [0] attractive, [1] S2G (incl. binning), [2] G2G (FFT
 * convolution), [3] G2S + forces, [4] update, [5] total, [6] iterations. */
int sgtsne_ctx_timing(sgtsne_ctx *ctx, double *out7);

/* Choice of the attractive-kernel autotuner (run at upload for the default
 * format, and again after the first Morton re-order): lanes per row, unroll
 * (0 = plain form), rows per work-group (0 = grid-stride form), points of Y
 * staged in LDS per work-group (0 = none; -1 = column-localised blocks: Y at
 * each row block's distinct columns staged in LDS; -2^k, k = 6..16 =
 * chunk-cached blocks: per row block, the chunks of 2^k consecutive points its
 * neighbours fall in most, 128 KB of them, staged in LDS), and the measured ms
 * of each candidate (up to max; *count receives how many were timed). */
int sgtsne_ctx_tuning(sgtsne_ctx *ctx, int *group, int *unroll, int *rows_per_block, int *window, double *ms,
                      int max, int *count);

/* Pins the attractive kernel of the CSR_SEG format to one shape (same meaning
 * as sgtsne_ctx_tuning's outputs) and stops re-tuning: shapes with the same
 * lanes per row sum every row in the same order, so pinning the lanes makes
 * runs bit-reproducible across machines whose autotuner would choose
 * differently.  Call after upload_P; an unsupported shape returns
 * SGTSNE_ERR_ARG and leaves the current one.  No reference counterpart (the
 * reference's kernel is fixed by its format, CU/pq.cu:25-57). */
int sgtsne_ctx_set_attractive(sgtsne_ctx *ctx, int group, int unroll, int rows_per_block, int window);

/* Morton re-orders of the points applied so far, and those tried but
 * declined because the attractive kernel measured slower on them. */
int sgtsne_ctx_reorders(sgtsne_ctx *ctx, int *applied, int *declined);

/* Enables (default) or disables the per-phase event timing of ctx_step. */
int sgtsne_ctx_set_timing(sgtsne_ctx *ctx, int on);

/* Relative L2 errors recorded by params.compute_error > 0 (one every 10
 * iterations); *count receives how many exist, up to max are copied. */
int sgtsne_ctx_errors(sgtsne_ctx *ctx, double *out, int max, int *count);

/* Current iteration counter, n, d, nnz. */
int sgtsne_ctx_info(sgtsne_ctx *ctx, int *iter, int *n, int *d, int64_t *nnz);

int sgtsne_ctx_destroy(sgtsne_ctx *ctx);

/* ---- batched engine ------------------------------------------------------
 * Many independent embedding jobs of uneven size (same params.d) advanced
 * together: one launch per phase for all jobs, iterations replayed from a
 * captured HIP graph, P, Y and the grids resident on the device.  Each job
 * runs the plain method's schedule (CU/gradient_descend.cu:102-245) from its
 * own admission; its result does not depend on which other jobs share the
 * batch (bit-identical alone and in any mix).  The reference has no batched
 * form: it runs one sgtsneCUDA call (CU/sgtsne.cu:26-128) per embedding. */
typedef struct sgtsne_batch sgtsne_batch;

/* ---- files and start data of the reference's command-line programs ----
 * Host only: none of these touches the device. */

/* Frees an array returned by the readers below. */
void sgtsne_free(void *p);

/* Point file: readXfromMTX (CU/utils.cu:216-256).  Text, a header "n dim"
 * then n*dim values, point after point, separated by any whitespace.  A
 * positive *n on entry reads only the first *n points (the reference's -n
 * option).  *X receives a malloc'ed float [n][dim] (free with sgtsne_free);
 * *n and *dim receive the sizes read. */
int sgtsne_read_X_text(const char *path, int *n, int *dim, float **X);

/* Matrix Market graph: buildPFromMTX (CU/utils.cu:335-409, HY/utils.cpp:
 * 104-170).  Lines starting with '%' are skipped, then "m n nnz" and nnz
 * 1-based "row col value" triplets.  Returns P in compressed columns, as the
 * reference does: colptr [n+1], row [nnz] (0-based) and val [nnz], entries of
 * a column in file order.  Read as compressed rows, the same arrays hold P^T,
 * which is what the drop-in sgtsne_run takes for a column-stochastic P (the
 * HYB convention).  Arrays are malloc'ed (free with sgtsne_free). */
int sgtsne_read_P_mtx(const char *path, int *m, int *n, int64_t *nnz, int32_t **colptr, int32_t **row,
                      double **val);

/* Embedding text: extractEmbeddingTextT (CU/utils.cu:121-138, y SoA [d][N])
 * and extractEmbeddingText (HY/utils.cpp:180-196, y AoS [N][d]).  One line
 * per point, each coordinate written as std::ostream writes it by default
 * (printf "%.6g") followed by one space. */
int sgtsne_write_embedding_text(const char *path, const float *y, int n, int d, int layout);
int sgtsne_write_embedding_text_f64(const char *path, const double *y, int n, int d, int layout);

/* Embedding binary: extractEmbedding (HY/utils.cpp:199-206), n*d doubles
 * row-major. */
int sgtsne_write_embedding_bin(const char *path, const double *y, int n, int d);

/* Timing report: ExtractTimeInfo (CU/utils.cu:72-102) for a time_info array
 * of 7*iters+3 doubles (sgtsne_run fills 7 per iteration and the loop time at
 * 7*iters; the caller adds the matrix and total times at 7*iters+1, +2).
 * As in the reference, the report is appended only when the file already
 * exists; otherwise an empty file is created. */
int sgtsne_write_time_info(const char *path, const double *time_info, int n, int d, int iters);

/* The reference's synthetic points for its -s option: loadSimulationData
 * (CU/utils.cu:310-332), n_clusters groups of consecutive points whose
 * coordinate a is drawn from N(20, 2) if a >= the group index, else from
 * N(-20, 2), with one default-seeded std::default_random_engine.  X is
 * float [n][dim]. */
int sgtsne_simulation_data(int n, int dim, int n_clusters, float *X);

/* The reference's start embedding for y_in == NULL: generateRandomGaussianCoord
 * (CU/utils.cu:185-192), 1e-4 * N(0, 1) from a default-seeded
 * std::default_random_engine, drawn in [d][N] order.  sgtsne_run uses it when
 * params.seed is 0. */
int sgtsne_random_start(int n, int d, int layout, float *y);

/* The locality order the engine applies at upload (reorder >= 1): reverse
 * Cuthill-McKee of the graph of P (rowptr [n+1], col [nnz], host),
 * perm[r] = the vertex placed at row r.  Host only; no device work.  The
 * reference reorders with METIS nested dissection for its CSB blocks
 * (HY/sgtsne.cpp:90-112); this replaces that step. */
int sgtsne_locality_order(int n, const int32_t *rowptr, const int32_t *col, int32_t *perm);

/* Device bandwidth probe: copies `bytes` (a multiple of 16) between two
 * device buffers with a plain streaming kernel `reps` times and returns the
 * ms of one copy.  Calibrates HBM counters and the roofline; no reference
 * counterpart. */
int sgtsne_copy_probe(int64_t bytes, int reps, double *ms_per_copy);

/* Capacity: job slots (<= 1024), points and P_sym nonzeros of the arena, and
 * the byte budget of the per-job grid buffers (<= 0: 64 GiB). */
int sgtsne_batch_create(sgtsne_batch **batch, const sgtsne_params *params, int max_jobs, int64_t max_points,
                        int64_t max_nnz, int64_t grid_budget_bytes);
int sgtsne_batch_destroy(sgtsne_batch *batch);

/* Admits a job given its row-stochastic kNN P (CSR, host; symmetrize: form
 * (P + P^T) / (2N) as sgtsneCUDA does, CU/sgtsne.cu:58-76) and y_in (layout
 * params.layout, or NULL for the seeded N(0, 1e-4^2) start).  max_iter <= 0:
 * params.max_iter; ng_cap <= 0: automatic grid capacity.  Returns
 * SGTSNE_ERR_FULL when the job does not fit now.  The job starts iterating
 * with the next sgtsne_batch_step. */
int sgtsne_batch_add(sgtsne_batch *batch, int n, int64_t nnz, const int32_t *rowptr, const int32_t *col,
                     const float *val, int symmetrize, const float *y_in, int max_iter, int ng_cap, int *job);

/* Admits a job from its points X [n][dim]: exact kNN (k), perplexity u and
 * P + P^T on the device (perplexityEqualization, CU/perplexity_search.cu:114-141,
 * then as sgtsne_batch_add).  Runs on a stream of its own, so it overlaps
 * iterations already queued by sgtsne_batch_step. */
int sgtsne_batch_add_points(sgtsne_batch *batch, const float *X, int n, int dim, int k, double u, const float *y_in,
                            int max_iter, int ng_cap, int *job);

/* sgtsne_batch_add_points without holding the host: _async queues the first
 * part of the front end (the kNN up to its certificate) on the batch's
 * front-end stream and returns a ticket at once; each sgtsne_batch_add_poll
 * (wait = 0) returns at once with *job = -1 until the device has passed the
 * queued part, then queues the next one (P and the merge of P and P^T, then
 * P_sym's CSR), and after the last one admits the job and sets *job.  wait = 1
 * runs the remaining parts to the admission.  Same kernels in the same order
 * as sgtsne_batch_add_points, so the same job bits.  Up to 8 front ends per
 * batch in flight (SGTSNE_ERR_STATE beyond): their parts queue one behind
 * another on the front-end stream, so the device runs one while the host takes
 * up the next.  X and y_in must stay valid and unchanged until the poll that
 * admits the job or returns an error.  While a front end is in flight, reads
 * of finished jobs (sgtsne_batch_get_Y, sgtsne_batch_job_status) wait for its
 * queued part, as they share its stream.  A server uses this to keep queuing
 * iterations while a job's kNN runs.  No counterpart in the reference. */
int sgtsne_batch_add_points_async(sgtsne_batch *batch, const float *X, int n, int dim, int k, double u,
                                  const float *y_in, int max_iter, int ng_cap, int64_t *ticket);
int sgtsne_batch_add_poll(sgtsne_batch *batch, int64_t ticket, int wait, int *job);

/* Queues iters batch iterations (asynchronous; returns before they finish).
 * A newly admitted job spends its first batch iteration centring Y and
 * sizing its grid, then advances one iteration per batch iteration until
 * its max_iter. */
int sgtsne_batch_step(sgtsne_batch *batch, int iters);
int sgtsne_batch_sync(sgtsne_batch *batch);

/* Pacing of queued iterations: sgtsne_batch_mark records a point on the
 * batch's iteration stream (after everything queued so far) and returns its
 * number; sgtsne_batch_wait(batch, mark) returns once the device has passed
 * it (or a later mark).  A server keeps a few steps in flight with these
 * instead of draining the stream.  No counterpart in the reference. */
int sgtsne_batch_mark(sgtsne_batch *batch, int64_t *mark);
int sgtsne_batch_wait(sgtsne_batch *batch, int64_t mark);
/* Non-blocking form of sgtsne_batch_wait: *passed = 1 once the device has
 * passed the mark, else 0.  No counterpart in the reference. */
int sgtsne_batch_passed(sgtsne_batch *batch, int64_t mark, int *passed);

/* Waits for queued work, then reports SGTSNE_JOB_RUNNING / SGTSNE_JOB_DONE
 * or a negative error of the job and its iteration counter. */
int sgtsne_batch_job_status(sgtsne_batch *batch, int job, int *status, int *iter);
/* Batch iterations the job still needs before its last one is queued (host
 * bookkeeping, no device access): max_iter + 1 right after admission, 0 once
 * a step has queued its last iteration.  No counterpart in the reference. */
int sgtsne_batch_job_left(sgtsne_batch *batch, int job, int64_t *left);

/* Threads: call a batch's functions from one host thread at a time. */
int sgtsne_batch_get_Y(sgtsne_batch *batch, int job, float *y, int layout);

/* The job's attractive term, repulsive term (normalised by Z) and Z at its
 * current Y, without stepping (same meaning as sgtsne_ctx_terms). */
int sgtsne_batch_terms(sgtsne_batch *batch, int job, float *Fattr, float *Frep, double *Z, int *ng, int layout);

/* Frees the job's slot and grid buffers; its rows are reclaimed by a later
 * compaction of the arena. */
int sgtsne_batch_remove(sgtsne_batch *batch, int job);

/* Live jobs, arena points and nonzeros in use, grid bytes in use, graph replays. */
int sgtsne_batch_info(sgtsne_batch *batch, int *jobs, int64_t *points, int64_t *nnz, int64_t *grid_bytes,
                      int64_t *graph_replays);

/* Attractive-kernel shape chosen by the batch autotuner (lanes per row fixed
 * at 16, so the choice never changes a result) and whether iterations are
 * graph-replayed. */
int sgtsne_batch_tuning(sgtsne_batch *batch, int *group, int *unroll, int *rows_per_block, int *window,
                        int *use_graph);

/* Pins the batch's attractive-kernel shape (unroll, rows per work-group,
 * points of Y staged in LDS; the same meaning as sgtsne_batch_tuning's
 * outputs) and stops re-tuning.  The lanes per row stay the engine's, so the
 * pin never changes a result; it lets a caller run, test or profile one
 * shape.  An unsupported shape returns SGTSNE_ERR_ARG and keeps the current
 * one.  No reference counterpart (the reference has no batched form). */
int sgtsne_batch_set_attractive(sgtsne_batch *batch, int unroll, int rows_per_block, int window);

/* Graph replay on (default) or off (iterations launched one by one). */
int sgtsne_batch_set_graph(sgtsne_batch *batch, int on);

#ifdef __cplusplus
}
#endif

#endif /* SGTSNE_H */
