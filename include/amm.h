/*
 * amm.h -- C ABI of the MI355X (gfx950) AMM power method for the largest
 * eigenvalue of A*A^T.
 *
 * A is d x n, column-major (column j at A + ld*j), FP32 or BF16, resident in
 * device memory and owned by the caller.  Every function returns 0 on
 * success or a negative AMM_E* code; amm_last_error() then holds a message
 * (per thread).  No function throws.  A context belongs to one stream and is
 * not shared between threads.
 *
 * Each entry point names the reference interface it replaces
 * (Takuya-Okuyama/eigenvalueComputation_by_AMM, file:line).  The reference
 * has no FFI; these are the operator boundaries of its hot path.
 */
#ifndef AMM_H_
#define AMM_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMM_ABI_VERSION 1

/* status codes */
#define AMM_OK 0
#define AMM_E_INVALID (-1)   /* bad argument */
#define AMM_E_HIP (-2)       /* HIP runtime error */
#define AMM_E_NOMEM (-3)     /* device allocation failed */
#define AMM_E_STATE (-4)     /* call out of order */

/* element types */
#define AMM_DTYPE_F32 0
#define AMM_DTYPE_BF16 1

/* kernels (the -kernel flag, host_memory.h:132-141) */
#define AMM_KERNEL_EXACT 0
#define AMM_KERNEL_PREVIOUS 1
#define AMM_KERNEL_PROPOSED 2
#define AMM_KERNEL_PROPOSED_V2 3

/* entry distributions of amm_generate_symmetric (-matrixType, device_memory.h:405-421) */
#define AMM_DIST_GAUSSIAN 0
#define AMM_DIST_LOGNORMAL 1
#define AMM_DIST_UNIFORM 2

/* flags of amm_params.flags */
#define AMM_FLAG_NO_GRAPH 1u     /* launch kernel by kernel instead of one hipGraph */
#define AMM_FLAG_KEEP_STATS 2u   /* reuse column statistics of an earlier run */
/* variance reduction off: the reference's variance_reduction = false, which
 * the last of -vr / -vr_adaptive decides (host_memory.h:108-120) and which
 * gates every variance-reduction exact step (powerMethod_by_previous.cuh:
 * 135-138, powerMethod_by_proposed.cuh:159-161, powerMethod_by_proposed_v2.cuh:
 * 305); proposedAMM_v2 still takes its transition step at vr_period
 * (powerMethod_by_proposed_v2.cuh:234-277). */
#define AMM_FLAG_NO_VR 4u

typedef struct amm_ctx amm_ctx;

/* Run parameters: the reference's host_memory fields (host_memory.h:26-50). */
typedef struct amm_params {
  int32_t kernel;               /* AMM_KERNEL_* */
  uint32_t c;                   /* sampled columns per product (ignored by exact) */
  uint64_t step;                /* power iterations */
  uint64_t vr_period;           /* -vr N: exact step every N iterations (0: off) */
  double vr_adaptive_threshold; /* -vr_adaptive thr: exact step when cos > 1 - thr */
  uint64_t seed;                /* Philox key */
  uint32_t trial;               /* Philox counter word: repetition / trial index */
  int32_t sanity;               /* -sanity: with c == n every product is exact */
  uint32_t flags;               /* AMM_FLAG_* */
  uint32_t reserved;
} amm_params;

/* Message of the last failing call on this thread ("" if none). */
const char *amm_last_error(void);
int amm_abi_version(void);

/* Create a context over a device-resident matrix.  stream: a hipStream_t
 * (NULL: the context creates its own).  Replaces the constructor of
 * struct device_memory (device_memory.h:172-276).  Requirements: d, n >= 1,
 * ld >= d, A 16-byte aligned, ld a multiple of 4 (FP32) / 8 (BF16). */
int amm_ctx_create(const void *A, int32_t dtype, uint64_t d, uint64_t n, uint64_t ld,
                   void *stream, amm_ctx **out);
int amm_ctx_destroy(amm_ctx *ctx);

/* Reference eigenvector v* (host, FP64, length d) for NormError
 * err = 1 - (y^T v* / ||y||)^2; replaces the read of data/exact_eigenvector
 * (device_memory.h:246-254).  NULL clears it (err then reads 1). */
int amm_set_exact_eigenvector(amm_ctx *ctx, const double *v_host, uint64_t len);

/* Column statistics alpha_j (mean), ||a_j||^2, ||a_j - alpha_j 1||^2.
 * Replaces get_norms (device_memory.h:34-59) and cublasSnrm2 of alpha
 * (device_memory.h:335-336).  Outputs are host arrays of length n (each may
 * be NULL); alpha2_host receives ||alpha||^2 (may be NULL). */
int amm_column_stats(amm_ctx *ctx, float *alpha_host, double *nrm2_host, double *cnrm2_host,
                     double *alpha2_host);

/* Sampling CDF over w_j = ||a_j||^2 (centred = 0) or ||a_j - alpha_j 1||^2
 * (centred = 1), exact in integers: q_j = floor(w_j 2^(62 - ceil(log2 n)) /
 * max w), cdf_j = q_0 + ... + q_j.  Replaces set_weights +
 * cub::DeviceScan::InclusiveSum (common.h:34-40,
 * powerMethod_by_previous.cuh:26-32).  Host outputs. */
int amm_build_cdf(amm_ctx *ctx, int32_t centred, uint64_t *cdf_host, uint64_t *total_host,
                  uint64_t *last_pos_host);

/* Draw the c samples of iteration `iter` and merge duplicates.  Replaces
 * pick_index + cub::DeviceRadixSort::SortKeys + update_weights
 * (device_memory.h:519-545, powerMethod_by_previous.cuh:56-70,
 * common.h:42-53).  Segment g (slots [8192 g, 8192 g + 8192)) writes its
 * unique columns, ascending, at idx_host[8192 g ...] with weights
 * mult * total / (c * q_j), q_j = cdf_j - cdf_{j-1}, and its count in
 * seg_count_host[g]. */
int amm_sample(amm_ctx *ctx, int32_t centred, uint64_t seed, uint32_t iter, uint32_t trial,
               uint32_t c, uint32_t *idx_host, float *wgt_host, uint32_t *seg_count_host);

/* One sampled product y = AMM(A A^T)(x - x_memo) + y_memo (previous) or the
 * mean-shifted form (proposed), with iteration `iter`'s samples.  Device
 * pointers, length d (x_memo, y_memo may be NULL).  Replaces
 * square_of_A_by_previous / square_of_A_by_proposed
 * (powerMethod_by_previous.cuh:7-98, powerMethod_by_proposed.cuh:7-103). */
int amm_square_of_A(amm_ctx *ctx, int32_t kernel, int32_t sanity, uint32_t c, uint64_t seed,
                    uint32_t iter, uint32_t trial, const float *x_dev, const float *x_memo_dev,
                    const float *y_memo_dev, float *y_dev);

/* Exact y = A (A^T x), device pointers of length d.  Replaces the two
 * cublasSgemv calls (powerMethod_by_exact.cuh:37-53). */
int amm_exact_product(amm_ctx *ctx, const float *x_dev, float *y_dev);

/* The same exact product y = A (A^T x) with no device allocation and no
 * context.  No counterpart in the reference: for callers that cannot
 * allocate (code under hipGraph capture, serving loops, harnesses that
 * forbid hipMalloc).  A as for amm_ctx_create; x_dev, y_dev: device, length
 * d.  Every intermediate lives in one 16 MiB device array per GPU, held
 * until the process ends; sizes whose intermediates do not fit are rejected
 * with AMM_E_INVALID.  Asynchronous on `stream` (a hipStream_t, NULL: the
 * null stream) on the current device.  Calls that overlap on one GPU share
 * that array and race: order them on one stream. */
int amm_exact_product_noalloc(const void *A, int32_t dtype, uint64_t d, uint64_t n, uint64_t ld,
                              const float *x_dev, float *y_dev, void *stream);

/* A full power-method run; host outputs of length p->step: cumulative device
 * time in ms from the run's first kernel, raw Rayleigh quotient
 * lambda_i = y^T x, and y^T v* / ||y||.  Replaces powerMethod_by_exact /
 * _previous / _proposed / _proposed_v2 (powerMethod_by_exact.cuh:4-100,
 * powerMethod_by_previous.cuh:100-210, powerMethod_by_proposed.cuh:105-235,
 * powerMethod_by_proposed_v2.cuh:199-377).  Any output may be NULL. */
int amm_power_method(amm_ctx *ctx, const amm_params *p, double *time_ms, double *eig,
                     double *err);

/* Step kinds of the last run (1 exact+memo, 2 previous, 3 proposed, 4 exact),
 * host array of length step. */
int amm_last_modes(amm_ctx *ctx, uint32_t *modes_host, uint64_t len);

/* Copy the current iterate x (which = 0) or the last product y (which = 1)
 * to host (length d). */
int amm_get_vector(amm_ctx *ctx, int32_t which, float *host, uint64_t len);

/* Wall time of the last amm_power_method, from HIP events around it (ms). */
int amm_last_run_ms(amm_ctx *ctx, double *ms);

/* Measurement helper (the reference times whole iterations only,
 * powerMethod_by_exact.cuh:64-69): launch the two contraction kernels of one
 * product `iters` times each on the context's stream, bracketed by HIP
 * events, and report the mean duration of each (ms) and the number of
 * distinct columns one launch reads (n for the exact product).
 * kernel: AMM_KERNEL_EXACT, _PREVIOUS or _PROPOSED; samples of iteration 0
 * of `seed`. */
int amm_time_product(amm_ctx *ctx, int32_t kernel, uint32_t c, uint64_t seed, uint32_t iters,
                     double *ms_dot, double *ms_axpy, uint64_t *columns);

/* Measurement helper: mean duration (ms) of the column-statistics pass
 * (one read of A) over `iters` launches, HIP events on the context stream. */
int amm_time_column_stats(amm_ctx *ctx, uint32_t iters, double *ms);

/* Measurement helper: mean duration (ms) over `iters` launches of one pass
 * as a run issues it (HIP events on the context stream).
 *   AMM_PASS_STATS_Z: column statistics with z = A alpha in the same read
 *     (get_norms, device_memory.h:34-59, plus powerMethod_by_proposed.cuh:
 *     132-139), the kernel and the partial sum; d <= 16384 FP32 / 32768 BF16.
 *   AMM_PASS_EXACT: the exact product y = A (A^T x) (the two cublasSgemv
 *     calls of powerMethod_by_exact.cuh:46-62) without its finish: one read
 *     of A for d <= 16384, else the two contraction kernels.
 *   AMM_PASS_READ: no counterpart in the reference; a plain streaming read
 *     of the matrix (every byte up to its last column's d-th row, 16-byte
 *     non-temporal loads, 1 MiB per workgroup) and nothing else: the read
 *     ceiling the passes above are measured against.
 * AMM_E_INVALID when the pass does not apply to the context's shape. */
enum { AMM_PASS_STATS_Z = 1, AMM_PASS_EXACT = 2, AMM_PASS_READ = 3 };
int amm_time_pass(amm_ctx *ctx, int32_t pass, uint32_t iters, double *ms);

/* Fill the m x m FP32 device matrix A (leading dimension ld) with a
 * symmetric random matrix -- entries gaussian(mean p1, std p2),
 * lognormal(p1, p2) or uniform[p1, p2) -- and add the Gershgorin offset
 * delta = max_j max(0, sum_i |a_ij| - 2 |a_jj|) to its diagonal; delta is
 * returned in *diagonal_offset.  Replaces generate_matrix's synthetic branch
 * (device_memory.h:402-431: curandGenerate*, convert_to_symmetry,
 * find_offset, cublasIsamax, add_offset).  Synchronises `stream`. */
int amm_generate_symmetric(void *A_dev, uint64_t m, uint64_t ld, int32_t dist, float p1, float p2,
                           uint64_t seed, void *stream, float *diagonal_offset);

/* sigma_max(A) with its singular vectors (no reference counterpart: the
 * reference reports sqrt(lambda) only, powerMethod_by_exact.cuh:83): runs
 * the power method of p, takes its final iterate u (device, length d), and
 * one exact pass t = A^T u gives sigma = ||t|| (FP64 norm, host) and
 * v = t / sigma (device, length n).  At convergence d sigma / d A = u v^T. */
int amm_sigma_max(amm_ctx *ctx, const amm_params *p, double *sigma, float *u_dev, float *v_dev);

/* out = g * u v^T for a column-major d x n device matrix (leading dimension
 * ld, FP32 or BF16): the backward of sigma_max.  Asynchronous on stream. */
int amm_rank1_update(void *out_dev, int32_t dtype, uint64_t d, uint64_t n, uint64_t ld, float g,
                     const float *u_dev, const float *v_dev, void *stream);

/* The same for a strided stack of `count` matrices in one launch (no
 * reference counterpart; the backward of a batched sigma_max): matrix k at
 * out_dev + k out_stride elements gets g_dev[k] u_k v_k^T with u_k = u_dev +
 * k u_stride, v_k = v_dev + k v_stride (g_dev: count FP32 values on the
 * device).  1 <= count < 65536.  Asynchronous on stream. */
int amm_rank1_update_batched(void *out_dev, int32_t dtype, uint64_t d, uint64_t n, uint64_t ld, uint64_t count,
                             uint64_t out_stride, const float *g_dev, const float *u_dev, uint64_t u_stride,
                             const float *v_dev, uint64_t v_stride, void *stream);

/* ---------------------------------------------------------------------------
 * Batched form (no reference counterpart: the reference solves one matrix
 * with one right-hand side per run, AMM.cu:7-72).  Many matrices of uneven
 * sizes, each with `trials` (<= 16) sampling seeds stacked as one block of
 * right-hand sides, so that both contractions of a product run as skinny
 * GEMMs on the matrix cores over the union of the trials' samples.  Trial k
 * of problem i reproduces amm_power_method on that matrix with p->seed =
 * problem.seed and p->trial = p->trial + k: the same sampled columns,
 * multiplicities and weights in every iteration (the summation order of
 * the products differs).  Every launch covers all problems; the whole run
 * is one hipGraph.
 * ------------------------------------------------------------------------- */
typedef struct amm_problem {
  const void *A;   /* device matrix, d x n column-major, leading dimension ld */
  int32_t dtype;   /* AMM_DTYPE_F32 or AMM_DTYPE_BF16 */
  uint32_t c;      /* sampled columns per trial and product, <= min(n, 65535) (0: exact runs only) */
  uint64_t d, n, ld;
  uint64_t seed;   /* Philox key of this problem */
} amm_problem;

typedef struct amm_batch amm_batch;

int amm_batch_create(const amm_problem *problems, uint32_t count, uint32_t trials, void *stream,
                     amm_batch **out);
int amm_batch_destroy(amm_batch *b);

/* Reference eigenvector of one problem for the err output (NULL clears). */
int amm_batch_set_exact_eigenvector(amm_batch *b, uint32_t problem, const double *v_host,
                                    uint64_t len);

/* One batched run (kernel, step, vr_period, vr_adaptive_threshold, trial,
 * flags of p; c and seed per problem; sanity mode is not supported).
 * time_ms: host, length step (device time from the run's first kernel to
 * the end of each iteration, all problems advance together); eig, err:
 * host, [count][trials][step] as in amm_power_method.  Any may be NULL. */
int amm_batch_power_method(amm_batch *b, const amm_params *p, double *time_ms, double *eig,
                           double *err);

/* Step kinds of the last run, host [count][trials][step]. */
int amm_batch_last_modes(amm_batch *b, uint32_t *modes_host, uint64_t len);

/* Current iterate (which = 0) or last product (1) of one (problem, trial). */
int amm_batch_get_vector(amm_batch *b, uint32_t problem, uint32_t trial, int32_t which,
                         float *host, uint64_t len);

/* Wall time of the last batched run from HIP events around it (ms). */
int amm_batch_last_run_ms(amm_batch *b, double *ms);

/* Launch shapes of one problem: workgroups the first contraction aims for
 * (its row slices per product follow from the union size), row block of
 * the second contraction = 256 * passes rows (passes 1 or 2), and the
 * number of chunks the union is split into for the second contraction
 * (1 .. ceil(n / 64)). */
int amm_batch_set_tuning(amm_batch *b, uint32_t problem, uint32_t dot_blocks, uint32_t passes,
                         uint32_t chunks);
int amm_batch_get_tuning(amm_batch *b, uint32_t problem, uint32_t *dot_blocks, uint32_t *passes,
                         uint32_t *chunks);

/* Autotuner: times candidate launch shapes of both contractions on the live
 * problems (all columns present, `iters` repetitions each) and keeps the
 * fastest; *best_ms = one pass of both contractions with it. */
int amm_batch_autotune(amm_batch *b, uint32_t iters, double *best_ms);

/* Autotuner on whole runs: times complete runs of `params` (graph launch,
 * device clock, best of `reps` after one warm-up each) for the shapes in
 * place and along a coordinate descent over the default targets (chunks of
 * the second contraction x 1/4 .. 16, then workgroups of the first x 1/4 ..
 * 4, then passes 1 or 2) and keeps the fastest, so that sampled and exact
 * steps weigh as often as the run takes them; *best_ms = that run's time.
 * Leaves no run results behind that a caller should read. */
int amm_batch_autotune_run(amm_batch *b, const amm_params *params, uint32_t reps, double *best_ms);

/* Mean time (ms) of one pass of both contractions over every column (all
 * trials exact) with the shapes in place; `iters` timed repetitions. */
int amm_batch_time_dense(amm_batch *b, uint32_t iters, double *ms);

/* Union sample of iteration `iter` for one problem with every trial in the
 * step kind of `kernel` (previousAMM: plain CDF, proposedAMM: centred),
 * trial words trial0 + k: the distinct columns (ascending, host, up to n)
 * and their per-trial weights mult * total / (c q_j) (host [count][trials],
 * 0 where a trial did not draw the column); *count = number of columns. */
int amm_batch_sample(amm_batch *b, int32_t kernel, uint32_t iter, uint32_t trial0,
                     uint32_t problem, uint32_t *cols_host, float *w_host, uint64_t *count);

/* sigma_max of every (problem, trial) of a batched run with parameters p
 * (no reference counterpart; see amm_sigma_max): runs the power method,
 * then one exact pass t = A^T x per trial.  sigma_host: host
 * [count][trials], sigma = ||t||; u_dev[i] (length d_i) and v_dev[i]
 * (length n_i): host arrays of device pointers receiving trial 0's unit
 * singular vectors u = x, v = t / sigma (either array may be NULL). */
int amm_batch_sigma_max(amm_batch *b, const amm_params *p, double *sigma_host, float *const *u_dev,
                        float *const *v_dev);

/* Round count FP32 device values to BF16 (nearest even) into dst_dev; the
 * storage form of the BF16 path (no reference counterpart: the reference is
 * FP32 only).  Asynchronous on `stream`. */
int amm_f32_to_bf16(const float *src_dev, void *dst_dev, uint64_t count, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* AMM_H_ */
