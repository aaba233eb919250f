/*
 * amm_oracle.c -- CPU restatement of the AMM power method for the largest
 * eigenvalue of A*A^T.  TEST INFRASTRUCTURE ONLY: this file is the checker
 * that the HIP library is compared against.  It is linked only by tests/,
 * by the smoke check and by the CPU-baseline leg of bench.py; the product
 * library never calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * The reference is square (m x m, symmetric); this restatement uses the
 * rectangular generalisation A in R^{d x n}, column-major with leading
 * dimension ld, column j = A + ld*j (the reference addresses column j as
 * A + m*j, core_previousAMM.cuh:47,219).  The exact step is y = A (A^T x),
 * which equals the reference's A (A x) when A is symmetric
 * (powerMethod_by_exact.cuh:37-53).
 *
 * Randomness is the library's counter-based Philox4x32-10 stream (Salmon et
 * al., "Parallel random numbers: as easy as 1, 2, 3", SC'11), not cuRAND:
 * the reference's XORWOW/xorshift trajectories are closed-library behaviour
 * and stay unpinned.  A restatement of the reference's own xorshift32 slot
 * sampler (device_memory.h:506-545) is included for statistical parity.
 *
 * Pins (tests/test_oracle_*.py): Philox against the published Random123
 * known-answer vectors; uniform53 exactness; the CDF against numpy FP64
 * cumsum; the inverse-CDF search against a brute-force linear scan; column
 * statistics against numpy FP64 closed forms; one AMM product in sanity mode
 * against the exact product; the proposed rank-2 identity; unbiasedness of
 * the estimator by Monte Carlo; the power method against LAPACK eigvalsh;
 * xorshift32 (13,17,5) full period by GF(2) matrix powers.
 */
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al. SC'11, Random123 reference constants).       */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += PHILOX_W0; k1 += PHILOX_W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0,1) from two 32-bit words: (a>>5)*2^26 + (b>>6), / 2^53.
 * Exactly representable; no rounding. */
double orc_uniform53(uint32_t a, uint32_t b)
{
  uint64_t v = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)v * (1.0 / 9007199254740992.0);
}

/* Stream identifiers in the fourth counter word. */
#define STREAM_SAMPLE 1u
#define STREAM_X0 2u

/* ------------------------------------------------------------------------ */
/* x0 ~ N(0, I), normalised.  Restates curandGenerateNormal + normalize<512> */
/* (powerMethod_by_exact.cuh:31-33, common.h:156-183) with Philox and       */
/* Box-Muller: row pair (2m, 2m+1) uses counter (m, 0, trial, STREAM_X0).   */
/* ------------------------------------------------------------------------ */
void orc_x0_raw(uint64_t seed, uint32_t trial, uint64_t d, float *g)
{
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (uint64_t m = 0; 2 * m < d; ++m) {
    uint32_t ctr[4] = {(uint32_t)m, 0u, trial, STREAM_X0}, r[4];
    orc_philox4x32_10(ctr, key, r);
    double u1 = orc_uniform53(r[0], r[1]);
    double u2 = orc_uniform53(r[2], r[3]);
    double rad = sqrt(-2.0 * log(1.0 - u1));
    double th = 6.283185307179586 * u2;
    g[2 * m] = (float)(rad * cos(th));
    if (2 * m + 1 < d) g[2 * m + 1] = (float)(rad * sin(th));
  }
}

void orc_x0(uint64_t seed, uint32_t trial, uint64_t d, float *x)
{
  orc_x0_raw(seed, trial, d, x);
  double s = 0.0;
  for (uint64_t i = 0; i < d; ++i) s += (double)x[i] * (double)x[i];
  double scale = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
  for (uint64_t i = 0; i < d; ++i) x[i] = (float)(scale * (double)x[i]);
}

/* ------------------------------------------------------------------------ */
/* Column statistics.  Restates get_norms (device_memory.h:34-59): mean     */
/* alpha_j, ||a_j||^2 and ||a_j - alpha_j 1||^2, FP64 accumulation.  The     */
/* mean divides by the true d (the reference divides by the padded m).      */
/* ------------------------------------------------------------------------ */
static inline float bf16_to_f32(uint16_t h)
{
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static void colstats_core(const void *A, int bf16, uint64_t d, uint64_t n, uint64_t ld,
                          float *alpha, double *nrm2, double *cnrm2)
{
#pragma omp parallel for schedule(static)
  for (uint64_t j = 0; j < n; ++j) {
    double s = 0.0, ss = 0.0;
    for (uint64_t i = 0; i < d; ++i) {
      double v = bf16 ? (double)bf16_to_f32(((const uint16_t *)A)[j * ld + i])
                      : (double)((const float *)A)[j * ld + i];
      s += v;
      ss += v * v;
    }
    double mean = s / (double)d;
    alpha[j] = (float)mean;
    nrm2[j] = ss;
    double c = ss - s * mean;
    cnrm2[j] = c > 0.0 ? c : 0.0;
  }
}

void orc_colstats_f32(const float *A, uint64_t d, uint64_t n, uint64_t ld,
                      float *alpha, double *nrm2, double *cnrm2)
{
  colstats_core(A, 0, d, n, ld, alpha, nrm2, cnrm2);
}

void orc_colstats_bf16(const uint16_t *A, uint64_t d, uint64_t n, uint64_t ld,
                       float *alpha, double *nrm2, double *cnrm2)
{
  colstats_core(A, 1, d, n, ld, alpha, nrm2, cnrm2);
}

/* ------------------------------------------------------------------------ */
/* Sampling CDF.  Restates set_weights + cub::DeviceScan::InclusiveSum      */
/* (common.h:34-40, powerMethod_by_previous.cuh:26-32) in exact integer     */
/* arithmetic:                                                               */
/*   L = ceil(log2 n), scale = 2^(62-L) / max_j w_j,                          */
/*   q_j = floor(w_j * scale)            (so sum_j q_j <= 2^62),              */
/*   cdf_j = q_0 + ... + q_j             (uint64, exact, order-free).         */
/* The sampling law is p_j = q_j / total exactly.  Every step is             */
/* independent of summation order, so a parallel device scan equals this     */
/* sequential one bit for bit, the CDF is monotone and q_j = 0 columns are   */
/* never drawn.  A non-finite weight makes the scale negative (error).       */
/* ------------------------------------------------------------------------ */
int orc_ceil_log2(uint64_t n)
{
  int L = 0;
  while (L < 63 && (1ull << L) < n) ++L;
  return L;
}

double orc_cdf_scale(const double *w, uint64_t n)
{
  double mx = 0.0;
  for (uint64_t j = 0; j < n; ++j) {
    if (!(w[j] >= 0.0) || isinf(w[j])) return -1.0;
    if (w[j] > mx) mx = w[j];
  }
  return mx > 0.0 ? ldexp(1.0, 62 - orc_ceil_log2(n)) / mx : 0.0;
}

void orc_cdf(const double *w, uint64_t n, uint64_t *cdf, uint64_t *total, uint64_t *last_pos)
{
  const double scale = orc_cdf_scale(w, n);
  uint64_t run = 0, lp = 0;
  for (uint64_t j = 0; j < n; ++j) {
    uint64_t q = scale > 0.0 ? (uint64_t)(w[j] * scale) : 0u;
    run += q;
    cdf[j] = run;
    if (q > 0) lp = j;
  }
  *total = run;
  *last_pos = lp;
}

/* Inverse-CDF search.  Restates pick_index (device_memory.h:519-545):
 * j = min{ j : cdf[j] > t }, over [0, n-1]; t < total always holds when
 * total > 0, and the fallback (total = 0) is the last positive column. */
uint64_t orc_search(const uint64_t *cdf, uint64_t n, uint64_t t, uint64_t last_pos)
{
  uint64_t lo = 0, hi = n; /* answer in [lo, hi] */
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (cdf[mid] > t) hi = mid; else lo = mid + 1;
  }
  return lo < n ? lo : last_pos;
}

/* t = floor(u * total / 2^64) for the 64-bit uniform u = (a << 32) | b. */
uint64_t orc_draw_target(uint32_t a, uint32_t b, uint64_t total)
{
  const uint64_t u = ((uint64_t)a << 32) | (uint64_t)b;
  return (uint64_t)(((unsigned __int128)u * (unsigned __int128)total) >> 64);
}

#define SAMPLE_SEG 8192u

static int cmp_u32(const void *a, const void *b)
{
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

/* Draw the c column samples of one iteration and reduce them to sorted
 * unique columns with weights mult * total / (c * w_j).  Restates
 * pick_index + cub::DeviceRadixSort::SortKeys + update_weights
 * (device_memory.h:519-545, powerMethod_by_previous.cuh:56-70,
 * common.h:42-53), with duplicates merged (same sum, one read per column).
 * Slot k uses counter (k>>1, iter, trial, STREAM_SAMPLE) words 2(k&1),
 * 2(k&1)+1 as a 64-bit uniform.  Slots are grouped in segments of
 * SAMPLE_SEG; segment g's unique list is written at idx[g*SAMPLE_SEG ...]
 * with its length in seg_count[g].  q_j = cdf[j] - cdf[j-1]; the weight is
 * mult * total / (c * q_j), FP64 then rounded to FP32. */
void orc_sample(const uint64_t *cdf, uint64_t n, uint64_t total, uint64_t last_pos,
                uint64_t seed, uint32_t iter, uint32_t trial, uint32_t c,
                uint32_t *idx, float *wgt, uint32_t *seg_count)
{
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (SAMPLE_SEG + 1));
  const uint32_t nseg = (c + SAMPLE_SEG - 1) / SAMPLE_SEG;
  for (uint32_t g = 0; g < nseg; ++g) {
    uint32_t b = g * SAMPLE_SEG, e = b + SAMPLE_SEG < c ? b + SAMPLE_SEG : c;
    for (uint32_t k = b; k < e; ++k) {
      uint32_t ctr[4] = {k >> 1, iter, trial, STREAM_SAMPLE}, r[4];
      orc_philox4x32_10(ctr, key, r);
      uint64_t t = orc_draw_target(r[2 * (k & 1)], r[2 * (k & 1) + 1], total);
      tmp[k - b] = (uint32_t)orc_search(cdf, n, t, last_pos);
    }
    uint32_t cnt = e - b;
    qsort(tmp, cnt, sizeof(uint32_t), cmp_u32);
    uint32_t u = 0;
    for (uint32_t p = 0; p < cnt;) {
      uint32_t q = p;
      while (q < cnt && tmp[q] == tmp[p]) ++q;
      uint32_t j = tmp[p];
      uint64_t qj = cdf[j] - (j ? cdf[j - 1] : 0u);
      double num = (double)(q - p) * (double)total;
      double den = (double)c * (double)qj;
      idx[b + u] = j;
      wgt[b + u] = den > 0.0 ? (float)(num / den) : 0.0f;
      ++u;
      p = q;
    }
    seg_count[g] = u;
  }
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* One AMM product (FP64 accumulation, FP32 storage of the vectors).         */
/* previous: y = y_memo + sum_k wgt_k a_k (a_k^T dx),  dx = x - x_memo       */
/*   (core_previousAMM.cuh:4-175 kernel1, :177-263 kernel2).                  */
/* proposed: t_k = wgt_k (a_k^T dx - alpha_k s),                              */
/*   y = y_memo + s v + q 1 + sum_k (a_k - alpha_k 1) t_k,  s = 1^T dx,        */
/*   q = z^T dx  (common.h:55-87, core_proposedAMM.cuh:17-313).                */
/* The list is either sampled (idx/wgt/seg_count) or, in sanity mode, all n  */
/* columns with weight 1 (device_memory.h:500-504,547-551).                  */
/* ------------------------------------------------------------------------ */
typedef struct {
  const float *A; /* FP32 view of the matrix (BF16 inputs are widened first) */
  uint64_t d, n, ld;
} orc_mat;

static double col_dot(const orc_mat *M, uint64_t j, const double *v)
{
  const float *a = M->A + j * M->ld;
  double s = 0.0;
  for (uint64_t i = 0; i < M->d; ++i) s += (double)a[i] * v[i];
  return s;
}

/* acc[i] += sum_e a_{cols[e]}[i] * t[e], every element summed in entry
 * order (the serial order); row blocks run in parallel (OpenMP), so the
 * result does not depend on the thread count. */
#define ORC_ROW_BLOCK 1024
static void axpy_rows(const float *A, uint64_t ld, uint64_t d, uint64_t m, const uint64_t *cols,
                      const double *t, double *acc)
{
  const int64_t nblk = (int64_t)((d + ORC_ROW_BLOCK - 1) / ORC_ROW_BLOCK);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nblk; ++b) {
    const uint64_t i0 = (uint64_t)b * ORC_ROW_BLOCK;
    const uint64_t i1 = i0 + ORC_ROW_BLOCK < d ? i0 + ORC_ROW_BLOCK : d;
    for (uint64_t e = 0; e < m; ++e) {
      const float *a = A + (cols ? cols[e] : e) * ld;
      const double te = t[e];
      for (uint64_t i = i0; i < i1; ++i) acc[i] += (double)a[i] * te;
    }
  }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_amm_product(const float *A, uint64_t d, uint64_t n, uint64_t ld, int proposed,
                     uint32_t c, const uint32_t *idx, const float *wgt, const uint32_t *seg_count,
                     int sanity, const float *x, const float *x_memo, const float *y_memo,
                     const float *alpha, const float *v, const float *z, float *y)
{
  orc_mat M = {A, d, n, ld};
  double *dx = (double *)malloc(sizeof(double) * d);
  double *acc = (double *)calloc(d, sizeof(double));
  double s = 0.0, q = 0.0;
  for (uint64_t i = 0; i < d; ++i) {
    dx[i] = (double)x[i] - (x_memo ? (double)x_memo[i] : 0.0);
  }
  if (proposed) {
    for (uint64_t i = 0; i < d; ++i) { s += dx[i]; q += (double)z[i] * dx[i]; }
  }
  /* the list in order: segment by segment (sampled) or all n columns (sanity) */
  const uint32_t nseg = sanity ? 1 : (c + SAMPLE_SEG - 1) / SAMPLE_SEG;
  uint64_t m = 0;
  for (uint32_t g = 0; g < nseg; ++g) m += sanity ? n : seg_count[g];
  uint64_t *cols = (uint64_t *)malloc(sizeof(uint64_t) * (m ? m : 1));
  double *w = (double *)malloc(sizeof(double) * (m ? m : 1));
  double *t = (double *)malloc(sizeof(double) * (m ? m : 1));
  uint64_t e = 0;
  for (uint32_t g = 0; g < nseg; ++g) {
    uint32_t cnt = sanity ? (uint32_t)n : seg_count[g];
    for (uint32_t p = 0; p < cnt; ++p, ++e) {
      cols[e] = sanity ? p : idx[g * SAMPLE_SEG + p];
      w[e] = sanity ? 1.0 : (double)wgt[g * SAMPLE_SEG + p];
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)m; ++k) {
    const double r = col_dot(&M, cols[k], dx);
    const double aj = proposed ? (double)alpha[cols[k]] : 0.0;
    t[k] = w[k] * (r - aj * s);
  }
  double corr = 0.0;
  for (uint64_t k = 0; k < m; ++k) corr += (proposed ? (double)alpha[cols[k]] : 0.0) * t[k];
  axpy_rows(A, ld, d, m, cols, t, acc);
  for (uint64_t i = 0; i < d; ++i) {
    double b = y_memo ? (double)y_memo[i] : 0.0;
    if (proposed) b += s * (double)v[i] + q - corr;
    y[i] = (float)(b + acc[i]);
  }
  free(cols);
  free(w);
  free(t);
  free(dx);
  free(acc);
}

/* Exact product y = A (A^T x), FP64 accumulation (powerMethod_by_exact.cuh:37-53). */
void orc_exact_product(const float *A, uint64_t d, uint64_t n, uint64_t ld, const float *x, float *y)
{
  double *xv = (double *)malloc(sizeof(double) * d);
  double *acc = (double *)calloc(d, sizeof(double));
  double *t = (double *)malloc(sizeof(double) * (n ? n : 1));
  orc_mat M = {A, d, n, ld};
  for (uint64_t i = 0; i < d; ++i) xv[i] = x[i];
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)n; ++j) t[j] = col_dot(&M, (uint64_t)j, xv);
  axpy_rows(A, ld, d, n, NULL, t, acc);
  for (uint64_t i = 0; i < d; ++i) y[i] = (float)acc[i];
  free(t);
  free(xv);
  free(acc);
}

/* ------------------------------------------------------------------------ */
/* Power method with the reference's schedules.                              */
/* kernel: 0 exact, 1 previousAMM, 2 proposedAMM, 3 proposedAMM_v2.          */
/* Restates powerMethod_by_exact.cuh:4-100, powerMethod_by_previous.cuh:     */
/* 100-210, powerMethod_by_proposed.cuh:105-235, powerMethod_by_proposed_v2. */
/* cuh:199-377 and dotProduct_and_generate_x (common.h:89-154):              */
/*   eig_i = y^T x,  err_i = y^T v* / ||y||,  cos_i = y^T x / ||y||,          */
/*   x <- y / ||y||.                                                          */
/* Weights are taken from w_plain / w_cent when given (so that the samples   */
/* are the device's to the bit), else computed here.                         */
/* ------------------------------------------------------------------------ */
typedef struct {
  int32_t kernel;
  uint32_t c;
  uint64_t step;
  uint64_t vr_period;
  double vr_adaptive_threshold;
  uint64_t seed;
  int32_t sanity;
  uint32_t trial;
  int32_t no_vr; /* the reference's variance_reduction = false (host_memory.h:108-120) */
  uint32_t pad;
} orc_params;

/* Returns 1 when iteration i is an exact (variance-reduction) step.
 * previous:  powerMethod_by_previous.cuh:135-138
 * proposed:  powerMethod_by_proposed.cuh:159-161 (gate i > last_exact + 1)  */
int orc_is_exact_step(int32_t kernel, uint64_t i, uint64_t vr_period, double thr,
                      double prev_cos, uint64_t last_exact)
{
  int vr = (vr_period > 0) || (thr > 0.0);
  int due = (vr_period > 0 && (i + 1) % vr_period == 0) || (thr > 0.0 && prev_cos > 1.0 - thr);
  if (kernel == 0) return 1;
  if (kernel == 1) return vr && due;
  if (kernel == 2) return (i > last_exact + 1) && vr && due;
  /* proposedAMM_v2 (powerMethod_by_proposed_v2.cuh:234,257-277,305) */
  if (i + 1 < vr_period) return 0;
  if (i + 1 == vr_period) return 1;
  return vr_period > 0 && (i + 1) % vr_period == 0;
}

int orc_power_method(const float *A, uint64_t d, uint64_t n, uint64_t ld, const orc_params *P,
                     const double *w_plain_in, const double *w_cent_in, const double *eigvec,
                     double *eig_out, double *err_out, uint8_t *mode_out, float *x_final)
{
  const int kern = P->kernel;
  if (kern == 3 && P->vr_period < 1) return -1; /* v2 needs 1 <= vr_period */
  float *alpha = (float *)calloc(n, sizeof(float));
  double *nrm2 = (double *)calloc(n, sizeof(double));
  double *cnrm2 = (double *)calloc(n, sizeof(double));
  uint64_t *cdf_p = (uint64_t *)calloc(n, sizeof(uint64_t));
  uint64_t *cdf_c = (uint64_t *)calloc(n, sizeof(uint64_t));
  uint64_t tot_p = 0, tot_c = 0;
  uint64_t lp_p = 0, lp_c = 0;
  /* the exact method samples nothing: no column statistics, as
   * powerMethod_by_exact.cuh:4-100 reads no norms */
  if (kern != 0) {
    orc_colstats_f32(A, d, n, ld, alpha, nrm2, cnrm2);
    const double *wp = w_plain_in ? w_plain_in : nrm2;
    const double *wc = w_cent_in ? w_cent_in : cnrm2;
    orc_cdf(wp, n, cdf_p, &tot_p, &lp_p);
    orc_cdf(wc, n, cdf_c, &tot_c, &lp_c);
  }

  float *x = (float *)malloc(sizeof(float) * d);
  float *y = (float *)malloc(sizeof(float) * d);
  float *xm = (float *)calloc(d, sizeof(float));
  float *ym = (float *)calloc(d, sizeof(float));
  float *z = (float *)calloc(d, sizeof(float));
  float *v = (float *)calloc(d, sizeof(float));
  uint32_t cap = P->c ? P->c : 1;
  uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * cap);
  float *wgt = (float *)malloc(sizeof(float) * cap);
  uint32_t *segc = (uint32_t *)malloc(sizeof(uint32_t) * ((cap + SAMPLE_SEG - 1) / SAMPLE_SEG));

  /* z = A alpha, v = z - ||alpha||^2 (powerMethod_by_proposed.cuh:132-145) */
  double a2 = 0.0;
  for (uint64_t j = 0; j < n; ++j) a2 += (double)alpha[j] * (double)alpha[j];
  if (kern >= 2) {
    double *acc = (double *)calloc(d, sizeof(double));
    double *al = (double *)malloc(sizeof(double) * (n ? n : 1));
    for (uint64_t j = 0; j < n; ++j) al[j] = (double)alpha[j];
    axpy_rows(A, ld, d, n, NULL, al, acc);
    free(al);
    for (uint64_t i = 0; i < d; ++i) {
      z[i] = (float)acc[i];
      v[i] = (float)((double)z[i] - a2);
    }
    free(acc);
  }

  orc_x0(P->seed, P->trial, d, x);
  double prev_cos = 0.0;
  uint64_t last_exact = 0;
  for (uint64_t i = 0; i < P->step; ++i) {
    int ex;
    if (P->no_vr && kern != 0) /* only v2's transition step stays (powerMethod_by_proposed_v2.cuh:255-277) */
      ex = kern == 3 && i + 1 == P->vr_period;
    else
      ex = orc_is_exact_step(kern, i, P->vr_period, kern == 3 ? 0.0 : P->vr_adaptive_threshold,
                             prev_cos, last_exact);
    int centred = (kern == 2) || (kern == 3 && i + 1 > P->vr_period);
    if (ex) {
      orc_exact_product(A, d, n, ld, x, y);
      int memo = (kern != 0) && !(kern == 3 && i + 1 == P->vr_period);
      if (memo) {
        memcpy(xm, x, sizeof(float) * d);
        memcpy(ym, y, sizeof(float) * d);
      }
      last_exact = i;
    } else {
      if (!P->sanity)
        orc_sample(centred ? cdf_c : cdf_p, n, centred ? tot_c : tot_p, centred ? lp_c : lp_p,
                   P->seed, (uint32_t)i, P->trial, P->c, idx, wgt, segc);
      orc_amm_product(A, d, n, ld, centred, P->c, idx, wgt, segc, P->sanity, x, xm, ym, alpha, v,
                      z, y);
    }
    double yx = 0.0, yy = 0.0, yv = 0.0;
    for (uint64_t r = 0; r < d; ++r) {
      double yr = y[r];
      yx += yr * (double)x[r];
      yy += yr * yr;
      if (eigvec) yv += yr * eigvec[r];
    }
    double scale = yy > 0.0 ? 1.0 / sqrt(yy) : 0.0;
    eig_out[i] = yx;
    err_out[i] = yv * scale;
    prev_cos = yx * scale;
    if (mode_out) mode_out[i] = (uint8_t)(ex ? 1 : (centred ? 3 : 2));
    for (uint64_t r = 0; r < d; ++r) x[r] = (float)(scale * (double)y[r]);
  }
  if (x_final) memcpy(x_final, x, sizeof(float) * d);
  free(alpha); free(nrm2); free(cnrm2); free(cdf_p); free(cdf_c);
  free(x); free(y); free(xm); free(ym); free(z); free(v); free(idx); free(wgt); free(segc);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* The reference's own slot sampler, for statistical parity only:            */
/* xorshift32 (13,17,5) on a per-slot state, 23-bit uniform                   */
/* float((y & 0x7FFFFF) | 0x3F800000) - 1, FP32 CDF, binary search           */
/* (device_memory.h:506-545).  Seeds come from a documented substitute       */
/* (orc_ref_seed_states), not from cuRAND.                                   */
/* ------------------------------------------------------------------------ */
uint32_t orc_xorshift32(uint32_t *state)
{
  uint32_t y = *state;
  y ^= (y << 13);
  y ^= (y >> 17);
  y ^= (y << 5);
  *state = y;
  return y;
}

float orc_ref_uniform23(uint32_t *state)
{
  uint32_t y = orc_xorshift32(state);
  uint32_t bits = (y & 0x007FFFFFu) | 0x3F800000u;
  float f;
  memcpy(&f, &bits, 4);
  return f - 1.0f;
}

/* Seed substitute: state_k = splitmix64(seed, k) truncated, never 0. */
void orc_ref_seed_states(uint64_t seed, uint32_t c, uint32_t *states)
{
  for (uint32_t k = 0; k < c; ++k) {
    uint64_t zz = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(k + 1);
    zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ull;
    zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBull;
    zz ^= zz >> 31;
    uint32_t s = (uint32_t)zz;
    states[k] = s ? s : 0x6D2B79F5u;
  }
}

/* FP32 sequential inclusive scan (the numerical content of the FP32 CUB
 * scan: powerMethod_by_previous.cuh:26-32). */
void orc_ref_cdf_f32(const float *w, uint64_t n, float *cdf)
{
  float s = 0.0f;
  for (uint64_t j = 0; j < n; ++j) { s += w[j]; cdf[j] = s; }
}

/* pick_index's search over [-1, n-1] (device_memory.h:529-543). */
uint32_t orc_ref_pick(const float *cdf, uint64_t n, float u)
{
  int64_t lo = -1, hi = (int64_t)n - 1;
  const float target = cdf[n - 1] * u;
  while (lo + 1 != hi) {
    int64_t mid = (lo + hi) / 2;
    if (cdf[mid] <= target) lo = mid; else hi = mid;
  }
  return (uint32_t)hi;
}

/* The reference's previous/proposed AMM power method, run with its own slot
 * sampler and FP32 weights w_j = total / (norm_j^2 * c)
 * (common.h:34-53), duplicates kept as separate terms.  Returns eig_i. */
int orc_ref_power_method(const float *A, uint64_t d, uint64_t n, uint64_t ld, int32_t kernel,
                         uint32_t c, uint64_t step, uint64_t seed, const float *x0,
                         double *eig_out)
{
  float *alpha = (float *)malloc(sizeof(float) * n);
  double *nrm2 = (double *)malloc(sizeof(double) * n);
  double *cnrm2 = (double *)malloc(sizeof(double) * n);
  orc_colstats_f32(A, d, n, ld, alpha, nrm2, cnrm2);
  int prop = kernel == 2;
  float *wf = (float *)malloc(sizeof(float) * n);
  float *cdf = (float *)malloc(sizeof(float) * n);
  for (uint64_t j = 0; j < n; ++j) {
    float nrm = (float)sqrt(prop ? cnrm2[j] : nrm2[j]);
    wf[j] = nrm * nrm;
  }
  orc_ref_cdf_f32(wf, n, cdf);
  float total = cdf[n - 1];
  uint32_t *states = (uint32_t *)malloc(sizeof(uint32_t) * c);
  orc_ref_seed_states(seed, c, states);
  float *x = (float *)malloc(sizeof(float) * d);
  float *y = (float *)malloc(sizeof(float) * d);
  float *z = (float *)calloc(d, sizeof(float));
  float *v = (float *)calloc(d, sizeof(float));
  double a2 = 0.0;
  for (uint64_t j = 0; j < n; ++j) a2 += (double)alpha[j] * alpha[j];
  if (prop) {
    double *acc = (double *)calloc(d, sizeof(double));
    for (uint64_t j = 0; j < n; ++j)
      for (uint64_t i = 0; i < d; ++i) acc[i] += (double)A[j * ld + i] * alpha[j];
    for (uint64_t i = 0; i < d; ++i) { z[i] = (float)acc[i]; v[i] = (float)(z[i] - a2); }
    free(acc);
  }
  memcpy(x, x0, sizeof(float) * d);
  uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * c);
  float *wgt = (float *)malloc(sizeof(float) * c);
  uint32_t segc[1];
  for (uint64_t it = 0; it < step; ++it) {
    for (uint32_t k = 0; k < c; ++k) {
      float u = orc_ref_uniform23(&states[k]);
      uint32_t j = orc_ref_pick(cdf, n, u);
      idx[k] = j;
      wgt[k] = total / (wf[j] * (float)c);
    }
    segc[0] = c;
    /* duplicates kept: feed the list as one segment of length c */
    {
      orc_mat M = {A, d, n, ld};
      double *dx = (double *)malloc(sizeof(double) * d);
      double *acc = (double *)calloc(d, sizeof(double));
      double s = 0.0, q = 0.0, corr = 0.0;
      for (uint64_t i = 0; i < d; ++i) dx[i] = x[i];
      if (prop) for (uint64_t i = 0; i < d; ++i) { s += dx[i]; q += (double)z[i] * dx[i]; }
      for (uint32_t k = 0; k < c; ++k) {
        uint64_t j = idx[k];
        double aj = prop ? alpha[j] : 0.0;
        double t = (double)wgt[k] * (col_dot(&M, j, dx) - aj * s);
        for (uint64_t i = 0; i < d; ++i) acc[i] += (double)A[j * ld + i] * t;
        corr += aj * t;
      }
      for (uint64_t i = 0; i < d; ++i) {
        double b = prop ? s * (double)v[i] + q - corr : 0.0;
        y[i] = (float)(b + acc[i]);
      }
      free(dx);
      free(acc);
    }
    double yx = 0.0, yy = 0.0;
    for (uint64_t r = 0; r < d; ++r) { yx += (double)y[r] * x[r]; yy += (double)y[r] * y[r]; }
    double scale = yy > 0.0 ? 1.0 / sqrt(yy) : 0.0;
    eig_out[it] = yx;
    for (uint64_t r = 0; r < d; ++r) x[r] = (float)(scale * y[r]);
  }
  free(alpha); free(nrm2); free(cnrm2); free(wf); free(cdf); free(states);
  free(x); free(y); free(z); free(v); free(idx); free(wgt);
  return 0;
}
