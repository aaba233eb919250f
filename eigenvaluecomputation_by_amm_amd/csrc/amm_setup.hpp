// Per-run setup on the device: column statistics and the sampling CDF.
// Device functions shared by the single-matrix kernels (one launch per
// matrix, amm_kernels.hip) and the grouped kernels of the batched engine
// (one launch over every problem, amm_batch.hip), so both produce the same
// bits for the same matrix.
#pragma once

#include "amm_device.hpp"

namespace amm {

// ---------------------------------------------------------------------------
// Column statistics (get_norms, device_memory.h:34-59): alpha_j = mean,
// nrm2_j = ||a_j||^2, cnrm2_j = ||a_j - alpha_j 1||^2 of column j (`col`).
// Called by a whole 256-thread workgroup: wave w streams quarter w of the
// rows (non-temporal 16-byte loads, four per trip), FP64 accumulation, the
// four wave sums added in LDS in a fixed order.  (Four contiguous streams
// per column measured 6.9 TB/s on a 4 GiB FP32 matrix against 6.6 TB/s for
// one wave walking the whole column.)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ double2 colstats_column(const T* __restrict__ col, uint32_t d, uint64_t j,
                                                   float* __restrict__ alpha, double* __restrict__ nrm2,
                                                   double* __restrict__ cnrm2) {
  constexpr int V = Elem<T>::VEC;
  __shared__ double part[2][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // quarter of the rows per wave, in whole 64-lane steps
  const uint32_t step = 64u * V;
  const uint32_t q = (((d + 3) / 4) + step - 1) / step * step;
  const uint32_t rb = w * q;
  const uint32_t re = rb + q < d ? rb + q : d;
  double s = 0.0, ss = 0.0;
  uint32_t row = rb + lane * V;
  // full trips: every lane of the wave has V in-range rows in all 4 loads
  for (; row + step * 3 + V <= re; row += step * 4) {
    float a0[V], a1[V], a2[V], a3[V];
    Elem<T>::template load_full<true>(col + row, a0);
    Elem<T>::template load_full<true>(col + row + step, a1);
    Elem<T>::template load_full<true>(col + row + 2 * step, a2);
    Elem<T>::template load_full<true>(col + row + 3 * step, a3);
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const double b0 = a0[k], b1 = a1[k], b2 = a2[k], b3 = a3[k];
      s += (b0 + b1) + (b2 + b3);
      ss += (b0 * b0 + b1 * b1) + (b2 * b2 + b3 * b3);
    }
  }
  for (; row < re; row += step) {
    float a0[V];
    load_col<T>(col, row, re, a0);
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const double b0 = a0[k];
      s += b0;
      ss += b0 * b0;
    }
  }
  s = wave_sum(s);
  ss = wave_sum(ss);
  if (lane == 0) {
    part[0][w] = s;
    part[1][w] = ss;
  }
  __syncthreads();
  double2 out = make_double2(0.0, 0.0);  // (||a_j||^2, centred) in thread 0
  if (threadIdx.x == 0) {
    const double S = (part[0][0] + part[0][1]) + (part[0][2] + part[0][3]);
    const double SS = (part[1][0] + part[1][1]) + (part[1][2] + part[1][3]);
    const double mean = S / (double)d;
    alpha[j] = (float)mean;
    nrm2[j] = SS;
    double c = SS - S * mean;
    c = c > 0.0 ? c : (c == c ? 0.0 : c);  // clamp rounding below 0, keep NaN
    cnrm2[j] = c;
    out = make_double2(SS, c);
  }
  return out;
}

// ---------------------------------------------------------------------------
// Column statistics of a short column (d * sizeof(T) < kShortColBytes: the
// four-wave split would leave each wave less than one trip of four loads
// per lane), by one wave: lane l adds rows l V + m 64 V in turn, up to eight
// 16-byte loads in flight, then a butterfly.  Callers give each wave of a
// workgroup its own column.  The same function for the same d in the
// single-matrix and the grouped kernels, so both produce the same bits.
// ---------------------------------------------------------------------------
constexpr uint32_t kShortColBytes = 16384;

template <typename T>
__host__ __device__ inline bool short_column(uint32_t d) {
  return (uint64_t)d * sizeof(T) < kShortColBytes;
}

__host__ __device__ inline bool short_column_dt(int dtype, uint32_t d) {
  return (uint64_t)d * (dtype == 1 ? 2u : 4u) < kShortColBytes;
}

template <typename T>
__device__ __forceinline__ void colstats_wave(const T* __restrict__ col, uint32_t d, uint64_t j,
                                              float* __restrict__ alpha, double* __restrict__ nrm2,
                                              double* __restrict__ cnrm2) {
  constexpr int V = Elem<T>::VEC;
  constexpr uint32_t step = 64u * V;
  const uint32_t lane = threadIdx.x & 63u;
  double s = 0.0, ss = 0.0;
  for (uint32_t row = lane * V; row < d; row += 8 * step) {
    float a[8][V];
    if (row - lane * V + 8 * step <= d) {
      // the whole trip lies within the column (wave-uniform): its eight
      // non-temporal loads go out together, with no bound per load
#pragma unroll
      for (int u = 0; u < 8; ++u) Elem<T>::template load_full<true>(col + row + u * step, a[u]);
    } else {
      // the last, partial trip: its whole 64-lane steps load as above, the
      // step that crosses d row by row
      const uint32_t full = (d - (row - lane * V)) / step;  // wave-uniform, < 8
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t r = row + u * step;
        if ((uint32_t)u < full) {
          Elem<T>::template load_full<true>(col + r, a[u]);
        } else if (r < d) {
          load_col<T>(col, r, d, a[u]);  // zeros at and beyond d
        } else {
#pragma unroll
          for (int k = 0; k < V; ++k) a[u][k] = 0.0f;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const double b = a[u][k];
        s += b;
        ss += b * b;
      }
    }
  }
  s = wave_sum(s);
  ss = wave_sum(ss);
  if (lane == 0) {
    const double mean = s / (double)d;
    alpha[j] = (float)mean;
    nrm2[j] = ss;
    double c = ss - s * mean;
    c = c > 0.0 ? c : (c == c ? 0.0 : c);  // clamp rounding below 0, keep NaN
    cnrm2[j] = c;
  }
}

// ---------------------------------------------------------------------------
// Sampling CDF (set_weights + cub::DeviceScan::InclusiveSum,
// common.h:34-40, powerMethod_by_previous.cuh:26-32) in exact integer
// arithmetic:
//   scale = 2^(62 - ceil(log2 n)) / max_j w_j,  q_j = floor(w_j * scale),
//   cdf_j = q_0 + ... + q_j  (uint64).
// Three passes over tiles of kCdfTile columns, one 256-thread workgroup per
// tile: cdf_max_tile (tile maxima of the weights' bit patterns --
// non-negative doubles order like their patterns, and NaN/Inf patterns
// exceed every finite one), cdf_scan_tile (global max from the tile maxima,
// q_j, in-tile scan), cdf_offsets_tile (add the sum of the preceding tiles).
// Max and integer sums are order-free, so the result equals any sequential
// replay bit for bit; the CDF is monotone and q_j = 0 columns can never be
// drawn.  ||alpha||^2 comes along when alpha is given (FP64, fixed order:
// 16 in a lane, tree over the tile, tiles in order).
// ---------------------------------------------------------------------------
constexpr uint32_t kCdfPer = 16;
constexpr uint32_t kCdfTile = 256 * kCdfPer;

__device__ __forceinline__ void cdf_max_tile(const double* __restrict__ w, uint64_t n,
                                             uint64_t* __restrict__ tile_max, uint32_t tile) {
  __shared__ uint64_t red[256];
  const uint32_t t = threadIdx.x;
  uint64_t m = 0;
  const uint64_t b = (uint64_t)tile * kCdfTile;
  for (uint64_t j = b + t; j < b + kCdfTile && j < n; j += 256) {
    const uint64_t v = (uint64_t)__double_as_longlong(w[j]);
    m = v > m ? v : m;
  }
  red[t] = m;
  __syncthreads();
  for (uint32_t off = 128; off > 0; off >>= 1) {
    if (t < off) red[t] = red[t] > red[t + off] ? red[t] : red[t + off];
    __syncthreads();
  }
  if (t == 0) tile_max[tile] = red[0];
}

__device__ __forceinline__ void cdf_scan_tile(const double* __restrict__ w, uint64_t n,
                                              uint64_t* __restrict__ cdf, DevState* st,
                                              const float* __restrict__ alpha,
                                              const uint64_t* __restrict__ tile_max, uint32_t ntiles,
                                              uint64_t* __restrict__ tile_sum,
                                              uint64_t* __restrict__ tile_lp,
                                              double* __restrict__ tile_a2, uint32_t tile) {
  __shared__ uint64_t sc[2][256];
  __shared__ double ad[256];
  const uint32_t t = threadIdx.x;
  {
    uint64_t m = 0;
    for (uint32_t i = t; i < ntiles; i += 256) m = tile_max[i] > m ? tile_max[i] : m;
    sc[0][t] = m;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
      if (t < off) sc[0][t] = sc[0][t] > sc[0][t + off] ? sc[0][t] : sc[0][t + off];
      __syncthreads();
    }
  }
  const double mx = __longlong_as_double((long long)sc[0][0]);
  const bool bad = !(mx <= 1.79769313486231570e+308);  // NaN or Inf
  const double scale = (!bad && mx > 0.0) ? ldexp(1.0, 62 - ceil_log2_u64(n)) / mx : 0.0;
  __syncthreads();
  if (bad && tile == 0 && t == 0) atomicOr(&st->status, kStatusNonFinite);
  const uint64_t b = (uint64_t)tile * kCdfTile + (uint64_t)t * kCdfPer;
  uint64_t q[kCdfPer];
  uint64_t run = 0, lp = 0;
  bool any = false;
  double a2 = 0.0;
#pragma unroll
  for (uint32_t k = 0; k < kCdfPer; ++k) {
    const uint64_t j = b + k;
    q[k] = j < n ? (uint64_t)(w[j] * scale) : 0ull;
    run += q[k];
    q[k] = run;  // in-lane inclusive
    if (j < n && q[k] != (k ? q[k - 1] : 0ull)) {
      lp = j;
      any = true;
    }
    if (alpha && j < n) {
      const double a = alpha[j];
      a2 += a * a;
    }
  }
  sc[0][t] = run;
  ad[t] = a2;
  __syncthreads();
  int cur = 0;
  for (uint32_t off = 1; off < 256; off <<= 1) {
    const uint64_t v = t >= off ? sc[cur][t] + sc[cur][t - off] : sc[cur][t];
    sc[cur ^ 1][t] = v;
    cur ^= 1;
    __syncthreads();
  }
  const uint64_t pre = t ? sc[cur][t - 1] : 0ull;
#pragma unroll
  for (uint32_t k = 0; k < kCdfPer; ++k)
    if (b + k < n) cdf[b + k] = pre + q[k];
  // tile-level: total, last positive column, alpha^2 partial
  __syncthreads();
  sc[cur ^ 1][t] = any ? lp + 1 : 0ull;  // 0 = none in this lane
  __syncthreads();
  for (uint32_t off = 128; off > 0; off >>= 1) {
    if (t < off) {
      const uint64_t a = sc[cur ^ 1][t], c = sc[cur ^ 1][t + off];
      sc[cur ^ 1][t] = a > c ? a : c;
      ad[t] = ad[t] + ad[t + off];
    }
    __syncthreads();
  }
  if (t == 0) {
    tile_sum[tile] = sc[cur][255];
    tile_lp[tile] = sc[cur ^ 1][0];
    tile_a2[tile] = ad[0];
  }
}

__device__ __forceinline__ void cdf_offsets_tile(uint64_t* __restrict__ cdf, uint64_t n,
                                                 const uint64_t* __restrict__ tile_sum,
                                                 const uint64_t* __restrict__ tile_lp,
                                                 const double* __restrict__ tile_a2, DevState* st,
                                                 int which, int want_a2, uint32_t tile, uint32_t ntiles) {
  __shared__ uint64_t red[256];
  const uint32_t t = threadIdx.x;
  uint64_t s = 0;
  for (uint32_t i = t; i < tile; i += 256) s += tile_sum[i];
  red[t] = s;
  __syncthreads();
  for (uint32_t off = 128; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  const uint64_t offset = red[0];
  const uint64_t b = (uint64_t)tile * kCdfTile;
  if (offset)
    for (uint64_t j = b + t; j < b + kCdfTile && j < n; j += 256) cdf[j] += offset;
  if (tile == ntiles - 1 && t == 0) {
    st->total[which] = offset + tile_sum[tile];
    uint64_t lp = 0;
    double a2 = 0.0;
    for (uint32_t i = 0; i <= tile; ++i) {
      lp = tile_lp[i] > lp ? tile_lp[i] : lp;
      a2 += tile_a2[i];
    }
    st->last_pos[which] = lp ? lp - 1 : 0ull;
    if (want_a2) st->alpha2 = a2;
  }
}

}  // namespace amm
