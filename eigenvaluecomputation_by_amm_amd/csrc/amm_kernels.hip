// amm_kernels.hip -- HIP kernels of the AMM power method for lambda_max(A A^T)
// on gfx950 (wave64).
//
// Matrix layout: A is d x n, column-major, leading dimension ld (column j is
// the contiguous run A + ld*j, as in the reference's "A + m*idx" addressing,
// core_previousAMM.cuh:47,219).  Vectors of length d are FP32 and padded to
// dpad = roundup(d, 8) with zeros, so 16-byte vector loads never leave them.
//
// A sampled power iteration is two launches where d allows (8192 <= d,
// d a multiple of the dot pass's row split, static step schedule):
//   gather_dot   P1[r][e] = a_{col(e)}[rows of split r]^T (x - x_memo)[rows],
//                carrying the previous iteration's tail: 1/||y||, lambda,
//                err, x = y / ||y|| (+ s = 1^T dx, q = z^T dx for proposed)
//   axpy_finish  t_e = wgt_e (sum_r P1[r][e] - alpha_col s); y = base +
//                sum_e a_{col(e)} t_e; block partials of y^T x, y^T y, y^T v*
// Otherwise the second contraction is chunked (gather_axpy + finish) and a
// tail kernel closes every iteration.  The exact step is k_gram (A (A^T x)
// in one read) up to d = 16384, else the same two contraction kernels over
// all n columns with unit weights (the sanity-mode identity of
// device_memory.h:500-551), then finish.
// Every reduction runs in a fixed order, so results do not depend on timing
// (the reference's kernel2 uses float atomics, core_previousAMM.cuh:259-262).
#include "amm_device.hpp"
#include "amm_launch.hpp"
#include "amm_setup.hpp"
#include "amm_resident.hpp"

#include <algorithm>
#include <map>
#include <mutex>

namespace amm {

// Column statistics: short columns (colstats_wave) a wave each, four per
// workgroup; longer ones (colstats_column) a workgroup each, workgroup b
// taking columns [b K, (b + 1) K) in turn (K from colstats_per_block).  The
// 4-wave split of each column and its sums do not depend on K, so neither do
// the bits.
template <typename T>
__global__ __launch_bounds__(256) void k_colstats(const T* __restrict__ A, uint64_t ld, uint32_t d,
                                                  uint64_t n, uint32_t K, float* __restrict__ alpha,
                                                  double* __restrict__ nrm2,
                                                  double* __restrict__ cnrm2, uint64_t* __restrict__ wmax) {
  if (short_column<T>(d)) {  // a wave per column, four columns per workgroup
    const uint64_t j = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j < n) colstats_wave<T>(A + j * ld, d, j, alpha, nrm2, cnrm2);
    return;
  }
  const uint64_t j0 = (uint64_t)blockIdx.x * K;
  const uint64_t j1 = j0 + K < n ? j0 + K : n;
  uint64_t m0 = 0, m1 = 0;  // bit patterns of the workgroup's largest weights (thread 0)
  for (uint64_t j = j0; j < j1; ++j) {
    const double2 w = colstats_column<T>(A + j * ld, d, j, alpha, nrm2, cnrm2);
    const uint64_t b0 = (uint64_t)__double_as_longlong(w.x), b1 = (uint64_t)__double_as_longlong(w.y);
    m0 = b0 > m0 ? b0 : m0;
    m1 = b1 > m1 ? b1 : m1;
    __syncthreads();  // the per-column exchange in LDS is reused by the next column
  }
  // per-workgroup maxima of both weights for the CDF (cdf_max_tile's
  // order: non-negative doubles order like their patterns, NaN and Inf
  // above every finite one), in place of its maximum pass
  if (wmax && threadIdx.x == 0) {
    wmax[blockIdx.x] = m0;
    wmax[gridDim.x + blockIdx.x] = m1;
  }
}

// ---------------------------------------------------------------------------
// Column-resident passes.  A 256-thread workgroup holds one column of A in
// registers: wave w owns rows [w q, (w + 1) q) with q = L x 64 lanes x VEC
// (d <= 4 q), each lane L non-temporal 16-byte loads (raw bits; rows at and
// beyond d read as zeros).  Persistent workgroups walk the columns
// j = blockIdx.x, blockIdx.x + gridDim.x, ... and prefetch the next column
// while the workgroup reduces the current one, so A is read once and a
// per-column scalar c_j (a workgroup reduction over the column) can be
// applied back to the same column: every lane adds c_j a_j to its rows of a
// private FP32 partial, written to OUT[blockIdx.x][0, dpad) at the end
// (ColRegs and the partial helpers: amm_resident.hpp).
// ---------------------------------------------------------------------------
// Column statistics with z = A alpha in the same read of A (get_norms,
// device_memory.h:34-59, and the z = A alpha product of
// powerMethod_by_proposed.cuh:132-139): c_j = alpha_j, partials of z in
// ZP[blockIdx.x] (k_sum_rows adds them in FP64).
template <typename T, int L>
__global__ __launch_bounds__(256) void k_colstats_z(const T* __restrict__ A, uint64_t ld, uint32_t d,
                                                    uint64_t n, uint64_t dpad, float* __restrict__ alpha,
                                                    double* __restrict__ nrm2, double* __restrict__ cnrm2,
                                                    float* __restrict__ ZP) {
  constexpr int V = Elem<T>::VEC;
  __shared__ double part[2][2][4];  // [column parity][sum | sum of squares][wave]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t q = resident_rows<V>(d);  // <= L * 64 * V (host)
  const uint32_t rb = w * q;
  const uint32_t re = rb + q < d ? rb + q : d;
  ColRegs<T, L> cr;
  float zp[L * V];
#pragma unroll
  for (int e = 0; e < L * V; ++e) zp[e] = 0.0f;
  uint64_t j = blockIdx.x;
  if (j < n) cr.load(A + j * ld, rb, re, lane);
  for (int par = 0; j < n; j += gridDim.x, par ^= 1) {
    u32x4 cur[L];
#pragma unroll
    for (int l = 0; l < L; ++l) cur[l] = cr.raw[l];
    if (j + gridDim.x < n) cr.load(A + (j + gridDim.x) * ld, rb, re, lane);  // streams in meanwhile
    double s = 0.0, ss = 0.0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      float v[V];
      ColRegs<T, L>::unpack(cur[l], v);
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const double b = v[k];
        s += b;
        ss += b * b;
      }
    }
    s = wave_sum(s);
    ss = wave_sum(ss);
    if (lane == 0) {
      part[par][0][w] = s;
      part[par][1][w] = ss;
    }
    __syncthreads();
    const double S = (part[par][0][0] + part[par][0][1]) + (part[par][0][2] + part[par][0][3]);
    const double SS = (part[par][1][0] + part[par][1][1]) + (part[par][1][2] + part[par][1][3]);
    const double mean = S / (double)d;
    if (threadIdx.x == 0) {
      alpha[j] = (float)mean;
      nrm2[j] = SS;
      double c = SS - S * mean;
      c = c > 0.0 ? c : (c == c ? 0.0 : c);  // clamp rounding below 0, keep NaN
      cnrm2[j] = c;
    }
    resident_axpy<T, L>(zp, (float)mean, cur);
  }
  resident_store<T, L>(zp, ZP + (uint64_t)blockIdx.x * dpad, rb, q, lane, dpad);
}

// The exact product y = A (A^T x) in one read of A (the two cublasSgemv
// calls of powerMethod_by_exact.cuh:46-62): c_j = a_j^T x (FP64 sum of
// exact FP32 products, rounded to FP32 like the separate pass's t_j),
// partials of y in P2[blockIdx.x] for k_finish.  x sits in LDS (zeros
// beyond d).  Skipped when the gate says the step is not exact.
template <typename T, int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_gram(const T* __restrict__ A, uint64_t ld, uint32_t d,
                                              uint64_t n, uint64_t dpad, const float* __restrict__ x,
                                              float* __restrict__ P2, const uint32_t* __restrict__ gate_mode,
                                              uint32_t it, uint32_t gate_mask) {
  if (gate_mode && !((gate_mask >> gate_mode[it]) & 1u)) return;
  constexpr int V = Elem<T>::VEC;
  constexpr uint32_t kRows = 4u * L * 64u * V;
  __shared__ float xs[kRows];
  __shared__ double part[2][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t q = resident_rows<V>(d);
  const uint32_t rb = w * q;
  const uint32_t re = rb + q < d ? rb + q : d;
  ColRegs<T, L> cr;
  uint64_t j = blockIdx.x;
  if (j < n) cr.load(A + j * ld, rb, re, lane);
  for (uint32_t r = threadIdx.x; r < kRows; r += 256) xs[r] = r < d ? x[r] : 0.0f;
  float yp[L * V];
#pragma unroll
  for (int e = 0; e < L * V; ++e) yp[e] = 0.0f;
  __syncthreads();
  for (int par = 0; j < n; j += gridDim.x, par ^= 1) {
    u32x4 cur[L];
#pragma unroll
    for (int l = 0; l < L; ++l) cur[l] = cr.raw[l];
    if (j + gridDim.x < n) cr.load(A + (j + gridDim.x) * ld, rb, re, lane);  // streams in meanwhile
    double t = 0.0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      float v[V];
      ColRegs<T, L>::unpack(cur[l], v);
      const float* xr = xs + rb + lane * V + l * ColRegs<T, L>::step;
#pragma unroll
      for (int k = 0; k < V; k += 4) {
        const float4 xv = *reinterpret_cast<const float4*>(xr + k);
        t += (double)v[k] * (double)xv.x;
        t += (double)v[k + 1] * (double)xv.y;
        t += (double)v[k + 2] * (double)xv.z;
        t += (double)v[k + 3] * (double)xv.w;
      }
    }
    t = wave_sum(t);
    if (lane == 0) part[par][w] = t;
    __syncthreads();
    const float tj = (float)((part[par][0] + part[par][1]) + (part[par][2] + part[par][3]));
    resident_axpy<T, L>(yp, tj, cur);
  }
  resident_store<T, L>(yp, P2 + (uint64_t)blockIdx.x * dpad, rb, q, lane, dpad);
}

// Read ceiling: a plain streaming read of n4 16-byte words (the matrix's
// n * ld elements) and nothing else; each workgroup reads one contiguous
// chunk with eight non-temporal loads per thread in flight.  The XOR of the
// words is written only if it equals a value no input produces in practice,
// which keeps the loads alive.
__global__ __launch_bounds__(256) void k_read(const u32x4* __restrict__ a, uint64_t n4, uint64_t per_block,
                                              uint32_t* __restrict__ sink) {
  const uint64_t b0 = (uint64_t)blockIdx.x * per_block;
  const uint64_t b1 = b0 + per_block < n4 ? b0 + per_block : n4;
  uint32_t acc = 0;
  for (uint64_t i = b0 + threadIdx.x; i < b1; i += 256 * 8) {
    u32x4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = i + u * 256 < b1 ? load16<true>(a + i + u * 256) : u32x4{0u, 0u, 0u, 0u};
#pragma unroll
    for (int u = 0; u < 8; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  if (acc == 0x9E3779B9u && threadIdx.x == 0) sink[0] = acc;
}

// Sampling CDF: three launches over tiles of kCdfTile columns (cdf_max_tile,
// cdf_scan_tile, cdf_offsets_tile).
__global__ __launch_bounds__(256) void k_cdf_max(const double* __restrict__ w, uint64_t n,
                                                 uint64_t* __restrict__ tile_max) {
  cdf_max_tile(w, n, tile_max, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_cdf_tiles(const double* __restrict__ w, uint64_t n,
                                                   uint64_t* __restrict__ cdf, DevState* st,
                                                   int which, const float* __restrict__ alpha,
                                                   const uint64_t* __restrict__ tile_max, uint32_t nmax,
                                                   uint64_t* __restrict__ tile_sum,
                                                   uint64_t* __restrict__ tile_lp,
                                                   double* __restrict__ tile_a2) {
  cdf_scan_tile(w, n, cdf, st, alpha, tile_max, nmax, tile_sum, tile_lp, tile_a2, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_cdf_offsets(uint64_t* __restrict__ cdf, uint64_t n,
                                                     const uint64_t* __restrict__ tile_sum,
                                                     const uint64_t* __restrict__ tile_lp,
                                                     const double* __restrict__ tile_a2,
                                                     DevState* st, int which, int want_a2) {
  cdf_offsets_tile(cdf, n, tile_sum, tile_lp, tile_a2, st, which, want_a2, blockIdx.x, gridDim.x);
}

// Value of lane (lane ^ J) within a wave, J a power of two below 64, by the
// cheapest exchange for J: DPP quad permutes (1, 2), swizzle (4, 8, 16),
// bpermute (32).
template <uint32_t J>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
  if constexpr (J == 1) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
  } else if constexpr (J == 2) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
  } else if constexpr (J < 32) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (J << 10));  // bit mode: lane ^ J within 32
  } else {
    return (uint32_t)__shfl_xor((int)v, (int)J);
  }
}

// Bitonic sort of keys[0, 1024 NE) by a 1024-thread workgroup, the network
// unrolled: thread tid holds the keys at positions tid + 1024 e, e < NE;
// partners 1024 or more apart swap in registers, closer than 64 exchange
// within the wave (xor_lane), the others through LDS.  Sorting all 1024 NE
// positions gives the first cnt sorted keys of any shorter sort: the
// positions past cnt hold sentinel keys, which sort last.
template <uint32_t NE, uint32_t K, uint32_t J>
__device__ __forceinline__ void bitonic_stage(uint32_t* keys, uint32_t tid, uint32_t (&v)[NE]) {
  if constexpr (J >= 1024) {
    constexpr uint32_t je = J / 1024;
#pragma unroll
    for (uint32_t e = 0; e < NE; ++e) {
      if ((e & je) == 0) {
        const bool up = ((tid + 1024 * e) & K) == 0;
        const uint32_t a = v[e], b = v[e | je];
        const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        v[e] = up ? lo : hi;
        v[e | je] = up ? hi : lo;
      }
    }
  } else {
    uint32_t q[NE];
    if constexpr (J >= 64) {
      __syncthreads();
#pragma unroll
      for (uint32_t e = 0; e < NE; ++e) keys[tid + 1024 * e] = v[e];
      __syncthreads();
#pragma unroll
      for (uint32_t e = 0; e < NE; ++e) q[e] = keys[(tid + 1024 * e) ^ J];
    } else {
#pragma unroll
      for (uint32_t e = 0; e < NE; ++e) q[e] = xor_lane<J>(v[e]);
    }
#pragma unroll
    for (uint32_t e = 0; e < NE; ++e) {
      const uint32_t p = tid + 1024 * e;
      const bool up = (p & K) == 0, lower = (p & J) == 0;
      v[e] = (lower == up) ? (v[e] < q[e] ? v[e] : q[e]) : (v[e] < q[e] ? q[e] : v[e]);
    }
  }
  if constexpr (J > 1) bitonic_stage<NE, K, J / 2>(keys, tid, v);
}

template <uint32_t NE, uint32_t K>
__device__ __forceinline__ void bitonic_merge_all(uint32_t* keys, uint32_t tid, uint32_t (&v)[NE]) {
  bitonic_stage<NE, K, K / 2>(keys, tid, v);
  if constexpr (K < 1024 * NE) bitonic_merge_all<NE, K * 2>(keys, tid, v);
}

template <uint32_t NE>
__device__ __forceinline__ void bitonic_sort_keys(uint32_t* keys, uint32_t tid) {
  uint32_t v[NE];
#pragma unroll
  for (uint32_t e = 0; e < NE; ++e) v[e] = keys[tid + 1024 * e];
  bitonic_merge_all<NE, 2>(keys, tid, v);
  __syncthreads();
#pragma unroll
  for (uint32_t e = 0; e < NE; ++e) keys[tid + 1024 * e] = v[e];
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Sampler (get_rand + pick_index + cub SortKeys + update_weights,
// device_memory.h:506-545, powerMethod_by_previous.cuh:56-70,
// common.h:42-53).  Slot k of iteration it takes the 64-bit uniform u from
// Philox counter (k>>1, it, trial, kStreamSample), words 2(k&1), 2(k&1)+1;
// t = floor(u * total / 2^64); column = min{j : cdf_j > t} (located in a
// coarse table of the CDF in LDS first, then within its bucket).  One
// workgroup per (segment, iteration) sorts its slots (bitonic, up to eight
// keys per thread: partners in one thread swap in registers, in one wave by
// shuffles, between waves through LDS), merges
// duplicates and writes the unique columns in ascending order with
// weight mult * total / (c * q_j), q_j = cdf_j - cdf_{j-1}.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(1024) void k_sample(const uint64_t* __restrict__ cdf, uint64_t n,
                                                 const DevState* st, int which, uint32_t it0,
                                                 uint32_t buf_it0, uint32_t c,
                                                 uint32_t* __restrict__ idx_out,
                                                 float* __restrict__ wgt_out,
                                                 uint32_t* __restrict__ seg_count) {
  __shared__ uint32_t keys[kSampleSeg];
  __shared__ uint32_t scan[2][1024];
  __shared__ uint64_t cs[kCoarseSample];
  const uint32_t tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  const uint32_t it = it0 + blockIdx.y;
  const uint32_t nseg = (c + kSampleSeg - 1) / kSampleSeg;
  const uint32_t b = g * kSampleSeg;
  const uint32_t cnt = (c - b) < kSampleSeg ? (c - b) : kSampleSeg;
  const uint64_t total = st->total[which];
  const uint64_t last_pos = st->last_pos[which];
  const uint64_t seed = st->seed;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const uint32_t trial = st->trial;
  coarse_fill<kCoarseSample>(cs, cdf, (uint32_t)n, tid, 1024);
  __syncthreads();

  // draw: one Philox call serves two consecutive slots
  for (uint32_t p = tid; 2 * p < kSampleSeg; p += 1024) {
    const uint32_t l0 = 2 * p;
    if (l0 < cnt) {
      const uint32_t k = b + l0;
      uint32_t r[4];
      philox4x32_10(k >> 1, it, trial, kStreamSample, k0, k1, r);
      keys[l0] = bcdf_search<kCoarseSample>(cdf, cs, (uint32_t)n, draw_target(r[0], r[1], total), last_pos);
      keys[l0 + 1] = (l0 + 1 < cnt)
                         ? bcdf_search<kCoarseSample>(cdf, cs, (uint32_t)n, draw_target(r[2], r[3], total), last_pos)
                         : 0xFFFFFFFFu;
    } else {
      keys[l0] = 0xFFFFFFFFu;
      keys[l0 + 1] = 0xFFFFFFFFu;
    }
  }
  __syncthreads();

  // bitonic sort of the first N = next_pow2(cnt) keys (the rest are
  // sentinels), NE = max(1, N / 1024) keys per thread
  uint32_t N = 1;
  while (N < cnt) N <<= 1;
  switch (N > 1024 ? N / 1024 : 1) {  // sentinels beyond N sort last
    case 1: bitonic_sort_keys<1>(keys, tid); break;
    case 2: bitonic_sort_keys<2>(keys, tid); break;
    case 4: bitonic_sort_keys<4>(keys, tid); break;
    default: bitonic_sort_keys<8>(keys, tid); break;
  }

  // run starts: thread tid owns positions [tid*8, tid*8+8)
  constexpr uint32_t E = kSampleSeg / 1024;
  uint32_t flags = 0, nf = 0;
#pragma unroll
  for (uint32_t q = 0; q < E; ++q) {
    const uint32_t p = tid * E + q;
    const bool f = p < cnt && (p == 0 || keys[p] != keys[p - 1]);
    flags |= (f ? 1u : 0u) << q;
    nf += f ? 1u : 0u;
  }
  scan[0][tid] = nf;
  __syncthreads();
  int cur = 0;
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    const uint32_t v = tid >= off ? scan[cur][tid] + scan[cur][tid - off] : scan[cur][tid];
    scan[cur ^ 1][tid] = v;
    cur ^= 1;
    __syncthreads();
  }
  uint32_t pos = scan[cur][tid] - nf;  // exclusive prefix
  const uint64_t base = (uint64_t)(it - buf_it0) * c + b;
  // every run start's multiplicity (LDS) and its two CDF entries first, so
  // the thread's global loads are in flight together
  uint32_t rkey[E], rmult[E];
  uint64_t ca[E], cb[E];
#pragma unroll
  for (uint32_t q = 0; q < E; ++q) {
    rkey[q] = 0u;
    rmult[q] = 0u;
    if (flags & (1u << q)) {
      const uint32_t p = tid * E + q;
      const uint32_t key = keys[p];
      // end of the run: first position in [p, cnt) whose key differs
      uint32_t lo = p, hi = cnt;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] > key) hi = mid; else lo = mid + 1;
      }
      rkey[q] = key;
      rmult[q] = lo - p;
    }
    ca[q] = cdf[rkey[q]];
    cb[q] = cdf[rkey[q] ? rkey[q] - 1 : 0u];
  }
#pragma unroll
  for (uint32_t q = 0; q < E; ++q) {
    if (flags & (1u << q)) {
      const uint32_t key = rkey[q];
      const uint64_t qj = ca[q] - (key ? cb[q] : 0ull);
      const double num = (double)rmult[q] * (double)total;
      const double den = (double)c * (double)qj;
      idx_out[base + pos] = key;
      wgt_out[base + pos] = den > 0.0 ? (float)(num / den) : 0.0f;
      ++pos;
    }
  }
  if (tid == 1023) seg_count[(uint64_t)(it - buf_it0) * nseg + g] = scan[cur][1023];
}

// ---------------------------------------------------------------------------
// Run prologue: zero x_memo / y_memo (zero_clear, device_memory.h:345-355),
// reset the scalar state, stamp the start time and draw x0 in the same
// launch.  Thread m owns rows 2m, 2m + 1.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_begin(float* __restrict__ xm, float* __restrict__ ym,
                                               uint64_t dpad, DevState* st,
                                               uint32_t* __restrict__ mode, uint32_t mode0,
                                               float* __restrict__ y, uint32_t d,
                                               double* __restrict__ BP) {
  __shared__ double red[4];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->t_start = wall_clock64();
    st->cos_prev = 0.0;
    st->last_exact = 0;
    st->sq_n = 0;
    st->status = 0;
    if (mode0) mode[0] = mode0;
  }
  // x0 ~ N(0, I) (curandGenerateNormal, powerMethod_by_exact.cuh:32):
  // Box-Muller on Philox counter (m, 0, trial, kStreamX0) for rows 2m,
  // 2m+1; the raw vector goes to y and per-block sums of squares to BP,
  // from which k_tail forms 1/||y|| and x (normalize<512>, common.h:156-183).
  const uint64_t m = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint64_t seed = st->seed;
  double acc = 0.0;
  if (2 * m < dpad) {  // dpad is even: rows 2m and 2m + 1 both exist
    xm[2 * m] = 0.0f;
    xm[2 * m + 1] = 0.0f;
    ym[2 * m] = 0.0f;
    ym[2 * m + 1] = 0.0f;
    float g0 = 0.0f, g1 = 0.0f;
    if (2 * m < d) {
      uint32_t r[4];
      philox4x32_10((uint32_t)m, 0u, st->trial, kStreamX0, (uint32_t)seed, (uint32_t)(seed >> 32),
                    r);
      const double u1 = uniform53(r[0], r[1]);
      const double u2 = uniform53(r[2], r[3]);
      const double rad = sqrt(-2.0 * log(1.0 - u1));
      const double th = 6.283185307179586 * u2;
      g0 = (float)(rad * cos(th));
      if (2 * m + 1 < d) g1 = (float)(rad * sin(th));
    }
    y[2 * m] = g0;
    y[2 * m + 1] = g1;
    acc = (double)g0 * (double)g0 + (double)g1 * (double)g1;
  }
  acc = block256_sum(acc, red);
  if (threadIdx.x == 0) BP[blockIdx.x] = acc;
}

__device__ __forceinline__ bool gate_open(const uint32_t* mode, uint32_t it, uint32_t mask) {
  return mode == nullptr || ((mask >> mode[it]) & 1u);
}

// Sums of block partials (the finish's y^T x, y^T y, y^T v*: BP[3 i + k];
// the run prologue's sums of squares of x0: BP[i]) over nb blocks, formed by
// one wave in a fixed order: lane l adds blocks l, l + 64, l + 128, ... in
// turn (batches of four in flight; blocks past nb read as 0), then a
// butterfly.  Every wave that forms them gets the same bits, so the tail
// kernel and a dot pass that carries the tail agree without a barrier.
// part_pre / part_post form one sum in the same order, with the first batch
// loaded ahead.
template <int S>
__device__ __forceinline__ double block_part(const double* __restrict__ BP, uint32_t nb, uint32_t i, int k) {
  const double v = BP[S * (i < nb ? i : nb - 1) + k];  // always a valid address
  return i < nb ? v : 0.0;
}

__device__ __forceinline__ double3 finish_sums_wave(const double* __restrict__ BP, uint32_t nb) {
  const uint32_t lane = threadIdx.x & 63u;
  double a = 0.0, b = 0.0, c = 0.0;
  for (uint32_t base = 0; base < nb; base += 256) {
    double v[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < 3; ++k) v[u][k] = block_part<3>(BP, nb, base + lane + 64 * u, k);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a += v[u][0];
      b += v[u][1];
      c += v[u][2];
    }
  }
  return make_double3(wave_sum(a), wave_sum(b), wave_sum(c));
}

template <int S, int K>
__device__ __forceinline__ void part_pre(const double* __restrict__ BP, uint32_t nb, double (&p)[4]) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int u = 0; u < 4; ++u) p[u] = block_part<S>(BP, nb, lane + 64 * u, K);
}

template <int S, int K>
__device__ __forceinline__ double part_post(const double* __restrict__ BP, uint32_t nb, const double (&p)[4]) {
  const uint32_t lane = threadIdx.x & 63u;
  double a = 0.0;
#pragma unroll
  for (int u = 0; u < 4; ++u) a += p[u];
  for (uint32_t base = 256; base < nb; base += 256) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = block_part<S>(BP, nb, base + lane + 64 * u, K);
#pragma unroll
    for (int u = 0; u < 4; ++u) a += v[u];
  }
  return wave_sum(a);
}

// the sum of squares of the raw x0 (stride 1), in the same order
__device__ __forceinline__ double x0_sumsq_wave(const double* __restrict__ BP, uint32_t nb) {
  double p[4];
  part_pre<1, 0>(BP, nb, p);
  return part_post<1, 0>(BP, nb, p);
}

// 1/||y|| from y^T y, as every consumer of the finish sums forms it
__device__ __forceinline__ double inv_norm(double yy) { return yy > 0.0 ? 1.0 / sqrt(yy) : 0.0; }

// ---------------------------------------------------------------------------
// Contraction 1 (previousAMM_kernel1 / proposedAMM_kernel1,
// core_previousAMM.cuh:4-175, core_proposedAMM.cuh:17-213): for list entry e,
// P1[r][e] = a_{col(e)}[split r]^T (x - x_memo)[split r].  One wave per CPW
// entries; lanes stride down the column with 16-byte loads; wave butterfly.
// LIST: col(e) = idx[e], valid when (e % seg) < seg_count[e / seg];
// dense: col(e) = e (exact step, sanity mode).
// FUSE: the previous iteration's tail rides in this pass (GatherTail; list
// mode with x_memo, d a multiple of the split): every wave
// reduces the finish partials to 1/||y|| and forms x = (float)(y / ||y||) as
// k_tail forms it; one extra workgroup per split writes x over the split's
// rows and, with z, part r of s and q, and split 0's writes lambda, err,
// 1/||y||, cos and the time stamp of the previous iteration.
// ---------------------------------------------------------------------------
template <typename T, int CPW, int U, bool LIST, bool NT, bool FUSE>
__global__ __launch_bounds__(256) void k_gather_dot(
    const T* __restrict__ A, uint64_t ld, uint32_t d, const uint32_t* __restrict__ idx,
    const uint32_t* __restrict__ seg_count, uint32_t cap, const float* __restrict__ x,
    const float* __restrict__ xm, uint32_t rows_per_split, float* __restrict__ P1,
    const uint32_t* __restrict__ mode, uint32_t it, uint32_t gate_mask, GatherTail gt) {
  if (!gate_open(mode, it, gate_mask)) return;
  constexpr int V = Elem<T>::VEC;
  const int lane = threadIdx.x & 63;
  const uint32_t e0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * CPW;
  const uint32_t r = blockIdx.y;
  const uint32_t rb = r * rows_per_split;
  const uint32_t re = (rb + rows_per_split < d) ? rb + rows_per_split : d;
  if constexpr (FUSE) {
    if (blockIdx.x == gridDim.x - 1) {
      // the extra workgroup of split r: x over the split's rows (V rows per
      // thread: the split is 256 V rows), with z its part of s and q; split
      // 0's also the scalars of the previous iteration
      __shared__ double red[2][4];
      const uint32_t row = rb + threadIdx.x * V;
      float yv[V], mv[V], zv[V];
      load_vecf<V>(gt.y + row, yv);
      load_vecf<V>(xm + row, mv);
      if (gt.z) load_vecf<V>(gt.z + row, zv);
      const double3 f = gt.x0 ? make_double3(0.0, x0_sumsq_wave(gt.BP, gt.nb), 0.0)
                              : finish_sums_wave(gt.BP, gt.nb);
      const double scale = inv_norm(f.y);
      double ps = 0.0, pq = 0.0;
#pragma unroll
      for (int k = 0; k < V; ++k) yv[k] = (float)(scale * (double)yv[k]);
      store_vecf<V>(gt.x + row, yv);
      if (gt.z) {
#pragma unroll
        for (int k = 0; k < V; ++k) {
          const double dd = (double)yv[k] - (double)mv[k];
          ps += dd;
          pq += (double)zv[k] * dd;
        }
        ps = wave_sum(ps);
        pq = wave_sum(pq);
        if (lane == 0) {
          red[0][threadIdx.x >> 6] = ps;
          red[1][threadIdx.x >> 6] = pq;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
          gt.st->sq_part[r][0] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
          gt.st->sq_part[r][1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        }
      }
      if (r == 0 && threadIdx.x == 0) {
        if (gt.z) gt.st->sq_n = gridDim.y;
        gt.st->scale = scale;
        if (!gt.x0) {
          gt.eig[gt.it] = f.x;
          gt.err[gt.it] = f.z * scale;
          gt.st->cos_prev = f.x * scale;
          if (gt.exact) gt.st->last_exact = gt.it;
          if (gt.ts) gt.ts[gt.it] = wall_clock64();
        }
      }
      return;
    }
  }
  // FUSE: y^T y of the previous product, its first loads issued beside the
  // sample list's (summed once the columns are in flight)
  double yy_pre[4];
  if constexpr (FUSE) {
    if (gt.x0)
      part_pre<1, 0>(gt.BP, gt.nb, yy_pre);
    else
      part_pre<3, 1>(gt.BP, gt.nb, yy_pre);
  }
  const T* col[CPW];
  bool valid[CPW];
  bool any = false;
#pragma unroll
  for (int c = 0; c < CPW; ++c) {
    const uint32_t e = e0 + c;
    bool v = e < cap;
    uint32_t j = e;
    if (LIST && v) {
      // idx[e] is in bounds for every e < cap: loaded beside the segment
      // count rather than after it, and used only where the entry is valid
      const uint32_t jr = idx[e];
      v = (e % kSampleSeg) < seg_count[e / kSampleSeg];
      j = jr;
    }
    valid[c] = v;
    col[c] = A + (uint64_t)(v ? j : 0) * ld;
    any |= v;
  }
  if (!any) return;
  float acc[CPW];
#pragma unroll
  for (int c = 0; c < CPW; ++c) acc[c] = 0.0f;
  constexpr uint32_t STEP = 64u * V;
  // FUSE: every split is exactly one trip (the host carries the tail only
  // where d is a multiple of the split), so each wave makes one trip: the
  // loads of y, x_memo (z for the x writer) and the columns, then x =
  // (float)(y / ||y||) as k_tail forms it, dx = x - x_memo, the FMAs
  if constexpr (FUSE) {
    const uint32_t row = rb + lane * V;
    float ys[U][V], ms[U][V];
#pragma unroll
    for (int u = 0; u < U; ++u) load_vecf<V>(gt.y + row + u * STEP, ys[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) load_vecf<V>(xm + row + u * STEP, ms[u]);
    float a[CPW][U][V];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      if (valid[c]) {
#pragma unroll
        for (int u = 0; u < U; ++u) Elem<T>::template load_full<NT>(col[c] + row + u * STEP, a[c][u]);
      }
    }
    const double scale =
        inv_norm(gt.x0 ? part_post<1, 0>(gt.BP, gt.nb, yy_pre) : part_post<3, 1>(gt.BP, gt.nb, yy_pre));
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int k = 0; k < V; ++k) ys[u][k] = (float)(scale * (double)ys[u][k]) - ms[u][k];
    }
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      if (valid[c]) {
        float s0 = 0.0f;
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
          for (int k = 0; k < V; ++k) s0 = fmaf(a[c][u][k], ys[u][k], s0);
        }
        acc[c] = s0;
      }
    }
  } else {
    uint32_t row = rb + lane * V;
    // U wave-steps per trip: all loads of the trip issued before the FMAs
    for (; row + (U - 1) * STEP + V <= re; row += U * STEP) {
      float xs[U][V];
#pragma unroll
      for (int u = 0; u < U; ++u) load_vecf<V>(x + row + u * STEP, xs[u]);
      if (xm) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          float m[V];
          load_vecf<V>(xm + row + u * STEP, m);
#pragma unroll
          for (int k = 0; k < V; ++k) xs[u][k] -= m[k];
        }
      }
      float a[CPW][U][V];
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        if (valid[c]) {
#pragma unroll
          for (int u = 0; u < U; ++u) Elem<T>::template load_full<NT>(col[c] + row + u * STEP, a[c][u]);
        }
      }
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        if (valid[c]) {
          float s0 = acc[c];
#pragma unroll
          for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int k = 0; k < V; ++k) s0 = fmaf(a[c][u][k], xs[u][k], s0);
          }
          acc[c] = s0;
        }
      }
    }
    for (; row < re; row += STEP) {
      float x0[V];
      load_vecf<V>(x + row, x0);
      if (xm) {
        float m0[V];
        load_vecf<V>(xm + row, m0);
#pragma unroll
        for (int k = 0; k < V; ++k) x0[k] -= m0[k];
      }
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        if (valid[c]) {
          float a0[V];
          load_col<T>(col[c], row, re, a0);
#pragma unroll
          for (int k = 0; k < V; ++k) acc[c] = fmaf(a0[k], x0[k], acc[c]);
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CPW; ++c) {
    const float s = wave_sum(acc[c]);
    if (lane == 0 && valid[c]) P1[(uint64_t)r * cap + e0 + c] = s;
  }
}

// sum_r P1[r][e] over the R row splits in split order (FP64), the loads of
// each batch of 8 issued together (a loop with a bound per element keeps one
// load in flight at a time)
__device__ __forceinline__ double split_sum(const float* __restrict__ P1, uint32_t R, uint32_t cap, uint32_t e) {
  double rr = 0.0;
  uint32_t q = 0;
  for (; q + 8 <= R; q += 8) {
    float p[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) p[u] = P1[(uint64_t)(q + u) * cap + e];
#pragma unroll
    for (int u = 0; u < 8; ++u) rr += (double)p[u];
  }
  for (; q < R; ++q) rr += (double)P1[(uint64_t)q * cap + e];
  return rr;
}

// ---------------------------------------------------------------------------
// Contraction 2 (previousAMM_kernel2 / proposedAMM_kernel2,
// core_previousAMM.cuh:177-263, core_proposedAMM.cuh:215-313): workgroup
// (row block, chunk s) forms t_e for the chunk's entries in LDS and
// accumulates P2[s][rows] = sum_e a_{col(e)}[rows] t_e.  The proposed
// mean shift (a_e - alpha_e 1) t_e is split into the raw product plus the
// scalar sum_e alpha_e t_e, written per chunk to P2s[s].
//   t_e = wgt_e * (sum_r P1[r][e] - alpha_col * s)     (P1 source)
//   t_e = tdirect[col]                                  (direct, z = A alpha)
// ---------------------------------------------------------------------------
constexpr uint32_t kMaxChunk = 2048;

template <typename T, bool LIST, bool NT>
__global__ __launch_bounds__(256) void k_gather_axpy(
    const T* __restrict__ A, uint64_t ld, uint32_t d, uint64_t dpad,
    const uint32_t* __restrict__ idx, const float* __restrict__ wgt,
    const uint32_t* __restrict__ seg_count, uint32_t cap, uint32_t cs, const float* __restrict__ P1,
    uint32_t R, const float* __restrict__ tdirect, const float* __restrict__ alpha,
    const DevState* __restrict__ st, float* __restrict__ P2, double* __restrict__ P2s,
    const uint32_t* __restrict__ mode, uint32_t it, uint32_t gate_mask) {
  if (!gate_open(mode, it, gate_mask)) return;
  constexpr int V = Elem<T>::VEC;
  __shared__ float tl[kMaxChunk];
  __shared__ uint32_t cl[kMaxChunk];
  __shared__ double red[4];
  const uint32_t s = blockIdx.y;
  const uint32_t e0 = s * cs;
  uint32_t e1 = e0 + cs < cap ? e0 + cs : cap;
  if (LIST && e0 < cap) {
    const uint32_t g = e0 / kSampleSeg;
    const uint32_t vend = g * kSampleSeg + seg_count[g];
    e1 = e1 < vend ? e1 : vend;
  }
  const uint32_t nv = e1 > e0 ? e1 - e0 : 0;
  const double sx = alpha ? sq_total(st).x : 0.0;
  double corr = 0.0;
  for (uint32_t k = threadIdx.x; k < nv; k += 256) {
    const uint32_t e = e0 + k;
    const uint32_t j = LIST ? idx[e] : e;
    float t;
    if (tdirect) {
      t = tdirect[j];
    } else {
      const double aj = alpha ? (double)alpha[j] : 0.0;
      const double wj = (LIST && wgt) ? (double)wgt[e] : 1.0;
      const double rr = split_sum(P1, R, cap, e);
      const double tt = wj * (rr - aj * sx);
      t = (float)tt;
      corr += aj * (double)t;
    }
    tl[k] = t;
    cl[k] = j;
  }
  if (alpha && blockIdx.x == 0) {
    corr = block256_sum(corr, red);
    if (threadIdx.x == 0) P2s[s] = corr;
  }
  __syncthreads();
  const uint64_t row64 = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * V;
  if (row64 >= dpad) return;
  const uint32_t row = (uint32_t)row64;
  float acc[V];
#pragma unroll
  for (int k = 0; k < V; ++k) acc[k] = 0.0f;
  const T* Ar = A + row;
  if (row + V <= d) {
    uint32_t k = 0;
    for (; k + 8 <= nv; k += 8) {
      float a[8][V];
#pragma unroll
      for (int u = 0; u < 8; ++u) Elem<T>::template load_full<NT>(Ar + (uint64_t)cl[k + u] * ld, a[u]);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const float t = tl[k + u];
#pragma unroll
        for (int q = 0; q < V; ++q) acc[q] = fmaf(a[u][q], t, acc[q]);
      }
    }
    for (; k < nv; ++k) {
      float a[V];
      Elem<T>::template load_full<NT>(Ar + (uint64_t)cl[k] * ld, a);
      const float t = tl[k];
#pragma unroll
      for (int q = 0; q < V; ++q) acc[q] = fmaf(a[q], t, acc[q]);
    }
  } else {
    for (uint32_t k = 0; k < nv; ++k) {
      float a[V];
      load_col<T>(A + (uint64_t)cl[k] * ld, row, d, a);
      const float t = tl[k];
#pragma unroll
      for (int q = 0; q < V; ++q) acc[q] = fmaf(a[q], t, acc[q]);
    }
  }
  store_vecf<V>(P2 + (uint64_t)s * dpad + row, acc);
}

// ---------------------------------------------------------------------------
// Finish one product and its Rayleigh quotient (dotProduct_and_generate_x,
// common.h:89-154): y = base + sum_s P2[s] with
//   exact:    base = 0
//   previous: base = y_memo
//   proposed: base = y_memo + s v + q - sum_s P2s[s]
// (core_proposedAMM.cuh:198-212), then per-block FP64 partials of y^T x,
// y^T y, y^T v* (k_tail finishes them).  Exact steps with memo snapshot
// [x, y] into [x_memo, y_memo] (powerMethod_by_previous.cuh:157).
// ---------------------------------------------------------------------------
constexpr uint32_t kFinishRows = 64;  // rows per finish block

__global__ __launch_bounds__(256) void k_finish(
    const float* __restrict__ P2, uint32_t S_exact, uint32_t S_amm, uint64_t dpad, uint32_t d,
    const float* __restrict__ x, float* __restrict__ y, float* __restrict__ xm,
    float* __restrict__ ym, const float* __restrict__ v, const double* __restrict__ P2s,
    const double* __restrict__ eigvec, double* __restrict__ BP, const DevState* st,
    const uint32_t* __restrict__ mode, uint32_t it, uint32_t gate_mask) {
  __shared__ double part[16][kFinishRows];
  __shared__ double red[4];
  __shared__ double corr_l;
  const uint32_t md = mode[it];
  if (gate_mask && !((gate_mask >> md) & 1u)) return;
  const bool prop = md == kModeAmmProposed;
  const bool memo_base = md == kModeAmmPrevious || md == kModeAmmProposed;
  const uint32_t S = memo_base ? S_amm : S_exact;
  if (prop) {
    double c = 0.0;
    for (uint32_t i = threadIdx.x; i < S; i += 256) c += P2s[i];
    c = block256_sum(c, red);
    if (threadIdx.x == 0) corr_l = c;
    __syncthreads();
  }
  const double2 sq = prop ? sq_total(st) : make_double2(0.0, 0.0);
  const double sx = sq.x, qx = sq.y;
  const double corr = prop ? corr_l : 0.0;
  // 16 groups of 16 lanes; group g sums chunks g, g+16, ... for 64 rows
  {
    const uint32_t g = threadIdx.x >> 4, l16 = threadIdx.x & 15;
    const uint64_t r0 = (uint64_t)blockIdx.x * kFinishRows + l16 * 4;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    if (r0 < dpad) {
      // batches of 16 partials with every load issued before the sums (a
      // loop with a bound per element gets one load in flight at a time);
      // the sums keep the order s = g, g + 16, ...
      uint32_t s = g;
      for (; s + 15 * 16 < S; s += 16 * 16) {
        float4 p[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) p[u] = *reinterpret_cast<const float4*>(P2 + (uint64_t)(s + 16 * u) * dpad + r0);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          a0 += p[u].x;
          a1 += p[u].y;
          a2 += p[u].z;
          a3 += p[u].w;
        }
      }
      // the remaining (at most 15) partials: every load issued before the
      // sums; absent ones read as +0, which leaves each sum unchanged
      {
        float4 p[15];
#pragma unroll
        for (int u = 0; u < 15; ++u)
          p[u] = (s + 16 * u < S) ? *reinterpret_cast<const float4*>(P2 + (uint64_t)(s + 16 * u) * dpad + r0)
                                  : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int u = 0; u < 15; ++u) {
          a0 += p[u].x;
          a1 += p[u].y;
          a2 += p[u].z;
          a3 += p[u].w;
        }
      }
    }
    part[g][l16 * 4 + 0] = a0;
    part[g][l16 * 4 + 1] = a1;
    part[g][l16 * 4 + 2] = a2;
    part[g][l16 * 4 + 3] = a3;
  }
  __syncthreads();
  double yx = 0.0, yy = 0.0, yv = 0.0;
  const uint64_t row = (uint64_t)blockIdx.x * kFinishRows + threadIdx.x;
  if (threadIdx.x < kFinishRows && row < dpad) {
    double acc = memo_base ? (double)ym[row] : 0.0;
    if (prop) acc += sx * (double)v[row] + qx - corr;
#pragma unroll
    for (int g = 0; g < 16; ++g) acc += part[g][threadIdx.x];
    const float yf = row < d ? (float)acc : 0.0f;
    const double yd = yf;
    const float xr = x[row];
    yx = yd * (double)xr;
    yy = yd * yd;
    if (eigvec && row < d) yv = yd * eigvec[row];
    y[row] = yf;
    if (md == kModeExactMemo) {
      xm[row] = xr;
      ym[row] = yf;
    }
  }
  yx = block256_sum(yx, red);
  yy = block256_sum(yy, red);
  yv = block256_sum(yv, red);
  if (threadIdx.x == 0) {
    BP[3 * blockIdx.x + 0] = yx;
    BP[3 * blockIdx.x + 1] = yy;
    BP[3 * blockIdx.x + 2] = yv;
  }
}

// ---------------------------------------------------------------------------
// Second contraction of a sampled product with the finish fused in
// (previousAMM_kernel2 / proposedAMM_kernel2, core_previousAMM.cuh:177-263,
// core_proposedAMM.cuh:215-313, then the y half of dotProduct_and_generate_x,
// common.h:89-154).  Workgroup b owns rows [64 b, 64 b + 64) and every entry
// of the list, so y is complete for its rows without partials in memory:
//   t_e = wgt_e (sum_r P1[r][e] - alpha_col s)      (chunks of kMaxChunk in LDS)
//   group g (64 / V lanes, 16-byte loads of 64 rows) accumulates entries
//   g, g + G, ... (G = 1024 V / 64 groups) in FP32;
//   the groups' sums are added in FP64 in a fixed order (16 parts of G/16
//   groups, then the parts), y = y_memo (+ s v + q - sum_e alpha_e t_e) + sum;
//   block partials of y^T x, y^T y, y^T v* as k_finish writes them.
// The list's column indices and t stay in LDS; A is read once per entry and
// row, as in the chunked form, and nothing of size d x chunks is written.
// ---------------------------------------------------------------------------
constexpr uint32_t kAfRows = 64;  // rows per workgroup (= kFinishRows: same block partials)

template <typename T, bool NT>
__global__ __launch_bounds__(1024) void k_axpy_finish(
    const T* __restrict__ A, uint64_t ld, uint32_t d, uint64_t dpad,
    const uint32_t* __restrict__ idx, const float* __restrict__ wgt,
    const uint32_t* __restrict__ seg_count, uint32_t cap, const float* __restrict__ P1, uint32_t R,
    const float* __restrict__ alpha, const DevState* __restrict__ st, const float* __restrict__ x,
    float* __restrict__ y, const float* __restrict__ ym, const float* __restrict__ v,
    const double* __restrict__ eigvec, double* __restrict__ BP, const uint32_t* __restrict__ mode,
    uint32_t it, uint32_t gate_mask) {
  if (!gate_open(mode, it, gate_mask)) return;
  constexpr int V = Elem<T>::VEC;
  constexpr uint32_t LPG = kAfRows / V;  // lanes per group: 16 (FP32), 8 (BF16)
  constexpr uint32_t G = 1024 / LPG;     // groups: 64 (FP32), 128 (BF16)
  constexpr uint32_t GP = G / 16;        // groups per part of the fixed-order sum
  __shared__ float tl[kMaxChunk];
  __shared__ uint32_t cl[kMaxChunk];
  __shared__ float gsum[G][kAfRows];
  __shared__ double psum[16][kAfRows];
  __shared__ double red[16];
  const uint32_t tid = threadIdx.x;
  const uint32_t g = tid / LPG, l = tid % LPG;
  const uint64_t row64 = (uint64_t)blockIdx.x * kAfRows + l * V;
  const bool live = row64 < dpad;
  const uint32_t row = live ? (uint32_t)row64 : 0u;
  const double2 sq = alpha ? sq_total(st) : make_double2(0.0, 0.0);
  const double sx = sq.x;
  double corr = 0.0;
  float acc[V];
#pragma unroll
  for (int q = 0; q < V; ++q) acc[q] = 0.0f;
  const T* Ar = A + row;
  const bool full = live && row + V <= d;
  for (uint32_t base = 0; base < cap; base += kMaxChunk) {
    // a chunk lies inside one sample segment (kMaxChunk divides kSampleSeg)
    const uint32_t sg = base / kSampleSeg;
    const uint32_t ce = min(base + kMaxChunk, cap);
    __syncthreads();  // the previous chunk's t and columns are consumed
    // every entry of the chunk up to cap (in bounds) is loaded without
    // waiting for the segment's count; entries at or beyond it go unused
    for (uint32_t k = tid; k < ce - base; k += 1024) cl[k] = idx[base + k];
    const uint32_t vend = sg * kSampleSeg + seg_count[sg];
    const uint32_t e1 = min(ce, vend);
    const uint32_t nv = e1 > base ? e1 - base : 0u;
    __syncthreads();
    // the group's first 8 entries are loaded before t is formed, so the
    // reads of A overlap the reads of the dot partials (absent entries: 0)
    float a0[8][V];
    if (full) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t k = g + u * G;
        if (k < nv) {
          Elem<T>::template load_full<NT>(Ar + (uint64_t)cl[k] * ld, a0[u]);
        } else {
#pragma unroll
          for (int q = 0; q < V; ++q) a0[u][q] = 0.0f;
        }
      }
    }
    for (uint32_t k = tid; k < nv; k += 1024) {
      const uint32_t e = base + k;
      const uint32_t j = cl[k];
      // the per-entry loads first: they do not wait for the partial sums
      const double aj = alpha ? (double)alpha[j] : 0.0;
      const double wj = (double)wgt[e];
      const double rr = split_sum(P1, R, cap, e);
      const float t = (float)(wj * (rr - aj * sx));
      corr += aj * (double)t;
      tl[k] = t;
    }
    __syncthreads();
    if (full) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t k = g + u * G;
        const float t = k < nv ? tl[k] : 0.0f;
#pragma unroll
        for (int q = 0; q < V; ++q) acc[q] = fmaf(a0[u][q], t, acc[q]);
      }
      uint32_t k = g + 8 * G;
      for (; k + 7 * G < nv; k += 8 * G) {
        float a[8][V];
#pragma unroll
        for (int u = 0; u < 8; ++u) Elem<T>::template load_full<NT>(Ar + (uint64_t)cl[k + u * G] * ld, a[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const float t = tl[k + u * G];
#pragma unroll
          for (int q = 0; q < V; ++q) acc[q] = fmaf(a[u][q], t, acc[q]);
        }
      }
      for (; k < nv; k += G) {
        float a[V];
        Elem<T>::template load_full<NT>(Ar + (uint64_t)cl[k] * ld, a);
        const float t = tl[k];
#pragma unroll
        for (int q = 0; q < V; ++q) acc[q] = fmaf(a[q], t, acc[q]);
      }
    } else if (live) {
      for (uint32_t k = g; k < nv; k += G) {
        float a[V];
        load_col<T>(A + (uint64_t)cl[k] * ld, row, d, a);
        const float t = tl[k];
#pragma unroll
        for (int q = 0; q < V; ++q) acc[q] = fmaf(a[q], t, acc[q]);
      }
    }
  }
  // proposed: sum_e alpha_e t_e over the whole list (every workgroup forms it)
  if (alpha) {
    corr = wave_sum(corr);
    if ((tid & 63) == 0) red[tid >> 6] = corr;
  }
#pragma unroll
  for (int q = 0; q < V; ++q) gsum[g][l * V + q] = acc[q];
  __syncthreads();
  {  // part p sums groups [p GP, (p + 1) GP) for row r, in order
    const uint32_t r = tid % kAfRows, p = tid / kAfRows;
    double a = 0.0;
#pragma unroll
    for (uint32_t k = 0; k < GP; ++k) a += (double)gsum[p * GP + k][r];
    psum[p][r] = a;
  }
  __syncthreads();
  if (tid >= 64) return;  // wave 0 finishes its 64 rows
  double corr_all = 0.0;
  if (alpha) {
#pragma unroll
    for (int w = 0; w < 16; ++w) corr_all += red[w];
  }
  const uint64_t r = (uint64_t)blockIdx.x * kAfRows + tid;
  double yx = 0.0, yy = 0.0, yv = 0.0;
  if (r < dpad) {
    double a = (double)ym[r];
    if (alpha) a += sx * (double)v[r] + sq.y - corr_all;
#pragma unroll
    for (int p = 0; p < 16; ++p) a += psum[p][tid];
    const float yf = r < d ? (float)a : 0.0f;
    const double yd = yf;
    yx = yd * (double)x[r];
    yy = yd * yd;
    if (eigvec && r < d) yv = yd * eigvec[r];
    y[r] = yf;
  }
  yx = wave_sum(yx);
  yy = wave_sum(yy);
  yv = wave_sum(yv);
  if (tid == 0) {
    BP[3 * blockIdx.x + 0] = yx;
    BP[3 * blockIdx.x + 1] = yy;
    BP[3 * blockIdx.x + 2] = yv;
  }
}

// ---------------------------------------------------------------------------
// Tail of every grid-wide reduction: sums the block partials in a fixed
// order and updates the scalar state.  A separate launch (rather than a
// last-block-done epilogue) keeps every block free of the device-scope
// fences that cross-XCD visibility would need.  With normalisation it runs
// one block per 1024 rows (at most kTailMaxBlocks), each repeating the
// reduction (a few KB of partials) so that x = y / ||y|| needs no second
// launch; block 0 writes the scalars, and with z every block writes its
// part of s and q (DevState::sq_part, summed by the next product).
//   kTailFinish: lambda_i = y^T x, err_i = y^T v* / ||y||, cos_i, 1/||y||,
//                last exact step, the next step kind (adaptive variance
//                reduction), the time stamp
//   kTailScale:  1/||y|| of the raw x0
//   kTailSQ:     s, q of a stand-alone product or a grid normalisation
// With norm != 0 the blocks then write x = y / ||y|| over their rows and,
// with z given, their parts of s = 1^T (x - x_memo), q = z^T (x - x_memo).
// ---------------------------------------------------------------------------
//   kTailNorm:   nothing to reduce; normalise with the stored 1/||y||
// (TailOp is declared in amm_launch.hpp.)

__global__ __launch_bounds__(256) void k_tail(
    uint32_t op, const double* __restrict__ BP, uint32_t nb, DevState* st, int norm,
    const float* __restrict__ y, float* __restrict__ x, const float* __restrict__ xm,
    const float* __restrict__ z, uint32_t d, uint64_t dpad, uint32_t* __restrict__ mode, uint32_t it,
    double* __restrict__ eig, double* __restrict__ err, int64_t* __restrict__ ts, int stamp,
    uint32_t adapt_kernel, uint64_t period, double thr, uint32_t step) {
  __shared__ double red[4];
  __shared__ double red3[3][4];
  __shared__ double scale_l;
  const uint32_t t = threadIdx.x;
  // Every block repeats the (small, fixed-order) reduction of BP and
  // normalises its own rows; block 0 alone writes the scalar state.
  const bool lead = blockIdx.x == 0;
  // The first rows this thread normalises (its only ones where the grid
  // covers dpad, as launch_tail arranges up to kTailMaxBlocks x 1024 rows):
  // their loads do not depend on the reduction, so they go out first.
  const uint64_t r0 = ((uint64_t)blockIdx.x * 256 + t) * 4;
  const bool pre = norm && r0 < dpad;
  float4 y0 = make_float4(0.f, 0.f, 0.f, 0.f), m0 = y0, z0 = y0;
  if (pre) {
    y0 = *reinterpret_cast<const float4*>(y + r0);
    if (z) {
      m0 = *reinterpret_cast<const float4*>(xm + r0);
      z0 = *reinterpret_cast<const float4*>(z + r0);
    }
  }
  if (op == kTailFinish) {
    // wave 0 alone reduces (the order a dot pass carrying this tail repeats)
    double a = 0.0, c = 0.0;
    if (t < 64) {
      const double3 f = finish_sums_wave(BP, nb);
      a = f.x;
      c = f.z;
      if (t == 0) scale_l = inv_norm(f.y);
    }
    if (t == 0 && lead) {
      const uint32_t md = mode[it];
      const double scale = scale_l;
      eig[it] = a;
      err[it] = c * scale;
      st->scale = scale;
      const double cs = a * scale;
      st->cos_prev = cs;
      uint64_t le = st->last_exact;
      if (md == kModeExactMemo || md == kModeExactNoMemo) {
        le = it;
        st->last_exact = le;
      }
      if (adapt_kernel && it + 1 < step) {
        const bool ex = exact_step_due(adapt_kernel, it + 1, period, thr, cs, le);
        mode[it + 1] = ex ? kModeExactMemo : (adapt_kernel == 2 ? kModeAmmProposed : kModeAmmPrevious);
      }
    }
  } else if (op == kTailScale) {
    // wave 0 alone reduces (the order a dot pass carrying this tail repeats)
    if (t < 64) {
      const double scale = inv_norm(x0_sumsq_wave(BP, nb));
      if (t == 0) {
        if (lead) st->scale = scale;
        scale_l = scale;
      }
    }
  } else if (op == kTailSQ) {
    double a = 0.0, b = 0.0;
    for (uint32_t i = t; i < nb; i += 256) {
      a += BP[2 * i];
      b += BP[2 * i + 1];
    }
    a = block256_sum(a, red);
    b = block256_sum(b, red);
    if (t == 0 && lead) {
      st->sq_part[0][0] = a;
      st->sq_part[0][1] = b;
      st->sq_n = 1;
    }
  } else if (op == kTailNorm2) {
    double s = 0.0;
    for (uint32_t i = t; i < nb; i += 256) s += BP[i];
    s = block256_sum(s, red);
    if (t == 0 && lead) st->sigma = sqrt(s);
  } else if (t == 0) {
    scale_l = st->scale;
  }
  __syncthreads();
  if (norm) {
    const double scale = scale_l;
    double ps = 0.0, pq = 0.0;
    for (uint64_t r = r0; r < dpad; r += (uint64_t)gridDim.x * 1024) {
      float4 yv4 = y0, m4 = m0, z4 = z0;
      if (r != r0) {
        yv4 = *reinterpret_cast<const float4*>(y + r);
        if (z) {
          m4 = *reinterpret_cast<const float4*>(xm + r);
          z4 = *reinterpret_cast<const float4*>(z + r);
        }
      }
      float xv[4] = {(float)(scale * (double)yv4.x), (float)(scale * (double)yv4.y),
                     (float)(scale * (double)yv4.z), (float)(scale * (double)yv4.w)};
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (r + k >= d) xv[k] = 0.0f;
      *reinterpret_cast<float4*>(x + r) = make_float4(xv[0], xv[1], xv[2], xv[3]);
      if (z) {
        const float mv[4] = {m4.x, m4.y, m4.z, m4.w};
        const float zv[4] = {z4.x, z4.y, z4.z, z4.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const double dx = (double)xv[k] - (double)mv[k];
          ps += dx;
          pq += (double)zv[k] * dx;
        }
      }
    }
    if (z) {  // this block's part of s and q (both in one exchange, each in block256_sum's order)
      ps = wave_sum(ps);
      pq = wave_sum(pq);
      if ((t & 63) == 0) {
        red3[0][t >> 6] = ps;
        red3[1][t >> 6] = pq;
      }
      __syncthreads();
      if (t == 0) {
        st->sq_part[blockIdx.x][0] = (red3[0][0] + red3[0][1]) + (red3[0][2] + red3[0][3]);
        st->sq_part[blockIdx.x][1] = (red3[1][0] + red3[1][1]) + (red3[1][2] + red3[1][3]);
        if (lead) st->sq_n = gridDim.x;
      }
    }
  }
  if (stamp && t == 0 && lead) ts[it] = wall_clock64();
}

// ---------------------------------------------------------------------------
// x = y / ||y|| over a grid (the generate-x half of
// dotProduct_and_generate_x, common.h:146-153) for large d; with z given,
// per-block partials of s = 1^T (x - x_memo) and q = z^T (x - x_memo) for
// the next proposed product (dots<128>, common.h:55-87), finished by k_tail.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_normalize(const float* __restrict__ y,
                                                   float* __restrict__ x, uint32_t d,
                                                   uint64_t dpad, const DevState* st,
                                                   const float* __restrict__ xm,
                                                   const float* __restrict__ z,
                                                   double* __restrict__ BP) {
  __shared__ double red[4];
  const double scale = st->scale;
  double ps = 0.0, pq = 0.0;
  const uint64_t r0 = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * 4;
  if (r0 < dpad) {
    const float4 yv = *reinterpret_cast<const float4*>(y + r0);
    float xv[4] = {(float)(scale * (double)yv.x), (float)(scale * (double)yv.y),
                   (float)(scale * (double)yv.z), (float)(scale * (double)yv.w)};
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (r0 + k >= d) xv[k] = 0.0f;
    *reinterpret_cast<float4*>(x + r0) = make_float4(xv[0], xv[1], xv[2], xv[3]);
    if (z) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double dx = (double)xv[k] - (double)xm[r0 + k];
        ps += dx;
        pq += (double)z[r0 + k] * dx;
      }
    }
  }
  if (!z) return;
  ps = block256_sum(ps, red);
  pq = block256_sum(pq, red);
  if (threadIdx.x == 0) {
    BP[2 * blockIdx.x] = ps;
    BP[2 * blockIdx.x + 1] = pq;
  }
}

// per-block partials of s = 1^T (x - x_memo), q = z^T (x - x_memo) for a
// stand-alone proposed product (dots<128>, common.h:55-87), finished by k_tail.
__global__ __launch_bounds__(256) void k_dots(const float* __restrict__ x,
                                              const float* __restrict__ xm,
                                              const float* __restrict__ z, uint64_t dpad,
                                              double* __restrict__ BP) {
  __shared__ double red[4];
  double ps = 0.0, pq = 0.0;
  const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (r < dpad) {
    const double dx = (double)x[r] - (double)xm[r];
    ps = dx;
    pq = (double)z[r] * dx;
  }
  ps = block256_sum(ps, red);
  pq = block256_sum(pq, red);
  if (threadIdx.x == 0) {
    BP[2 * blockIdx.x] = ps;
    BP[2 * blockIdx.x + 1] = pq;
  }
}

// z = sum_s P2[s] (z = A alpha, powerMethod_by_proposed.cuh:132-139) and
// v = z - ||alpha||^2 (generate_v, core_proposedAMM.cuh:4-15).  A workgroup
// owns 256 rows (a float4 per lane); its 16 waves take the partials
// s = w, w + 16, ... in FP64 and are added in wave order, so the result
// depends only on S and the partials.
constexpr uint32_t kSumRowsWaves = 16;
__global__ __launch_bounds__(1024) void k_sum_rows(const float* __restrict__ P2, uint32_t S,
                                                   uint64_t dpad, uint32_t d, float* __restrict__ z,
                                                   float* __restrict__ v, const DevState* st) {
  __shared__ double red[kSumRowsWaves][256];
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t r0 = (uint64_t)blockIdx.x * 256 + lane * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (r0 < dpad) {  // dpad is a multiple of 8, so r0 + 4 <= dpad
    // batches of 16 partials per wave, loads first (the sums keep the order
    // s = w, w + waves, ...)
    uint32_t s = w;
    for (; s + 15 * kSumRowsWaves < S; s += 16 * kSumRowsWaves) {
      float4 q[16];
#pragma unroll
      for (int u = 0; u < 16; ++u)
        q[u] = *reinterpret_cast<const float4*>(P2 + (uint64_t)(s + kSumRowsWaves * u) * dpad + r0);
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        acc[0] += (double)q[u].x;
        acc[1] += (double)q[u].y;
        acc[2] += (double)q[u].z;
        acc[3] += (double)q[u].w;
      }
    }
    for (; s < S; s += kSumRowsWaves) {
      const float4 q = *reinterpret_cast<const float4*>(P2 + (uint64_t)s * dpad + r0);
      acc[0] += (double)q.x;
      acc[1] += (double)q.y;
      acc[2] += (double)q.z;
      acc[3] += (double)q.w;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) red[w][lane * 4 + k] = acc[k];
  __syncthreads();
  if (threadIdx.x >= 256) return;
  const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= dpad) return;
  double t = 0.0;
#pragma unroll
  for (uint32_t k = 0; k < kSumRowsWaves; ++k) t += red[k][threadIdx.x];
  const float zf = r < d ? (float)t : 0.0f;
  z[r] = zf;
  if (v) v[r] = r < d ? (float)((double)zf - st->alpha2) : 0.0f;
}

// t_j = sum_r P1[r][j] (A^T x from a split exact pass) and per-block FP64
// partials of ||t||^2 (finished by k_tail, op kTailNorm2).
__global__ __launch_bounds__(256) void k_reduce_p1(const float* __restrict__ P1, uint32_t R,
                                                   uint64_t n, float* __restrict__ t,
                                                   double* __restrict__ BP) {
  __shared__ double red[4];
  const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  double acc2 = 0.0;
  if (j < n) {
    double a = 0.0;
    for (uint32_t r = 0; r < R; ++r) a += (double)P1[(uint64_t)r * n + j];
    const float tf = (float)a;
    t[j] = tf;
    acc2 = (double)tf * (double)tf;
  }
  acc2 = block256_sum(acc2, red);
  if (threadIdx.x == 0) BP[blockIdx.x] = acc2;
}

// v = t / ||t|| with ||t|| from the state (sigma)
__global__ __launch_bounds__(256) void k_unit(const float* __restrict__ t, uint64_t n,
                                              const DevState* st, float* __restrict__ v) {
  const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const double s = st->sigma;
  if (j < n) v[j] = s > 0.0 ? (float)((double)t[j] / s) : 0.0f;
}

// Backward of sigma_max: out[:, j] = g * a * b_j for a column-major d x n
// matrix (the rank-1 adjoint u v^T of d sigma / d A), FP32 products rounded
// once (BF16 output: round to nearest even); 16-byte loads and stores when
// vec (every matrix and vector 16-byte aligned, ld and the strides
// multiples of the vector width).  Matrix k = blockIdx.y of a strided stack
// uses g_k = gs[k] (or g), a + k astride, b + k bstride, out + k ostride.  A wave per column; its lanes stride over
// the column's 16-byte chunks.
template <typename T>
__global__ __launch_bounds__(256) void k_rank1(T* __restrict__ out, uint32_t d, uint64_t ld, uint64_t n,
                                               float g, const float* __restrict__ gs,
                                               const float* __restrict__ a, const float* __restrict__ b,
                                               int vec, uint64_t ostride, uint64_t astride,
                                               uint64_t bstride) {
  constexpr uint32_t V = Elem<T>::VEC;
  const uint32_t cpc = (d + V - 1) / V;  // chunks per column
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t k = blockIdx.y;
  const float gk = gs ? gs[k] : g;
  T* base = out + k * ostride;
  const float* ak = a + k * astride;
  const float* bk = b + k * bstride;
  // wave per column, lanes over the column's 16-byte chunks
  for (uint64_t j = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); j < n; j += (uint64_t)gridDim.x * 4) {
    const float s = gk * bk[j];
    T* colp = base + j * ld;
    for (uint32_t c = lane; c < cpc; c += 64) {
      const uint32_t row = c * V;
      T* col = colp + row;
      float v[V];
      if (vec && row + V <= d) {
#pragma unroll
        for (uint32_t e = 0; e < V; e += 4) {  // 16-byte loads of a (cached: every column reuses it)
          const float4 a4 = *reinterpret_cast<const float4*>(ak + row + e);
          v[e] = s * a4.x;
          v[e + 1] = s * a4.y;
          v[e + 2] = s * a4.z;
          v[e + 3] = s * a4.w;
        }
      } else {
#pragma unroll
        for (uint32_t e = 0; e < V; ++e) v[e] = (row + e < d) ? s * ak[row + e] : 0.0f;
      }
      if (vec && row + V <= d) {
        if constexpr (V == 4) {
          *reinterpret_cast<float4*>(col) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
          uint32_t w[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) w[q] = bf16_rne(v[2 * q]) | ((uint32_t)bf16_rne(v[2 * q + 1]) << 16);
          *reinterpret_cast<uint4*>(col) = make_uint4(w[0], w[1], w[2], w[3]);
        }
      } else {
        for (uint32_t e = 0; e < V && row + e < d; ++e) {
          if constexpr (V == 4)
            reinterpret_cast<float*>(col)[e] = v[e];
          else
            reinterpret_cast<uint16_t*>(col)[e] = bf16_rne(v[e]);
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t val) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = val;
}

// ===========================================================================
// Host launchers
// ===========================================================================
static inline unsigned grid1(uint64_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

// The column-resident passes hold q = L x 64 x VEC rows per wave in
// registers with L a power of two up to maxL: the statistics up to L = 16
// (d <= 16384 FP32, 32768 BF16); the exact product keeps x in 64 KB of LDS
// as well (d <= 16384 either way).
static int resident_L(int dtype, uint32_t d, int maxL) {
  const uint32_t V = dtype == 1 ? 8u : 4u, step = 64u * V;
  const uint32_t q = ((d + 3) / 4 + step - 1) / step * step;
  const uint32_t L = q / step;
  int r = 1;
  while ((uint32_t)r < L) r *= 2;
  return r <= maxL ? r : 0;
}
static int gram_maxL(int dtype) { return dtype == 1 ? 8 : 16; }

// Persistent grid: as many workgroups as stay resident on the device (the
// occupancy of the instance at 256 threads, times the CUs), at most one per
// column.  Cached per kernel and device; the partial count of a pass is this
// number, so it must not vary between calls.
static uint32_t resident_grid(const void* fn, uint64_t n) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, uint32_t> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  uint32_t g = 0;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({fn, dev});
    if (it != cache.end()) {
      g = it->second;
    } else {
      int per = 0, cus = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 256, 0) != hipSuccess || per < 1) per = 1;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 1;
      g = (uint32_t)(per * cus);
      cache[{fn, dev}] = g;
    }
  }
  return (uint32_t)std::min<uint64_t>(n, g);
}

// Workgroups of a column-resident pass (k_gram, k_colstats_z) over columns of
// d rows: as many as are resident, except FP32 columns of 16 KB or more,
// where fewer column streams in flight read HBM faster -- one workgroup per
// CU from 32 KB (16384 x 65536: 655 -> 625 us for the one-read exact
// product, 670 -> 630 us with z; 8192 x 32768: 170 -> 160 and 185 -> 161
// us), two per CU from 16 KB (4096 x 65536: 165 -> 158 and 191 -> 179 us).
// BF16 columns of 8-32 KB measured as fast or faster with every resident
// workgroup.
static uint32_t resident_blocks(const void* fn, int dtype, uint32_t d, uint64_t n) {
  const uint32_t g = resident_grid(fn, n);
  if (dtype == 1) return g;
  const uint64_t col = (uint64_t)d * 4u;
  uint32_t per_cu = col >= 32768 ? 1u : (col >= 16384 ? 2u : 0u);
  if (!per_cu) return g;
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) return g;
  return std::min<uint32_t>(g, (uint32_t)std::min<uint64_t>(n, (uint64_t)per_cu * (uint32_t)cus));
}

template <typename T>
static const void* colstats_z_fn(int L) {
  switch (L) {
    case 1: return (const void*)&k_colstats_z<T, 1>;
    case 2: return (const void*)&k_colstats_z<T, 2>;
    case 4: return (const void*)&k_colstats_z<T, 4>;
    case 8: return (const void*)&k_colstats_z<T, 8>;
    default: return (const void*)&k_colstats_z<T, 16>;
  }
}

template <typename T>
static const void* gram_fn(int L) {
  switch (L) {
    case 1: return (const void*)&k_gram<T, 1>;
    case 2: return (const void*)&k_gram<T, 2>;
    case 4: return (const void*)&k_gram<T, 4>;
    case 8: return (const void*)&k_gram<T, 8>;
    default:
      if constexpr (Elem<T>::VEC == 4) return (const void*)&k_gram<T, 16>;
      return nullptr;
  }
}

bool colstats_fuses_z(int dtype, uint32_t d) { return resident_L(dtype, d, 16) != 0; }

uint32_t colstats_z_blocks(int dtype, uint32_t d, uint64_t n) {
  const int L = resident_L(dtype, d, 16);
  if (!L) return 0;
  return resident_blocks(dtype == 1 ? colstats_z_fn<__hip_bfloat16>(L) : colstats_z_fn<float>(L), dtype, d, n);
}

bool gram_fits(int dtype, uint32_t d) { return resident_L(dtype, d, gram_maxL(dtype)) != 0; }

uint32_t gram_blocks(int dtype, uint32_t d, uint64_t n) {
  const int L = resident_L(dtype, d, gram_maxL(dtype));
  if (!L) return 0;
  return resident_blocks(dtype == 1 ? gram_fn<__hip_bfloat16>(L) : gram_fn<float>(L), dtype, d, n);
}

template <typename T>
static void cs_reg(int L, unsigned g, const void* A, uint64_t ld, uint32_t d, uint64_t n, uint64_t dpad,
                   float* alpha, double* nrm2, double* cnrm2, float* ZP, hipStream_t s) {
  const T* a = (const T*)A;
  switch (L) {
    case 1: k_colstats_z<T, 1><<<g, 256, 0, s>>>(a, ld, d, n, dpad, alpha, nrm2, cnrm2, ZP); break;
    case 2: k_colstats_z<T, 2><<<g, 256, 0, s>>>(a, ld, d, n, dpad, alpha, nrm2, cnrm2, ZP); break;
    case 4: k_colstats_z<T, 4><<<g, 256, 0, s>>>(a, ld, d, n, dpad, alpha, nrm2, cnrm2, ZP); break;
    case 8: k_colstats_z<T, 8><<<g, 256, 0, s>>>(a, ld, d, n, dpad, alpha, nrm2, cnrm2, ZP); break;
    default: k_colstats_z<T, 16><<<g, 256, 0, s>>>(a, ld, d, n, dpad, alpha, nrm2, cnrm2, ZP); break;
  }
}

hipError_t launch_colstats_z(int dtype, const void* A, uint64_t ld, uint32_t d, uint64_t n, uint64_t dpad,
                             float* alpha, double* nrm2, double* cnrm2, float* ZP, DevState* st,
                             hipStream_t s) {
  (void)st;
  const int L = resident_L(dtype, d, 16);
  if (!L || !ZP || n == 0) return hipErrorInvalidValue;  // z cannot be fused at this d
  const unsigned g = colstats_z_blocks(dtype, d, n);
  if (dtype == 1)
    cs_reg<__hip_bfloat16>(L, g, A, ld, d, n, dpad, alpha, nrm2, cnrm2, ZP, s);
  else
    cs_reg<float>(L, g, A, ld, d, n, dpad, alpha, nrm2, cnrm2, ZP, s);
  return hipGetLastError();
}

template <typename T>
static void gram_launch(int L, unsigned g, const void* A, uint64_t ld, uint32_t d, uint64_t n, uint64_t dpad,
                        const float* x, float* P2, const uint32_t* gm, uint32_t it, uint32_t mask,
                        hipStream_t s) {
  const T* a = (const T*)A;
  switch (L) {
    case 1: k_gram<T, 1><<<g, 256, 0, s>>>(a, ld, d, n, dpad, x, P2, gm, it, mask); break;
    case 2: k_gram<T, 2><<<g, 256, 0, s>>>(a, ld, d, n, dpad, x, P2, gm, it, mask); break;
    case 4: k_gram<T, 4><<<g, 256, 0, s>>>(a, ld, d, n, dpad, x, P2, gm, it, mask); break;
    case 8: k_gram<T, 8><<<g, 256, 0, s>>>(a, ld, d, n, dpad, x, P2, gm, it, mask); break;
    default:
      if constexpr (Elem<T>::VEC == 4) k_gram<T, 16><<<g, 256, 0, s>>>(a, ld, d, n, dpad, x, P2, gm, it, mask);
      break;
  }
}

hipError_t launch_gram(int dtype, const void* A, uint64_t ld, uint32_t d, uint64_t n, uint64_t dpad,
                       const float* x, float* P2, const uint32_t* gate_mode, uint32_t it, uint32_t gate_mask,
                       hipStream_t s) {
  const int L = resident_L(dtype, d, gram_maxL(dtype));
  if (!L || !P2 || n == 0) return hipErrorInvalidValue;
  const unsigned g = gram_blocks(dtype, d, n);
  if (dtype == 1)
    gram_launch<__hip_bfloat16>(L, g, A, ld, d, n, dpad, x, P2, gate_mode, it, gate_mask, s);
  else
    gram_launch<float>(L, g, A, ld, d, n, dpad, x, P2, gate_mode, it, gate_mask, s);
  return hipGetLastError();
}

// Columns per workgroup of the statistics pass: about 512 workgroups (two
// per CU) for FP32 columns of 64 KB or more (measured 2-6% faster at d =
// 16384, 32768, 65536), else one workgroup per column (FP32 columns of 16 KB
// and BF16 columns of 32-64 KB measured 5-20% slower blocked).  Without z
// this streamed form is as fast or faster than the column-resident one at
// every shape measured.
static uint32_t colstats_per_block(int dtype, uint32_t d, uint64_t n) {
  if (dtype == 1 || (uint64_t)d * 4u < 65536) return 1;
  return (uint32_t)std::max<uint64_t>(1, n / 512);
}
uint32_t colstats_max_blocks(int dtype, uint32_t d, uint64_t n) {
  if (short_column_dt(dtype, d) || n == 0) return 0;
  const uint32_t K = colstats_per_block(dtype, d, n);
  if (K < 2) return 0;
  const uint64_t g = (n + K - 1) / K;
  return g <= kColstatsMaxBlocks ? (uint32_t)g : 0u;
}

hipError_t launch_colstats(int dtype, const void* A, uint64_t ld, uint32_t d, uint64_t n,
                           float* alpha, double* nrm2, double* cnrm2, DevState* st, uint64_t* wmax,
                           hipStream_t s) {
  (void)st;
  const uint32_t K = colstats_per_block(dtype, d, n);
  const unsigned g = (unsigned)(short_column_dt(dtype, d) ? (n + 3) / 4 : (n + K - 1) / K);
  if (!colstats_max_blocks(dtype, d, n)) wmax = nullptr;
  if (dtype == 1)
    k_colstats<__hip_bfloat16><<<g, 256, 0, s>>>((const __hip_bfloat16*)A, ld, d, n, K, alpha, nrm2, cnrm2,
                                                  wmax);
  else
    k_colstats<float><<<g, 256, 0, s>>>((const float*)A, ld, d, n, K, alpha, nrm2, cnrm2, wmax);
  return hipGetLastError();
}

uint64_t cdf_tiles(uint64_t n) { return grid1(n, kCdfTile); }

hipError_t launch_cdf(const double* w, uint64_t n, uint64_t* cdf, DevState* st, int which,
                      const float* alpha, uint64_t* tile_max, uint64_t* tile_sum, uint64_t* tile_lp,
                      double* tile_a2, const uint64_t* wmax, uint32_t nwmax, hipStream_t s) {
  const unsigned nt = (unsigned)cdf_tiles(n);
  if (wmax && nwmax) {  // the statistics pass left the maxima of its workgroups
    k_cdf_tiles<<<nt, 256, 0, s>>>(w, n, cdf, st, which, alpha, wmax + (size_t)which * nwmax, nwmax, tile_sum,
                                   tile_lp, tile_a2);
  } else {
    k_cdf_max<<<nt, 256, 0, s>>>(w, n, tile_max);
    k_cdf_tiles<<<nt, 256, 0, s>>>(w, n, cdf, st, which, alpha, tile_max, nt, tile_sum, tile_lp, tile_a2);
  }
  k_cdf_offsets<<<nt, 256, 0, s>>>(cdf, n, tile_sum, tile_lp, tile_a2, st, which, alpha ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_sample(const uint64_t* cdf, uint64_t n, const DevState* st,
                         int which, uint32_t it0, uint32_t nit, uint32_t buf_it0, uint32_t c,
                         uint32_t* idx, float* wgt, uint32_t* seg_count, hipStream_t s) {
  if (nit == 0) return hipSuccess;
  const dim3 g((c + kSampleSeg - 1) / kSampleSeg, nit);
  k_sample<<<g, 1024, 0, s>>>(cdf, n, st, which, it0, buf_it0, c, idx, wgt, seg_count);
  return hipGetLastError();
}


hipError_t launch_begin(float* xm, float* ym, uint64_t dpad, DevState* st, uint32_t* mode,
                        uint32_t mode0, float* y, uint32_t d, double* BP, hipStream_t s) {
  k_begin<<<(unsigned)x0_blocks(dpad), 256, 0, s>>>(xm, ym, dpad, st, mode, mode0, y, d, BP);
  return hipGetLastError();
}

// Dense passes stream all of A once (non-temporal loads).  A sampled list
// reads its columns twice in an iteration (dot, then axpy) and heavy
// columns recur across iterations, so the list passes load non-temporally
// only where a.nt says so.
template <typename T, int CPW, int U, bool LIST, bool NT>
static void gd(const GatherDotArgs& a, hipStream_t s) {
  const dim3 g((a.cap + 4 * CPW - 1) / (4 * CPW), a.R);
  if constexpr (LIST) {
    if (a.tail) {  // one more workgroup per split: the x writer
      const dim3 gw(g.x + 1, g.y);
      k_gather_dot<T, CPW, U, LIST, NT, true><<<gw, 256, 0, s>>>(
          (const T*)a.A, a.ld, a.d, a.idx, a.seg_count, a.cap, a.x, a.xm, a.rows_per_split, a.P1,
          a.mode, a.it, a.gate_mask, *a.tail);
      return;
    }
  }
  k_gather_dot<T, CPW, U, LIST, NT, false><<<g, 256, 0, s>>>((const T*)a.A, a.ld, a.d, a.idx,
                                                             a.seg_count, a.cap, a.x, a.xm,
                                                             a.rows_per_split, a.P1, a.mode, a.it,
                                                             a.gate_mask, GatherTail{});
}

template <typename T>
static void gd_any(const GatherDotArgs& a, hipStream_t s) {
  // dense: 4 columns per wave (each x load serves 4 columns), two steps per
  // trip; measured faster than a workgroup per column or per 4 columns
  // split by rows.  List: 2 columns per wave and kListDotSteps steps per
  // trip, with the plan's splits sized so that a wave makes one trip.
  if (!a.idx) gd<T, 4, 2, false, true>(a, s);
  else if (a.nt) gd<T, 2, kListDotSteps, true, true>(a, s);
  else gd<T, 2, kListDotSteps, true, false>(a, s);
}

hipError_t launch_gather_dot(int dtype, const GatherDotArgs& a, hipStream_t s) {
  // the carried tail needs whole one-trip splits and writes one s, q part per
  // split (at most kTailMaxBlocks)
  if (a.tail && (!a.idx || !a.xm || a.R > kTailMaxBlocks || a.rows_per_split == 0 ||
                 a.d % a.rows_per_split != 0 ||
                 a.rows_per_split != 64u * (dtype == 1 ? 8u : 4u) * (uint32_t)kListDotSteps))
    return hipErrorInvalidValue;
  if (dtype == 1) gd_any<__hip_bfloat16>(a, s);
  else gd_any<float>(a, s);
  return hipGetLastError();
}

template <typename T, bool LIST, bool NT>
static void ga(const GatherAxpyArgs& a, hipStream_t s) {
  constexpr int V = Elem<T>::VEC;
  const dim3 g(grid1(a.dpad, 256 * V), a.S);
  k_gather_axpy<T, LIST, NT><<<g, 256, 0, s>>>((const T*)a.A, a.ld, a.d, a.dpad, a.idx, a.wgt,
                                               a.seg_count, a.cap, a.cs, a.P1, a.R, a.tdirect,
                                               a.alpha, a.st, a.P2, a.P2s, a.mode, a.it,
                                               a.gate_mask);
}

template <typename T>
static void ga_any(const GatherAxpyArgs& a, hipStream_t s) {
  if (!a.idx) ga<T, false, true>(a, s);
  else if (a.nt) ga<T, true, true>(a, s);
  else ga<T, true, false>(a, s);
}

hipError_t launch_gather_axpy(int dtype, const GatherAxpyArgs& a, hipStream_t s) {
  if (a.cs == 0 || a.cs > kMaxChunk) return hipErrorInvalidValue;
  if (dtype == 1) ga_any<__hip_bfloat16>(a, s);
  else ga_any<float>(a, s);
  return hipGetLastError();
}

hipError_t launch_finish(const FinishArgs& a, hipStream_t s) {
  k_finish<<<grid1(a.dpad, kFinishRows), 256, 0, s>>>(a.P2, a.S_exact, a.S_amm, a.dpad, a.d, a.x,
                                                       a.y, a.xm, a.ym, a.v, a.P2s, a.eigvec, a.BP,
                                                       a.st, a.mode, a.it, a.gate_mask);
  return hipGetLastError();
}

bool axpy_finish_applies(uint64_t dpad) { return grid1(dpad, kAfRows) >= kAfMinBlocks; }

template <typename T, bool NT>
static void af(const AxpyFinishArgs& a, hipStream_t s) {
  k_axpy_finish<T, NT><<<grid1(a.dpad, kAfRows), 1024, 0, s>>>(
      (const T*)a.A, a.ld, a.d, a.dpad, a.idx, a.wgt, a.seg_count, a.cap, a.P1, a.R, a.alpha, a.st, a.x,
      a.y, a.ym, a.v, a.eigvec, a.BP, a.mode, a.it, a.gate_mask);
}

hipError_t launch_axpy_finish(int dtype, const AxpyFinishArgs& a, hipStream_t s) {
  if (!a.idx || !a.wgt || !a.seg_count || !a.ym) return hipErrorInvalidValue;
  if (dtype == 1) {
    if (a.nt) af<__hip_bfloat16, true>(a, s);
    else af<__hip_bfloat16, false>(a, s);
  } else {
    if (a.nt) af<float, true>(a, s);
    else af<float, false>(a, s);
  }
  return hipGetLastError();
}

uint64_t finish_blocks(uint64_t dpad) { return grid1(dpad, kFinishRows); }
uint64_t x0_blocks(uint64_t dpad) { return grid1(dpad / 2, 256); }
uint64_t normalize_blocks(uint64_t dpad) { return grid1(dpad, 1024); }
uint64_t dots_blocks(uint64_t dpad) { return grid1(dpad, 256); }

hipError_t launch_tail(const TailArgs& a, hipStream_t s) {
  // normalisation spreads over blocks of 1024 rows (each block repeats the
  // reduction and writes its own part of the s, q dots)
  unsigned g = 1;
  if (a.norm && (a.op == kTailFinish || a.op == kTailScale || a.op == kTailNorm))
    g = (unsigned)std::min<uint64_t>(std::max<uint64_t>(1, (a.dpad + 1023) / 1024), kTailMaxBlocks);
  k_tail<<<g, 256, 0, s>>>(a.op, a.BP, a.nb, a.st, a.norm, a.y, a.x, a.xm, a.z, a.d, a.dpad, a.mode,
                           a.it, a.eig, a.err, a.ts, a.stamp, a.adapt_kernel, a.period, a.thr, a.step);
  return hipGetLastError();
}

hipError_t launch_read(const void* A, uint64_t bytes, uint32_t* sink, hipStream_t s) {
  const uint64_t n4 = bytes / 16;
  if (n4 == 0) return hipSuccess;
  const uint64_t per_block = 65536;  // 1 MiB per workgroup
  const uint64_t blocks = (n4 + per_block - 1) / per_block;
  k_read<<<(unsigned)blocks, 256, 0, s>>>(reinterpret_cast<const u32x4*>(A), n4, per_block, sink);
  return hipGetLastError();
}

hipError_t launch_normalize(const float* y, float* x, uint32_t d, uint64_t dpad, const DevState* st,
                            const float* xm, const float* z, double* BP, hipStream_t s) {
  k_normalize<<<(unsigned)normalize_blocks(dpad), 256, 0, s>>>(y, x, d, dpad, st, xm, z, BP);
  return hipGetLastError();
}

hipError_t launch_sum_rows(const float* P2, uint32_t S, uint64_t dpad, uint32_t d, float* z, float* v,
                           const DevState* st, hipStream_t s) {
  k_sum_rows<<<grid1(dpad, 256), 1024, 0, s>>>(P2, S, dpad, d, z, v, st);
  return hipGetLastError();
}

hipError_t launch_dots(const float* x, const float* xm, const float* z, uint64_t dpad, double* BP,
                       hipStream_t s) {
  k_dots<<<(unsigned)dots_blocks(dpad), 256, 0, s>>>(x, xm, z, dpad, BP);
  return hipGetLastError();
}

uint64_t reduce_p1_blocks(uint64_t n) { return grid1(n, 256); }

hipError_t launch_reduce_p1(const float* P1, uint32_t R, uint64_t n, float* t, double* BP,
                            hipStream_t s) {
  k_reduce_p1<<<(unsigned)reduce_p1_blocks(n), 256, 0, s>>>(P1, R, n, t, BP);
  return hipGetLastError();
}

hipError_t launch_unit(const float* t, uint64_t n, const DevState* st, float* v, hipStream_t s) {
  k_unit<<<(unsigned)grid1(n, 256), 256, 0, s>>>(t, n, st, v);
  return hipGetLastError();
}

hipError_t launch_rank1(int dtype, void* out, uint32_t d, uint64_t n, uint64_t ld, float g,
                        const float* a, const float* b, hipStream_t s) {
  return launch_rank1_batched(dtype, out, d, n, ld, 1, 0, g, nullptr, a, 0, b, 0, s);
}

hipError_t launch_rank1_batched(int dtype, void* out, uint32_t d, uint64_t n, uint64_t ld, uint64_t count,
                                uint64_t ostride, float g, const float* gs, const float* a, uint64_t astride,
                                const float* b, uint64_t bstride, hipStream_t s) {
  if (n == 0 || d == 0 || count == 0) return hipSuccess;
  const uint64_t V = dtype == 1 ? 8 : 4;
  const uint64_t esz = dtype == 1 ? 2 : 4;
  const int vec = (ld % V == 0) && (((uintptr_t)out & 15u) == 0) && (((uintptr_t)a & 15u) == 0) &&
                  (count == 1 || ((ostride * esz) % 16 == 0 && astride % 4 == 0));
  // about 2048 workgroups over the whole stack, a wave per column
  const uint64_t per = std::max<uint64_t>(1, (2048 + count - 1) / count);
  const dim3 grid((unsigned)std::min<uint64_t>((n + 3) / 4, per), (unsigned)count);
  if (dtype == 1)
    k_rank1<__hip_bfloat16><<<grid, 256, 0, s>>>((__hip_bfloat16*)out, d, ld, n, g, gs, a, b, vec, ostride, astride,
                                                bstride);
  else
    k_rank1<float><<<grid, 256, 0, s>>>((float*)out, d, ld, n, g, gs, a, b, vec, ostride, astride, bstride);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t val, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_fill_u32<<<grid1(n, 256), 256, 0, s>>>(p, n, val);
  return hipGetLastError();
}

}  // namespace amm
