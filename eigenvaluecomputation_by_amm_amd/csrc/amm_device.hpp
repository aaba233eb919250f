// amm_device.hpp -- device helpers shared by the AMM power-method kernels
// (gfx950, wave64).  Philox4x32-10 follows Salmon et al., SC'11; the
// reductions are fixed butterflies so every result is reproducible run to
// run.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace amm {

constexpr int kWave = 64;

// Philox counter word 3 names the stream.
constexpr uint32_t kStreamSample = 1u;
constexpr uint32_t kStreamX0 = 2u;

// Samples are deduplicated in segments of this many slots (one workgroup
// sorts one segment in LDS: 8192 x 4 B = 32 KiB).
constexpr uint32_t kSampleSeg = 8192u;

// Per-iteration step kinds (device mode array).
enum StepMode : uint32_t {
  kModeExactMemo = 1,    // exact A(A^T x), snapshot [x, y] -> [x_memo, y_memo]
  kModeAmmPrevious = 2,  // sampled A_S W A_S^T (x - x_memo) + y_memo
  kModeAmmProposed = 3,  // mean-shifted sampled product + rank-2 part
  kModeExactNoMemo = 4,  // exact product, memo untouched
};

// Adaptive variance reduction: is iteration i an exact step?  Mirrors
// powerMethod_by_previous.cuh:135-138 (kernel 1) and
// powerMethod_by_proposed.cuh:159-161 (kernel 2, at most every other step).
__device__ __forceinline__ bool exact_step_due(uint32_t kernel, uint64_t i, uint64_t period,
                                               double thr, double prev_cos, uint64_t last_exact) {
  const bool vr = period > 0 || thr > 0.0;
  const bool due = (period > 0 && (i + 1) % period == 0) || (thr > 0.0 && prev_cos > 1.0 - thr);
  if (kernel == 1) return vr && due;
  if (kernel == 2) return (i > last_exact + 1) && vr && due;
  return false;
}

// Status bits of DevState::status.
constexpr uint32_t kStatusNonFinite = 1u;  // a column weight was NaN or Inf

// blocks of a multi-block tail (k_tail), each writing one part of s and q
constexpr uint32_t kTailMaxBlocks = 64;

struct alignas(16) DevState {
  uint64_t total[2];    // integer CDF totals: [0] plain ||a_j||^2, [1] centred
  uint64_t last_pos[2]; // last positive-weight column per CDF
  double alpha2;        // ||alpha||^2
  double scale;         // 1 / ||y|| of the latest product
  double cos_prev;      // y^T x / ||y|| of the latest iteration
  double sigma;         // ||A^T u|| of the last sigma_max evaluation
  uint64_t last_exact;  // last exact iteration (proposed gate)
  uint64_t seed;        // Philox key of the run
  uint32_t trial;       // Philox counter word 2 of the run
  uint32_t status;      // kStatus* bits, sticky until the host clears them
  uint32_t pad0;
  int64_t t_start;      // wall clock at the run's first kernel
  // proposed: s = 1^T dx and q = z^T dx for the next product, as sq_n
  // per-block parts (summed by sq_total); no parts means s = q = 0
  uint32_t sq_n;
  uint32_t pad1;
  double sq_part[kTailMaxBlocks][2];
};

// s and q from their per-block parts, summed in a fixed order (lane i takes
// part i; the butterfly gives every lane the same bits).  Call with the
// whole wave active.
__device__ __forceinline__ double2 sq_total(const DevState* st) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n = st->sq_n;
  double a = 0.0, b = 0.0;
  if (lane < n) {
    a = st->sq_part[lane][0];
    b = st->sq_part[lane][1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  return make_double2(a, b);
}

__device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
  return __umulhi(a, b);
}

__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = mulhi32(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = mulhi32(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// t = floor(u * total / 2^64) for the 64-bit uniform u = (a << 32) | b.
__device__ __forceinline__ uint64_t draw_target(uint32_t a, uint32_t b, uint64_t total) {
  return __umul64hi(((uint64_t)a << 32) | (uint64_t)b, total);
}

__device__ __forceinline__ int ceil_log2_u64(uint64_t n) {
  int L = 0;
  while (L < 63 && (1ull << L) < n) ++L;
  return L;
}

// 53-bit uniform in [0,1): ((a>>5)*2^26 + (b>>6)) * 2^-53, exact.
__device__ __forceinline__ double uniform53(uint32_t a, uint32_t b) {
  const uint64_t v = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)v * (1.0 / 9007199254740992.0);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum over a 256-thread block (4 waves); result valid in thread 0.
template <typename T>
__device__ __forceinline__ T block256_sum(T v, T* lds4) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds4[wave] = v;
  __syncthreads();
  T r = v;
  if (threadIdx.x == 0) r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
  __syncthreads();
  return r;
}

// FP32 -> BF16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t bf16_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf16_bits_to_f32(uint32_t h16) {
  return __uint_as_float(h16 << 16);
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 16-byte load; NT marks it non-temporal (streamed data read once: measured
// ~7.0 TB/s on MI355X for a 4 GiB read against ~6.2 TB/s for plain loads).
template <bool NT>
__device__ __forceinline__ u32x4 load16(const void* p) {
  if (NT) return __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
  return *reinterpret_cast<const u32x4*>(p);
}

// Element-type traits: VEC elements per 16-byte load.
template <typename T>
struct Elem;

template <>
struct Elem<float> {
  static constexpr int VEC = 4;
  template <bool NT = false>
  __device__ __forceinline__ static void load_full(const float* p, float (&v)[4]) {
    const u32x4 q = load16<NT>(p);
    v[0] = __uint_as_float(q.x);
    v[1] = __uint_as_float(q.y);
    v[2] = __uint_as_float(q.z);
    v[3] = __uint_as_float(q.w);
  }
  __device__ __forceinline__ static float load_one(const float* p) { return *p; }
};

template <>
struct Elem<__hip_bfloat16> {
  static constexpr int VEC = 8;
  template <bool NT = false>
  __device__ __forceinline__ static void load_full(const __hip_bfloat16* p, float (&v)[8]) {
    const u32x4 q = load16<NT>(p);
    v[0] = bf16_bits_to_f32(q.x & 0xFFFFu);
    v[1] = bf16_bits_to_f32(q.x >> 16);
    v[2] = bf16_bits_to_f32(q.y & 0xFFFFu);
    v[3] = bf16_bits_to_f32(q.y >> 16);
    v[4] = bf16_bits_to_f32(q.z & 0xFFFFu);
    v[5] = bf16_bits_to_f32(q.z >> 16);
    v[6] = bf16_bits_to_f32(q.w & 0xFFFFu);
    v[7] = bf16_bits_to_f32(q.w >> 16);
  }
  __device__ __forceinline__ static float load_one(const __hip_bfloat16* p) {
    return bf16_bits_to_f32(*reinterpret_cast<const uint16_t*>(p));
  }
};

// Load VEC consecutive elements starting at row `row` of a column; rows at
// or beyond d read as zero.
template <typename T>
__device__ __forceinline__ void load_col(const T* col, uint32_t row, uint32_t d,
                                         float (&v)[Elem<T>::VEC]) {
  constexpr int V = Elem<T>::VEC;
  if (row + V <= d) {
    Elem<T>::load_full(col + row, v);
  } else {
#pragma unroll
    for (int k = 0; k < V; ++k) v[k] = (row + k < d) ? Elem<T>::load_one(col + row + k) : 0.0f;
  }
}

// Load VEC floats of a padded FP32 vector (length a multiple of 8).
template <int V>
__device__ __forceinline__ void load_vecf(const float* p, float (&v)[V]) {
#pragma unroll
  for (int k = 0; k < V; k += 4) {
    const float4 q = *reinterpret_cast<const float4*>(p + k);
    v[k] = q.x;
    v[k + 1] = q.y;
    v[k + 2] = q.z;
    v[k + 3] = q.w;
  }
}

template <int V>
__device__ __forceinline__ void store_vecf(float* p, const float (&v)[V]) {
#pragma unroll
  for (int k = 0; k < V; k += 4)
    *reinterpret_cast<float4*>(p + k) = make_float4(v[k], v[k + 1], v[k + 2], v[k + 3]);
}

// Coarse table: K bucket ends e_m = floor((m + 1) n / K) - 1 and their CDF
// values in LDS; a target is first located among the buckets in LDS, then
// within its bucket (about n / K entries) in global memory.  Same result as
// the plain search for any K: the smallest j with cdf[j] > t.
constexpr uint32_t kCoarse = 1024;        // the batched sampler: 8 KB of LDS beside its 64 KB count row
constexpr uint32_t kCoarseSample = 8192;  // the single-vector sampler: 64 KB of LDS, buckets of n / 8192

template <uint32_t K = kCoarse>
__device__ __forceinline__ uint32_t bucket_end(uint32_t m, uint32_t n) {
  return (uint32_t)(((uint64_t)(m + 1) * n) / K) - 1u;  // may wrap for empty leading buckets
}

template <uint32_t K = kCoarse>
__device__ __forceinline__ uint32_t bcdf_search(const uint64_t* __restrict__ cdf, const uint64_t* cs,
                                                uint32_t n, uint64_t t, uint64_t last_pos) {
  // smallest bucket m with cs[m] > t (cs[m] = cdf at its end; empty buckets
  // repeat the previous end's value)
  uint32_t lo = 0, hi = K;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cs[mid] > t) hi = mid; else lo = mid + 1;
  }
  if (lo >= K) return (uint32_t)last_pos;
  uint32_t jlo = lo ? bucket_end<K>(lo - 1, n) + 1u : 0u;
  uint32_t jhi = bucket_end<K>(lo, n) + 1u;
  while (jlo < jhi) {
    const uint32_t mid = jlo + ((jhi - jlo) >> 1);
    if (cdf[mid] > t) jhi = mid; else jlo = mid + 1;
  }
  return jlo < n ? jlo : (uint32_t)last_pos;
}

// Fill the coarse table cs[K] (LDS) by threads tid = 0, stride, ...; an
// empty bucket before the first column ends at "-1": CDF value 0.  Loads in
// batches of eight (a loop with a load per trip keeps one in flight).
template <uint32_t K = kCoarse>
__device__ __forceinline__ void coarse_fill(uint64_t* cs, const uint64_t* __restrict__ cdf, uint32_t n,
                                            uint32_t tid, uint32_t stride) {
  uint32_t m = tid;
  for (; m + 7 * stride < K; m += 8 * stride) {
    uint64_t v[8];
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) {
      const uint32_t mu = m + u * stride;
      const bool empty = ((uint64_t)(mu + 1) * n) / K == 0;
      v[u] = cdf[empty ? 0u : bucket_end<K>(mu, n)];
      v[u] = empty ? 0ull : v[u];
    }
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) cs[m + u * stride] = v[u];
  }
  for (; m < K; m += stride) cs[m] = ((uint64_t)(m + 1) * n) / K == 0 ? 0ull : cdf[bucket_end<K>(m, n)];
}

}  // namespace amm
