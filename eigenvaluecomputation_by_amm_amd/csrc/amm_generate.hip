// amm_generate.hip -- synthetic symmetric test matrices of the AMM command
// line (device_memory.h:402-431): gaussian / lognormal / uniform entries,
// symmetrised, then shifted by the Gershgorin offset
//   delta = max_j max(0, sum_i |a_ij| - 2 |a_jj|)   (find_offset, :89-105)
// added to the diagonal (add_offset, :107-114) so that every eigenvalue is
// non-negative and sqrt(lambda_max(A^2)) - delta = lambda_max(A).
//
// Entry (i, j) and (j, i) come from the same Philox counter
// (min(i,j), max(i,j), 0, kStreamMatrix), so the matrix is symmetric by
// construction (the reference generates, then copies one triangle:
// convert_to_symmetry, :61-76).  The offset is the true maximum (the
// reference reads the element after the cublasIsamax result, which is
// 1-based: device_memory.h:427-429).
#include <hip/hip_runtime.h>

#include "amm.h"
#include "amm_device.hpp"

namespace amm {

constexpr uint32_t kStreamMatrix = 3u;

__global__ __launch_bounds__(256) void k_gen_sym(float* __restrict__ A, uint64_t m, uint64_t ld,
                                                 int dist, float p1, float p2, uint64_t seed) {
  const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= m * m) return;
  const uint64_t j = e / m, i = e - j * m;  // row i of column j
  const uint32_t lo = (uint32_t)(i < j ? i : j), hi = (uint32_t)(i < j ? j : i);
  uint32_t r[4];
  philox4x32_10(lo, hi, 0u, kStreamMatrix, (uint32_t)seed, (uint32_t)(seed >> 32), r);
  const double u1 = uniform53(r[0], r[1]);
  float v;
  if (dist == AMM_DIST_UNIFORM) {
    v = p1 + (p2 - p1) * (float)u1;
  } else {
    const double u2 = uniform53(r[2], r[3]);
    const double g = sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
    const double z = (double)p1 + (double)p2 * g;
    v = dist == AMM_DIST_LOGNORMAL ? (float)exp(z) : (float)z;
  }
  A[j * ld + i] = v;
}

// one wave per column: s_j = sum_i |a_ij| - 2 |a_jj|; max over columns by an
// atomic max on the bit pattern (valid for non-negative floats)
__global__ __launch_bounds__(256) void k_gersh(const float* __restrict__ A, uint64_t m, uint64_t ld,
                                               uint32_t* __restrict__ delta_bits) {
  const int lane = threadIdx.x & 63;
  const uint64_t j = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= m) return;
  const float* col = A + j * ld;
  double s = 0.0;
  for (uint64_t i = lane; i < m; i += 64) s += fabs((double)col[i]);
  s = wave_sum(s);
  if (lane == 0) {
    const float off = (float)(s - 2.0 * fabs((double)col[j]));
    const float pos = off > 0.0f ? off : 0.0f;
    atomicMax(delta_bits, __float_as_uint(pos));
  }
}

__global__ void k_fill_u32_one(uint32_t* p) {
  if (threadIdx.x == 0) *p = 0u;
}

__global__ __launch_bounds__(256) void k_add_diag(float* __restrict__ A, uint64_t m, uint64_t ld,
                                                  const uint32_t* __restrict__ delta_bits) {
  const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (j < m) A[j * ld + j] += __uint_as_float(*delta_bits);
}

__global__ void k_f32_to_bf16(const float* __restrict__ src, uint16_t* __restrict__ dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t u = __float_as_uint(src[i]);
  // round to nearest even; NaN stays NaN
  const uint32_t r = ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu))
                         ? ((u >> 16) | 0x40u)
                         : ((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
  dst[i] = (uint16_t)r;
}

hipError_t launch_generate_symmetric(float* A, uint64_t m, uint64_t ld, int dist, float p1, float p2,
                                    uint64_t seed, uint32_t* delta_bits, hipStream_t s) {
  const uint64_t total = m * m;
  k_fill_u32_one<<<1, 64, 0, s>>>(delta_bits);
  k_gen_sym<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(A, m, ld, dist, p1, p2, seed);
  k_gersh<<<(unsigned)((m + 3) / 4), 256, 0, s>>>(A, m, ld, delta_bits);
  k_add_diag<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(A, m, ld, delta_bits);
  return hipGetLastError();
}

hipError_t launch_f32_to_bf16(const float* src, uint16_t* dst, uint64_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  k_f32_to_bf16<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(src, dst, count);
  return hipGetLastError();
}

}  // namespace amm
