// amm_resident.hpp -- column-resident building blocks shared by the
// single-matrix one-read passes (k_colstats_z, k_gram) and the batched
// one-read exact step (k_bgram): a 256-thread workgroup holds one column in
// registers, wave w owning rows [w q, (w + 1) q) with q = resident_rows(d),
// each lane L 16-byte loads; c a_j is added into a private FP32 partial of
// the rows, stored at the end.
#pragma once

#include "amm_device.hpp"

namespace amm {

template <typename T, int L>
struct ColRegs {
  static constexpr int V = Elem<T>::VEC;
  static constexpr uint32_t step = 64u * V;
  u32x4 raw[L];

  __device__ __forceinline__ void load(const T* col, uint32_t rb, uint32_t re, int lane) {
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const uint32_t row = rb + lane * V + l * step;
      if (row + V <= re) {
        raw[l] = load16<true>(col + row);
      } else {
        float v[V];
        load_col<T>(col, row, re, v);  // zeros at and beyond re
        if constexpr (V == 4) {
          raw[l] = u32x4{__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2]),
                         __float_as_uint(v[3])};
        } else {
          uint32_t h[4];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            h[k] = (__float_as_uint(v[2 * k]) >> 16) | (__float_as_uint(v[2 * k + 1]) & 0xFFFF0000u);
          raw[l] = u32x4{h[0], h[1], h[2], h[3]};
        }
      }
    }
  }

  static __device__ __forceinline__ void unpack(const u32x4& r, float (&v)[V]) {
    if constexpr (V == 4) {
      v[0] = __uint_as_float(r.x);
      v[1] = __uint_as_float(r.y);
      v[2] = __uint_as_float(r.z);
      v[3] = __uint_as_float(r.w);
    } else {
      v[0] = bf16_bits_to_f32(r.x & 0xFFFFu);
      v[1] = bf16_bits_to_f32(r.x >> 16);
      v[2] = bf16_bits_to_f32(r.y & 0xFFFFu);
      v[3] = bf16_bits_to_f32(r.y >> 16);
      v[4] = bf16_bits_to_f32(r.z & 0xFFFFu);
      v[5] = bf16_bits_to_f32(r.z >> 16);
      v[6] = bf16_bits_to_f32(r.w & 0xFFFFu);
      v[7] = bf16_bits_to_f32(r.w >> 16);
    }
  }
};

// rows per wave of the column-resident passes (a multiple of 64 VEC)
template <int V>
__device__ __forceinline__ uint32_t resident_rows(uint32_t d) {
  constexpr uint32_t step = 64u * V;
  return (((d + 3) / 4) + step - 1) / step * step;
}

// acc += c * column, then (at the end) the partial out to OUT[blockIdx.x]
template <typename T, int L>
__device__ __forceinline__ void resident_axpy(float (&acc)[L * Elem<T>::VEC], float c,
                                              const u32x4 (&col)[L]) {
  constexpr int V = Elem<T>::VEC;
#pragma unroll
  for (int l = 0; l < L; ++l) {
    float v[V];
    ColRegs<T, L>::unpack(col[l], v);
#pragma unroll
    for (int k = 0; k < V; ++k) acc[l * V + k] = fmaf(c, v[k], acc[l * V + k]);
  }
}

template <typename T, int L>
__device__ __forceinline__ void resident_store(const float (&acc)[L * Elem<T>::VEC], float* out,
                                               uint32_t rb, uint32_t q, int lane, uint64_t dpad) {
  constexpr int V = Elem<T>::VEC;
  // Each wave writes its own rows [rb, rb + q) only (L x 64 x VEC may exceed
  // q when L was rounded up to a power of two); together they cover
  // [0, dpad): 4 q is a multiple of 1024 >= d and dpad <= 4 q.
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const uint32_t row = rb + lane * V + l * ColRegs<T, L>::step;
    if (row < rb + q && row < dpad) {
      float v[V];
#pragma unroll
      for (int k = 0; k < V; ++k) v[k] = acc[l * V + k];
      store_vecf<V>(out + row, v);
    }
  }
}

}  // namespace amm
