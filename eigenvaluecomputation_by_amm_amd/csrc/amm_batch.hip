// amm_batch.hip -- kernels of the batched AMM power method (gfx950, wave64).
//
// Trial k of problem p is, sample for sample, the single-vector run with
// Philox key seed_p and counter word 2 = trial0 + k (k_sample, k_x0 in
// amm_kernels.hip): the same columns, multiplicities and weights; only the
// summation order of the contractions differs.
#include "amm_batch.hpp"

namespace amm {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// largest p with off[p] <= b (off is non-decreasing, off[0] = 0, b < off[np])
__device__ __forceinline__ uint32_t find_problem(const uint32_t* __restrict__ off, uint32_t np,
                                                 uint32_t b) {
  uint32_t lo = 0, hi = np;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// sum over the 16 lanes of a 16-lane group (groups aligned to 16 lanes)
template <typename T>
__device__ __forceinline__ T sum16(T v) {
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 16);
  return v;
}

__device__ __forceinline__ uint16_t bf16_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ bool is_amm(uint32_t md) {
  return md == kModeAmmPrevious || md == kModeAmmProposed;
}

// ---------------------------------------------------------------------------
// Run prologue: x_memo = y_memo = 0 (zero_clear, device_memory.h:345-355),
// per-trial state, start stamp.  Items: nvblk row blocks per problem.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bbegin(const BProb* __restrict__ P, const uint32_t* off,
                                                BRun r) {
  const uint32_t p = find_problem(off, r.np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t blk = blockIdx.x - off[p];
  const uint32_t k = threadIdx.x >> 4;
  const uint64_t row = (uint64_t)blk * kBVecRows + (threadIdx.x & 15) * 4;
  const uint64_t o = (uint64_t)k * pr.dstage + row;
  *reinterpret_cast<float4*>(pr.XM + o) = make_float4(0.f, 0.f, 0.f, 0.f);
  *reinterpret_cast<float4*>(pr.YM + o) = make_float4(0.f, 0.f, 0.f, 0.f);
  if (blk == 0 && threadIdx.x < kBT) {
    BTrial t{};
    t.mode = r.modes_static[0];
    pr.bt[threadIdx.x] = t;
    if (threadIdx.x == 0) pr.st->status = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *r.t_start = wall_clock64();
}

// ---------------------------------------------------------------------------
// x0 ~ N(0, I) per trial: Box-Muller on Philox counter (m, 0, trial, x0
// stream) for rows 2m, 2m+1 (as k_x0).  Raw vector into Y; per-block sums of
// squares into BP[., ., 1].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bx0(const BProb* __restrict__ P, const uint32_t* off,
                                             BRun r) {
  const uint32_t p = find_problem(off, r.np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t blk = blockIdx.x - off[p];
  const uint32_t k = threadIdx.x >> 4;
  const uint32_t row = blk * kBVecRows + (threadIdx.x & 15) * 4;
  float g[4] = {0.f, 0.f, 0.f, 0.f};
  if (k < r.ntr) {
    const uint32_t k0 = (uint32_t)pr.seed, k1 = (uint32_t)(pr.seed >> 32);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t r0 = row + 2 * h;
      if (r0 < pr.d) {
        uint32_t q[4];
        philox4x32_10(r0 >> 1, 0u, r.trial0 + k, kStreamX0, k0, k1, q);
        const double u1 = uniform53(q[0], q[1]);
        const double u2 = uniform53(q[2], q[3]);
        const double rad = sqrt(-2.0 * log(1.0 - u1));
        const double th = 6.283185307179586 * u2;
        g[2 * h] = (float)(rad * cos(th));
        if (r0 + 1 < pr.d) g[2 * h + 1] = (float)(rad * sin(th));
      }
    }
  }
  *reinterpret_cast<float4*>(pr.Y + (uint64_t)k * pr.dstage + row) = make_float4(g[0], g[1], g[2], g[3]);
  double yy = 0.0;
#pragma unroll
  for (int e = 0; e < 4; ++e) yy += (double)g[e] * (double)g[e];
  yy = sum16(yy);
  if ((threadIdx.x & 15) == 0) {
    double* b = pr.BP + ((uint64_t)blk * kBT + k) * 3;
    b[0] = 0.0;
    b[1] = yy;
    b[2] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Per-problem tail (dotProduct_and_generate_x, common.h:89-154, for 16
// trials): thread (k, sub) sums blocks sub, sub + 16, ... of trial k in a
// fixed order.  kBTailFinish records lambda_i = y^T x, y^T v* / ||y||,
// 1/||y||, cos, and decides the next step kind; kBTailScale forms 1/||y||
// of the raw x0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_btail(const BProb* __restrict__ P, BRun r, uint32_t op,
                                               uint32_t it) {
  const BProb& pr = P[blockIdx.x];
  const uint32_t k = threadIdx.x >> 4, sub = threadIdx.x & 15;
  double a = 0.0, b = 0.0, c = 0.0;
  for (uint32_t blk = sub; blk < pr.nvblk; blk += 16) {
    const double* q = pr.BP + ((uint64_t)blk * kBT + k) * 3;
    a += q[0];
    b += q[1];
    c += q[2];
  }
  a = sum16(a);
  b = sum16(b);
  c = sum16(c);
  if (sub != 0 || k >= r.ntr) return;
  BTrial& t = pr.bt[k];
  const double scale = b > 0.0 ? 1.0 / sqrt(b) : 0.0;
  t.scale = scale;
  if (op == kBTailScale) return;
  const uint32_t md = t.mode;
  pr.eig[(uint64_t)k * r.step + it] = a;
  pr.err[(uint64_t)k * r.step + it] = c * scale;
  pr.modes[(uint64_t)k * r.step + it] = md;
  const double cs = a * scale;
  t.cos_prev = cs;
  uint64_t le = t.last_exact;
  if (md == kModeExactMemo || md == kModeExactNoMemo) {
    le = it;
    t.last_exact = le;
  }
  if (it + 1 < r.step) {
    uint32_t nx;
    if (r.adaptive) {
      const bool ex = exact_step_due(r.kernel, it + 1, r.period, r.thr, cs, le);
      nx = ex ? kModeExactMemo : (r.kernel == 2 ? kModeAmmProposed : kModeAmmPrevious);
    } else {
      nx = r.modes_static[it + 1];
    }
    t.mode = nx;
  }
  if (blockIdx.x == 0 && k == 0 && r.ts) r.ts[it] = wall_clock64() - *r.t_start;
}

// ---------------------------------------------------------------------------
// x = y / ||y|| per trial, then stage the operand of the next product:
// dx = x - x_memo (sampled step) or x (exact step) as FP32 or as a BF16
// pair hi + lo (hi = bf16(dx), lo = bf16(dx - hi)); for a proposed step also
// block partials of 1^T dx and z^T dx in FP64 (as k_tail's norm pass).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bnorm(const BProb* __restrict__ P, const uint32_t* off,
                                               BRun r) {
  const uint32_t p = find_problem(off, r.np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t blk = blockIdx.x - off[p];
  const uint32_t k = threadIdx.x >> 4;
  const uint32_t row = blk * kBVecRows + (threadIdx.x & 15) * 4;
  const uint64_t o = (uint64_t)k * pr.dstage + row;
  const BTrial& t = pr.bt[k];
  const double scale = k < r.ntr ? t.scale : 0.0;
  const uint32_t md = k < r.ntr ? t.mode : kModeExactNoMemo;
  const bool memo = is_amm(md);
  const float4 y4 = *reinterpret_cast<const float4*>(pr.Y + o);
  float4 m4 = make_float4(0.f, 0.f, 0.f, 0.f);
  if (memo) m4 = *reinterpret_cast<const float4*>(pr.XM + o);
  const float yv[4] = {y4.x, y4.y, y4.z, y4.w};
  const float mv[4] = {m4.x, m4.y, m4.z, m4.w};
  float xv[4], dv[4];
  double s = 0.0, q = 0.0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    xv[e] = (row + e < pr.d) ? (float)(scale * (double)yv[e]) : 0.0f;
    dv[e] = xv[e] - mv[e];
    if (md == kModeAmmProposed && row + e < pr.d) {
      const double dx = (double)xv[e] - (double)mv[e];
      s += dx;
      q += (double)pr.z[row + e] * dx;
    }
  }
  *reinterpret_cast<float4*>(pr.X + o) = make_float4(xv[0], xv[1], xv[2], xv[3]);
  if (pr.dtype == 1) {
    uint16_t h[4], l[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      h[e] = bf16_rne(dv[e]);
      l[e] = bf16_rne(dv[e] - bf16_bits_to_f32(h[e]));
    }
    *reinterpret_cast<uint2*>(pr.DXh + o) =
        make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
    *reinterpret_cast<uint2*>(pr.DXl + o) =
        make_uint2((uint32_t)l[0] | ((uint32_t)l[1] << 16), (uint32_t)l[2] | ((uint32_t)l[3] << 16));
  } else {
    *reinterpret_cast<float4*>(pr.DX + o) = make_float4(dv[0], dv[1], dv[2], dv[3]);
  }
  if (md == kModeAmmProposed) {
    s = sum16(s);
    q = sum16(q);
    if ((threadIdx.x & 15) == 0) {
      double* b = pr.BQ + ((uint64_t)blk * kBT + k) * 2;
      b[0] = s;
      b[1] = q;
    }
  }
}

// ---------------------------------------------------------------------------
// Draws of iteration `it` (get_rand + pick_index, device_memory.h:506-545):
// slot i of trial k takes the 64-bit uniform of Philox counter (i >> 1, it,
// trial0 + k, sample stream), words 2(i & 1), 2(i & 1) + 1, the integer
// target floor(u * total / 2^64) and column min{j : cdf_j > target}, then
// counts the draw in cnt[j][k].  Items: trials x ceil(c / 2048) per problem.
// The first block of a proposed trial also finishes s, q of its product.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t bcdf_search(const uint64_t* __restrict__ cdf, uint32_t n,
                                                uint64_t t, uint64_t last_pos) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (cdf[mid] > t) hi = mid; else lo = mid + 1;
  }
  return lo < n ? lo : (uint32_t)last_pos;
}

__global__ __launch_bounds__(256) void k_bsample(const BProb* __restrict__ P, const uint32_t* off,
                                                 BRun r, uint32_t it) {
  __shared__ double red[4];
  const uint32_t p = find_problem(off, r.np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t local = blockIdx.x - off[p];
  const uint32_t per = (pr.c + 2047) / 2048;
  const uint32_t k = local / per, part = local % per;
  BTrial& t = pr.bt[k];
  const uint32_t md = t.mode;
  if (!is_amm(md)) return;
  const int which = md == kModeAmmProposed ? 1 : 0;
  if (which && part == 0) {
    double s = 0.0, q = 0.0;
    for (uint32_t blk = threadIdx.x; blk < pr.nvblk; blk += 256) {
      const double* b = pr.BQ + ((uint64_t)blk * kBT + k) * 2;
      s += b[0];
      q += b[1];
    }
    s = block256_sum(s, red);
    q = block256_sum(q, red);
    if (threadIdx.x == 0) {
      t.s = s;
      t.q = q;
    }
  }
  const uint64_t total = pr.st->total[which];
  const uint64_t last_pos = pr.st->last_pos[which];
  const uint64_t* cdf = which ? pr.cdf1 : pr.cdf0;
  const uint32_t k0 = (uint32_t)pr.seed, k1 = (uint32_t)(pr.seed >> 32);
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const uint32_t pair = part * 1024 + h * 256 + threadIdx.x;
    const uint32_t slot = 2 * pair;
    if (slot < pr.c) {
      uint32_t q[4];
      philox4x32_10(pair, it, r.trial0 + k, kStreamSample, k0, k1, q);
      const uint32_t j0 = bcdf_search(cdf, pr.n, draw_target(q[0], q[1], total), last_pos);
      atomicAdd(pr.cnt + (uint64_t)j0 * kBT + k, 1u);
      if (slot + 1 < pr.c) {
        const uint32_t j1 = bcdf_search(cdf, pr.n, draw_target(q[2], q[3], total), last_pos);
        atomicAdd(pr.cnt + (uint64_t)j1 * kBT + k, 1u);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Union of the trials' samples, one wave per tile of 64 columns: column j is
// present when a sampled trial drew it or a trial takes an exact step.  The
// wave compacts the present columns of its tile (ballot + prefix count) into
// list[tile*64 ...], writes their weights W[slot][k] -- mult * total /
// (c * q_j) for a sampled trial (update_weights, common.h:42-53, as
// k_sample), 1 for an exact trial, 0 otherwise -- and clears the counts.
// Items: ceil(ntiles / 4) per problem.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bprep(const BProb* __restrict__ P, const uint32_t* off,
                                               BRun r) {
  const uint32_t p = find_problem(off, r.np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t tile = (blockIdx.x - off[p]) * 4 + (threadIdx.x >> 6);
  if (tile >= pr.ntiles) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t j = tile * kBTile + lane;
  const bool inb = j < pr.n;
  uint32_t cnt[kBT];
  uint4* cp = reinterpret_cast<uint4*>(pr.cnt + (uint64_t)j * kBT);
  if (inb) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint4 c4 = cp[v];
      cnt[4 * v] = c4.x;
      cnt[4 * v + 1] = c4.y;
      cnt[4 * v + 2] = c4.z;
      cnt[4 * v + 3] = c4.w;
    }
  } else {
#pragma unroll
    for (int v = 0; v < (int)kBT; ++v) cnt[v] = 0;
  }
  float w[kBT];
  bool present = false;
  uint64_t q0 = 0, q1 = 0;
  bool have_q = false;
#pragma unroll
  for (int k = 0; k < (int)kBT; ++k) {
    w[k] = 0.0f;
    if ((uint32_t)k >= r.ntr || !inb) continue;
    const uint32_t md = pr.bt[k].mode;
    if (md == kModeExactMemo || md == kModeExactNoMemo) {
      w[k] = 1.0f;
      present = true;
    } else if (cnt[k]) {
      if (!have_q) {
        q0 = pr.cdf0 ? pr.cdf0[j] - (j ? pr.cdf0[j - 1] : 0ull) : 0ull;
        q1 = pr.cdf1 ? pr.cdf1[j] - (j ? pr.cdf1[j - 1] : 0ull) : 0ull;
        have_q = true;
      }
      const int which = md == kModeAmmProposed ? 1 : 0;
      const uint64_t qj = which ? q1 : q0;
      const double num = (double)cnt[k] * (double)pr.st->total[which];
      const double den = (double)pr.c * (double)qj;
      w[k] = den > 0.0 ? (float)(num / den) : 0.0f;
      present = true;
    }
  }
  const uint64_t mask = __ballot(present);
  const uint32_t pos = __popcll(mask & ((1ull << lane) - 1ull));
  if (present) {
    const uint64_t slot = (uint64_t)tile * kBTile + pos;
    pr.list[slot] = j;
    float4* wp = reinterpret_cast<float4*>(pr.W + slot * kBT);
#pragma unroll
    for (int v = 0; v < 4; ++v) wp[v] = make_float4(w[4 * v], w[4 * v + 1], w[4 * v + 2], w[4 * v + 3]);
  }
  if (lane == 0) pr.tcount[tile] = (uint32_t)__popcll(mask);
  if (inb) {
#pragma unroll
    for (int v = 0; v < 4; ++v) cp[v] = make_uint4(0u, 0u, 0u, 0u);
  }
}

// ---------------------------------------------------------------------------
// 16 bytes of column `col` from row `row` (rows at or beyond d read 0).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ u32x4 load_a16(const T* col, uint32_t row, uint32_t d) {
  constexpr uint32_t V = Elem<T>::VEC;
  if (row + V <= d) return *reinterpret_cast<const u32x4*>(col + row);
  u32x4 z = {0u, 0u, 0u, 0u};
  if (row >= d) return z;
  if constexpr (V == 4) {
    const uint32_t* c32 = reinterpret_cast<const uint32_t*>(col + row);
    for (uint32_t e = 0; e < 4 && row + e < d; ++e) z[e] = c32[e];
  } else {
    const uint16_t* c16 = reinterpret_cast<const uint16_t*>(col + row);
    for (uint32_t e = 0; e < 8 && row + e < d; ++e) z[e >> 1] |= (uint32_t)c16[e] << (16 * (e & 1));
  }
  return z;
}

// ---------------------------------------------------------------------------
// Contraction 1 (previousAMM_kernel1 / proposedAMM_kernel1 for 16 trials):
// P1[ks][slot][k] = a_{list[slot]}[rows of slice ks]^T dx_k.  Workgroup =
// (tile, row slice); wave w takes a quarter of the slice's rows and all
// present columns of the tile in groups of 16 (one MFMA tile 16 columns x
// 16 trials).  BF16: v_mfma_f32_16x16x32_bf16 with dx = hi + lo (two MFMAs
// per 32 rows); FP32: v_mfma_f32_16x16x4_f32, four per 16-byte load.
// A-operand lane (i, g) holds column slot 16G + i, rows 8g.. (BF16) or
// 4g.. (FP32) of the step; the B operand holds the same rows of dx for trial
// i; accumulator lane (i, g) element e is slot 16G + 4g + e, trial i.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_bdot(const BProb* __restrict__ P, const uint32_t* off,
                                              uint32_t np) {
  __shared__ float red[4][kBTile][kBT + 1];
  const uint32_t p = find_problem(off, np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t local = blockIdx.x - off[p];
  const uint32_t tile = local / pr.kslices, ks = local % pr.kslices;
  const uint32_t cnt = pr.tcount[tile];
  if (cnt == 0) return;
  const uint32_t ngroups = (cnt + 15) / 16;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t li = lane & 15, lg = lane >> 4;
  const uint32_t r0 = ks * pr.kslice_rows;
  const uint32_t r1 = min(r0 + pr.kslice_rows, pr.dstage);
  const uint32_t Q = pr.kslice_rows / 4;
  const uint32_t w0 = r0 + wave * Q;
  const uint32_t w1 = min(w0 + Q, r1);
  const T* A = reinterpret_cast<const T*>(pr.A);
  const T* col[4];
  bool val[4];
#pragma unroll
  for (int G = 0; G < 4; ++G) {
    const uint32_t slot = G * 16 + li;
    val[G] = slot < cnt;
    col[G] = A + (uint64_t)(val[G] ? pr.list[(uint64_t)tile * kBTile + slot] : 0u) * pr.ld;
  }
  f32x4_t acc[4];
#pragma unroll
  for (int G = 0; G < 4; ++G) acc[G] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const uint32_t d = pr.d;
  if constexpr (Elem<T>::VEC == 8) {
    const uint16_t* xh = pr.DXh + (uint64_t)li * pr.dstage;
    const uint16_t* xl = pr.DXl + (uint64_t)li * pr.dstage;
    // two 32-row steps per trip: 64 rows (lengths are multiples of 64)
    for (uint32_t kb = w0; kb < w1; kb += 64) {
      const uint32_t ra = kb + 8 * lg, rb = kb + 32 + 8 * lg;
      const u32x4 bh0 = *reinterpret_cast<const u32x4*>(xh + ra);
      const u32x4 bl0 = *reinterpret_cast<const u32x4*>(xl + ra);
      const u32x4 bh1 = *reinterpret_cast<const u32x4*>(xh + rb);
      const u32x4 bl1 = *reinterpret_cast<const u32x4*>(xl + rb);
      u32x4 a0[4], a1[4];
#pragma unroll
      for (int G = 0; G < 4; ++G) {
        if ((uint32_t)G < ngroups) {
          a0[G] = val[G] ? load_a16<T>(col[G], ra, d) : u32x4{0u, 0u, 0u, 0u};
          a1[G] = val[G] ? load_a16<T>(col[G], rb, d) : u32x4{0u, 0u, 0u, 0u};
        }
      }
#pragma unroll
      for (int G = 0; G < 4; ++G) {
        if ((uint32_t)G < ngroups) {
          acc[G] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a0[G]),
                                                           __builtin_bit_cast(bf16x8_t, bh0), acc[G], 0, 0, 0);
          acc[G] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a0[G]),
                                                           __builtin_bit_cast(bf16x8_t, bl0), acc[G], 0, 0, 0);
          acc[G] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a1[G]),
                                                           __builtin_bit_cast(bf16x8_t, bh1), acc[G], 0, 0, 0);
          acc[G] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a1[G]),
                                                           __builtin_bit_cast(bf16x8_t, bl1), acc[G], 0, 0, 0);
        }
      }
    }
  } else {
    const float* xs = pr.DX + (uint64_t)li * pr.dstage;
    // four 16-row steps per trip: 64 rows
    for (uint32_t kb = w0; kb < w1; kb += 64) {
      u32x4 b[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) b[s] = *reinterpret_cast<const u32x4*>(xs + kb + 16 * s + 4 * lg);
      u32x4 a[4][4];
#pragma unroll
      for (int G = 0; G < 4; ++G) {
        if ((uint32_t)G < ngroups) {
#pragma unroll
          for (int s = 0; s < 4; ++s)
            a[G][s] = val[G] ? load_a16<T>(col[G], kb + 16 * s + 4 * lg, d) : u32x4{0u, 0u, 0u, 0u};
        }
      }
#pragma unroll
      for (int G = 0; G < 4; ++G) {
        if ((uint32_t)G < ngroups) {
#pragma unroll
          for (int s = 0; s < 4; ++s) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
              acc[G] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a[G][s][u]),
                                                            __uint_as_float(b[s][u]), acc[G], 0, 0, 0);
          }
        }
      }
    }
  }
  // cross-wave sum in LDS, then one coalesced write of the tile's partials
#pragma unroll
  for (int G = 0; G < 4; ++G) {
#pragma unroll
    for (int e = 0; e < 4; ++e) red[wave][G * 16 + 4 * lg + e][li] = acc[G][e];
  }
  __syncthreads();
  float* out = pr.P1 + ((uint64_t)ks * pr.ntiles * kBTile + (uint64_t)tile * kBTile) * kBT;
  for (uint32_t i = threadIdx.x; i < cnt * kBT; i += 256) {
    const uint32_t slot = i / kBT, k = i % kBT;
    out[i] = (red[0][slot][k] + red[1][slot][k]) + (red[2][slot][k] + red[3][slot][k]);
  }
}

// ---------------------------------------------------------------------------
// Contraction 2 (previousAMM_kernel2 / proposedAMM_kernel2 for 16 trials):
// P2[ch][k][rows] = sum_{slots of chunk ch} a_slot[rows] t_slot,k with
//   t = W (sum_ks P1[ks] - alpha_j s_k)       (s_k = 0 unless proposed)
// staged per tile in LDS.  Workgroup = (row block, column chunk); wave w
// owns `passes` runs of 64 (FP32) or 128 (BF16) rows; MFMA
// v_mfma_f32_16x16x4_f32 with K = 4 columns: A-operand lane (i, g) holds
// column 4q + g at row base + V i + u (u-th element of its 16-byte load),
// B-operand lane (i, g) holds t of column 4q + g, trial i; accumulator u
// lane (i, g) element e is row base + V (4g + e) + u, trial i.  The proposed
// scalar sum_j alpha_j t_j of the chunk goes to P2s (row block 0).
// ---------------------------------------------------------------------------
constexpr uint32_t kMaxPasses = 2;

template <typename T, uint32_t PASSES>
__global__ __launch_bounds__(256) void k_baxpy(const BProb* __restrict__ P, const uint32_t* off,
                                               uint32_t np, uint32_t proposed_any) {
  constexpr uint32_t V = Elem<T>::VEC;
  constexpr uint32_t PR = 16 * V;  // rows per pass
  __shared__ float tl[kBTile][kBT];
  __shared__ uint32_t cl[kBTile];
  __shared__ double cred[256];
  const uint32_t p = find_problem(off, np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t local = blockIdx.x - off[p];
  const uint32_t rb = local / pr.chunks, ch = local % pr.chunks;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t li = lane & 15, lg = lane >> 4;
  const uint32_t t0 = ch * pr.tiles_per_chunk;
  const uint32_t t1 = min(t0 + pr.tiles_per_chunk, pr.ntiles);
  const uint32_t wrow = rb * 4 * PASSES * PR + wave * PASSES * PR;
  const T* A = reinterpret_cast<const T*>(pr.A);
  const uint32_t d = pr.d;
  // per-trial shift s_k (proposed trials only)
  const uint32_t kk = threadIdx.x & 15;
  double sk = 0.0;
  if (proposed_any && pr.bt[kk].mode == kModeAmmProposed) sk = pr.bt[kk].s;
  double corr = 0.0;
  f32x4_t acc[PASSES][V];
#pragma unroll
  for (uint32_t m = 0; m < PASSES; ++m)
#pragma unroll
    for (uint32_t u = 0; u < V; ++u) acc[m][u] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const uint64_t slab = (uint64_t)pr.ntiles * kBTile * kBT;
  for (uint32_t tile = t0; tile < t1; ++tile) {
    const uint32_t cnt = pr.tcount[tile];
    if (cnt == 0) continue;
    __syncthreads();  // the previous tile's readers are done with tl / cl
    for (uint32_t i = threadIdx.x; i < cnt * kBT; i += 256) {
      const uint32_t slot = i >> 4;
      const uint64_t e = ((uint64_t)tile * kBTile + slot) * kBT + (i & 15);
      double acc1 = 0.0;
      for (uint32_t ks = 0; ks < pr.kslices; ++ks) acc1 += (double)pr.P1[ks * slab + e];
      const uint32_t j = pr.list[(uint64_t)tile * kBTile + slot];
      const double aj = sk != 0.0 ? (double)pr.alpha[j] : 0.0;
      const float tt = (float)((double)pr.W[e] * (acc1 - aj * sk));
      tl[slot][i & 15] = tt;
      if (proposed_any && pr.alpha) corr += (double)pr.alpha[j] * (double)tt;
      if ((i & 15) == 0) cl[slot] = j;
    }
    __syncthreads();
    for (uint32_t q = 0; q < cnt; q += 4) {
      const uint32_t slot = q + lg;
      const bool v = slot < cnt;
      const float b = v ? tl[slot][li] : 0.0f;
      const T* col = A + (uint64_t)(v ? cl[slot] : 0u) * pr.ld;
      u32x4 a[PASSES];
#pragma unroll
      for (uint32_t m = 0; m < PASSES; ++m)
        a[m] = v ? load_a16<T>(col, wrow + m * PR + V * li, d) : u32x4{0u, 0u, 0u, 0u};
#pragma unroll
      for (uint32_t m = 0; m < PASSES; ++m) {
#pragma unroll
        for (uint32_t u = 0; u < V; ++u) {
          float av;
          if constexpr (V == 4) av = __uint_as_float(a[m][u]);
          else av = bf16_bits_to_f32((a[m][u >> 1] >> (16 * (u & 1))) & 0xFFFFu);
          acc[m][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b, acc[m][u], 0, 0, 0);
        }
      }
    }
  }
  // partial rows of trial li: row base + V (4 lg + e) + u
  float* out = pr.P2 + ((uint64_t)ch * kBT + li) * pr.p2pitch;
#pragma unroll
  for (uint32_t m = 0; m < PASSES; ++m) {
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
      const uint32_t rbase = wrow + m * PR + V * (4 * lg + e);
#pragma unroll
      for (uint32_t u = 0; u < V; u += 4)
        *reinterpret_cast<float4*>(out + rbase + u) =
            make_float4(acc[m][u][e], acc[m][u + 1][e], acc[m][u + 2][e], acc[m][u + 3][e]);
    }
  }
  if (proposed_any && rb == 0) {
    // thread t accumulated trial t % 16: sum the 16 threads of each trial
    cred[threadIdx.x] = corr;
    __syncthreads();
    if (threadIdx.x < kBT) {
      double c = 0.0;
      for (uint32_t i = threadIdx.x; i < 256; i += kBT) c += cred[i];
      pr.P2s[(uint64_t)ch * kBT + threadIdx.x] = c;
    }
  }
}

// ---------------------------------------------------------------------------
// Finish the products (dotProduct_and_generate_x, common.h:89-154, and the
// proposed rank-2 part, core_proposedAMM.cuh:198-212): y_k = base_k +
// sum_ch P2[ch][k] with base 0 (exact), y_memo (previous) or y_memo + s v + q
// - sum_ch P2s[ch][k] (proposed); block partials of y^T x, y^T y, y^T v*;
// exact steps with memo snapshot [x, y] -> [x_memo, y_memo].
// Items: nvblk per problem; thread (k, 4 rows).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bfinish(const BProb* __restrict__ P, const uint32_t* off,
                                                 BRun r) {
  const uint32_t p = find_problem(off, r.np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t blk = blockIdx.x - off[p];
  const uint32_t k = threadIdx.x >> 4;
  const uint32_t row = blk * kBVecRows + (threadIdx.x & 15) * 4;
  const uint64_t o = (uint64_t)k * pr.dstage + row;
  const uint32_t md = k < r.ntr ? pr.bt[k].mode : kModeExactNoMemo;
  const bool memo_base = is_amm(md);
  const bool prop = md == kModeAmmProposed;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t ch = 0; ch < pr.chunks; ++ch) {
    const float4 q = *reinterpret_cast<const float4*>(pr.P2 + ((uint64_t)ch * kBT + k) * pr.p2pitch + row);
    a[0] += q.x;
    a[1] += q.y;
    a[2] += q.z;
    a[3] += q.w;
  }
  double sx = 0.0, qx = 0.0, corr = 0.0;
  if (prop) {
    sx = pr.bt[k].s;
    qx = pr.bt[k].q;
    for (uint32_t ch = 0; ch < pr.chunks; ++ch) corr += pr.P2s[(uint64_t)ch * kBT + k];
  }
  float4 ym4 = make_float4(0.f, 0.f, 0.f, 0.f);
  if (memo_base) ym4 = *reinterpret_cast<const float4*>(pr.YM + o);
  const float ymv[4] = {ym4.x, ym4.y, ym4.z, ym4.w};
  const float4 x4 = *reinterpret_cast<const float4*>(pr.X + o);
  const float xv[4] = {x4.x, x4.y, x4.z, x4.w};
  float yv[4];
  double yx = 0.0, yy = 0.0, yw = 0.0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    double acc = memo_base ? (double)ymv[e] : 0.0;
    if (prop && row + e < pr.d) acc += sx * (double)pr.v[row + e] + qx - corr;
    acc += a[e];
    yv[e] = (row + e < pr.d && k < r.ntr) ? (float)acc : 0.0f;
    const double yd = yv[e];
    yx += yd * (double)xv[e];
    yy += yd * yd;
    if (pr.eigvec && row + e < pr.d) yw += yd * pr.eigvec[row + e];
  }
  *reinterpret_cast<float4*>(pr.Y + o) = make_float4(yv[0], yv[1], yv[2], yv[3]);
  if (md == kModeExactMemo && k < r.ntr) {
    *reinterpret_cast<float4*>(pr.XM + o) = x4;
    *reinterpret_cast<float4*>(pr.YM + o) = make_float4(yv[0], yv[1], yv[2], yv[3]);
  }
  yx = sum16(yx);
  yy = sum16(yy);
  yw = sum16(yw);
  if ((threadIdx.x & 15) == 0) {
    double* b = pr.BP + ((uint64_t)blk * kBT + k) * 3;
    b[0] = yx;
    b[1] = yy;
    b[2] = yw;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_bbegin(const BProb* P, BGroup g, BRun r, hipStream_t s) {
  if (g.total) k_bbegin<<<g.total, 256, 0, s>>>(P, g.off, r);
  return hipGetLastError();
}

hipError_t launch_bx0(const BProb* P, BGroup g, BRun r, hipStream_t s) {
  if (g.total) k_bx0<<<g.total, 256, 0, s>>>(P, g.off, r);
  return hipGetLastError();
}

hipError_t launch_btail(const BProb* P, BRun r, uint32_t op, uint32_t it, hipStream_t s) {
  k_btail<<<r.np, 256, 0, s>>>(P, r, op, it);
  return hipGetLastError();
}

hipError_t launch_bnorm(const BProb* P, BGroup g, BRun r, hipStream_t s) {
  if (g.total) k_bnorm<<<g.total, 256, 0, s>>>(P, g.off, r);
  return hipGetLastError();
}

hipError_t launch_bsample(const BProb* P, BGroup g, BRun r, uint32_t it, hipStream_t s) {
  if (g.total) k_bsample<<<g.total, 256, 0, s>>>(P, g.off, r, it);
  return hipGetLastError();
}

hipError_t launch_bprep(const BProb* P, BGroup g, BRun r, hipStream_t s) {
  if (g.total) k_bprep<<<g.total, 256, 0, s>>>(P, g.off, r);
  return hipGetLastError();
}

hipError_t launch_bdot(int dtype, const BProb* P, BGroup g, uint32_t np, hipStream_t s) {
  if (!g.total) return hipSuccess;
  if (dtype == 1)
    k_bdot<__hip_bfloat16><<<g.total, 256, 0, s>>>(P, g.off, np);
  else
    k_bdot<float><<<g.total, 256, 0, s>>>(P, g.off, np);
  return hipGetLastError();
}

hipError_t launch_baxpy(int dtype, uint32_t passes, const BProb* P, BGroup g, uint32_t np,
                        uint32_t proposed_any, hipStream_t s) {
  if (!g.total) return hipSuccess;
  if (dtype == 1) {
    if (passes == 1)
      k_baxpy<__hip_bfloat16, 1><<<g.total, 256, 0, s>>>(P, g.off, np, proposed_any);
    else
      k_baxpy<__hip_bfloat16, 2><<<g.total, 256, 0, s>>>(P, g.off, np, proposed_any);
  } else {
    if (passes == 1)
      k_baxpy<float, 1><<<g.total, 256, 0, s>>>(P, g.off, np, proposed_any);
    else
      k_baxpy<float, 2><<<g.total, 256, 0, s>>>(P, g.off, np, proposed_any);
  }
  return hipGetLastError();
}

hipError_t launch_bfinish(const BProb* P, BGroup g, BRun r, hipStream_t s) {
  if (g.total) k_bfinish<<<g.total, 256, 0, s>>>(P, g.off, r);
  return hipGetLastError();
}

}  // namespace amm
