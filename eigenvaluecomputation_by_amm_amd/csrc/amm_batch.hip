// amm_batch.hip -- kernels of the batched AMM power method (gfx950, wave64).
//
// Trial k of problem p is, sample for sample, the single-vector run with
// Philox key seed_p and counter word 2 = trial0 + k (k_sample, k_begin in
// amm_kernels.hip): the same columns, multiplicities and weights; only the
// summation order of the contractions differs.
#include "amm_batch.hpp"
#include "amm_setup.hpp"
#include "amm_resident.hpp"

namespace amm {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr uint32_t kMaxSlicesDev = 16;  // contraction-1 row slices at most

// Problem of workgroup b of a grouped launch: the group's offsets
// off[0 .. np] (off[p] = first workgroup of problem p) are followed by a
// map off[np + 1 + b] = problem of workgroup b, so a workgroup finds its
// problem with one load (a binary search over the offsets made each
// workgroup wait on about log2(np) dependent loads before its first read
// of A; with 256 problems that was a sizeable share of a short workgroup).
__device__ __forceinline__ uint32_t find_problem(const uint32_t* __restrict__ off, uint32_t np,
                                                 uint32_t b) {
  return off[np + 1 + b];
}

// sum over the 16 lanes of a 16-lane group (groups aligned to 16 lanes)
template <typename T>
__device__ __forceinline__ T sum16(T v) {
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 16);
  return v;
}

__device__ __forceinline__ bool is_amm(uint32_t md) {
  return md == kModeAmmPrevious || md == kModeAmmProposed;
}

// the products of this problem are k_bgram's (a one-trial batch, columns
// that k_bgram streams well)
__device__ __forceinline__ bool gram_step(const BProb& pr) { return pr.gram_chunks != 0; }

// ---------------------------------------------------------------------------
// Run prologue: x_memo = y_memo = 0 (zero_clear, device_memory.h:345-355),
// per-trial state, start stamp, and x0 ~ N(0, I) per trial in the same
// launch: Box-Muller on Philox counter (m, 0, trial, x0 stream) for rows
// 2m, 2m+1 (as k_begin); raw vector into Y, per-block sums of squares into
// BP[., ., 1].  Items: nvblk row blocks per problem.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bbegin(const BProb* __restrict__ P, const uint32_t* off,
                                                BRun r) {
  const uint32_t p = find_problem(off, r.np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t blk = blockIdx.x - off[p];
  const uint32_t k = threadIdx.x >> 4;
  const uint32_t k0 = (uint32_t)pr.seed, k1 = (uint32_t)(pr.seed >> 32);
  double yy = 0.0;
  for (uint32_t sub = 0; sub < kBVecRows / 64; ++sub) {
    const uint32_t row = blk * kBVecRows + sub * 64 + (threadIdx.x & 15) * 4;
    const uint64_t o = (uint64_t)k * pr.dstage + row;
    *reinterpret_cast<float4*>(pr.XM + o) = make_float4(0.f, 0.f, 0.f, 0.f);
    *reinterpret_cast<float4*>(pr.YM + o) = make_float4(0.f, 0.f, 0.f, 0.f);
    float g[4] = {0.f, 0.f, 0.f, 0.f};
    if (k < r.ntr) {
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
    *reinterpret_cast<float4*>(pr.Y + o) = make_float4(g[0], g[1], g[2], g[3]);
#pragma unroll
    for (int e = 0; e < 4; ++e) yy += (double)g[e] * (double)g[e];
  }
  yy = sum16(yy);
  if ((threadIdx.x & 15) == 0) {
    double* b = pr.BP + ((uint64_t)blk * kBT + k) * 3;
    b[0] = 0.0;
    b[1] = yy;
    b[2] = 0.0;
  }
  if (blk == 0 && threadIdx.x < kBT) {
    BTrial t{};
    t.mode = r.modes_static[0];
    pr.bt[threadIdx.x] = t;
    if (threadIdx.x == 0) pr.st->status = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *r.t_start = wall_clock64();
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
#pragma unroll 8
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
  // lane group threadIdx.x >> 4 takes trial threadIdx.x >> 4 of the
  // workgroup's slab, or, with one trial, trial 0 of slab 16 local +
  // (threadIdx.x >> 4); a group past the last slab leaves (no barrier here)
  const uint32_t local = blockIdx.x - off[p];
  const bool one = r.ntr == 1;
  const uint32_t k = one ? 0u : threadIdx.x >> 4;
  const uint32_t blk = one ? local * 16u + (threadIdx.x >> 4) : local;
  if (blk >= pr.nvblk) return;
  const BTrial& t = pr.bt[k];
  const double scale = k < r.ntr ? t.scale : 0.0;
  const uint32_t md = k < r.ntr ? t.mode : kModeExactNoMemo;
  const bool memo = is_amm(md);
  double s = 0.0, q = 0.0;
  for (uint32_t sub = 0; sub < kBVecRows / 64; ++sub) {
    const uint32_t row = blk * kBVecRows + sub * 64 + (threadIdx.x & 15) * 4;
    const uint64_t o = (uint64_t)k * pr.dstage + row;
    const float4 y4 = k < r.ntr ? *reinterpret_cast<const float4*>(pr.Y + o) : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 m4 = make_float4(0.f, 0.f, 0.f, 0.f);
    if (memo) m4 = *reinterpret_cast<const float4*>(pr.XM + o);
    const float yv[4] = {y4.x, y4.y, y4.z, y4.w};
    const float mv[4] = {m4.x, m4.y, m4.z, m4.w};
    float xv[4], dv[4];
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
    // dx in the trial-interleaved staging layout (dx_offset)
    if (pr.dtype == 1) {
      const uint64_t od = dx_offset<32>(k, row);
      uint16_t h[4], l[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        h[e] = bf16_rne(dv[e]);
        l[e] = bf16_rne(dv[e] - bf16_bits_to_f32(h[e]));
      }
      *reinterpret_cast<uint2*>(pr.DXh + od) =
          make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
      *reinterpret_cast<uint2*>(pr.DXl + od) =
          make_uint2((uint32_t)l[0] | ((uint32_t)l[1] << 16), (uint32_t)l[2] | ((uint32_t)l[3] << 16));
    } else {
      *reinterpret_cast<float4*>(pr.DX + dx_offset<16>(k, row)) = make_float4(dv[0], dv[1], dv[2], dv[3]);
    }
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
// counts the draw in trial k's row of 16-bit counts, cnt[k][j].  One
// workgroup per (problem, trial) draws all c slots, 2048 per pass, and
// rewrites the trial's whole row (so rows need no clearing between
// iterations; a row is read only while its trial samples):
//   n <= kDenseCols: the row is counted in LDS (16-bit halves of 32-bit
//     words, LDS atomics) and written out whole;
//   otherwise: an LDS open-addressing table (4096 entries, at most half
//     full for 2048 draws) counts each pass, the row is zero-filled first
//     and each pass adds its distinct columns to it.
// (Scattered 4-byte count updates into a column-major [n][16] array cost
// half of the sampler's time on 256 small problems: every update was a
// partial write of its own 64-byte row.)  The workgroup of a proposed
// trial also finishes s, q of its product.
// ---------------------------------------------------------------------------
constexpr uint32_t kSlotsPerPass = kBSampleSlots;
constexpr uint32_t kHash = 2 * kBSampleSlots;
constexpr uint32_t kSampleThreads = kBSampleSlots / 2;
static_assert(kDenseCols / 2 >= 2 * kHash, "the hash table shares the dense row's LDS");

__device__ __forceinline__ void hash_count(uint32_t* hkey, uint32_t* hcnt, uint32_t j) {
  const uint32_t key = j + 1u;
  uint32_t h = (j * 2654435761u) >> 20;  // top 12 bits: [0, 4096)
  while (true) {
    const uint32_t old = atomicCAS(&hkey[h], 0u, key);
    if (old == 0u || old == key) {
      atomicAdd(&hcnt[h], 1u);
      return;
    }
    h = (h + 1u) & (kHash - 1u);
  }
}

__global__ __launch_bounds__(kSampleThreads) void k_bsample(const BProb* __restrict__ P, const uint32_t* off,
                                                            BRun r, uint32_t it) {
  __shared__ double red[2][kSampleThreads / 64];
  __shared__ uint64_t cs[kCoarse];
  __shared__ __align__(16) uint32_t lc[kDenseCols / 2];  // dense row, or hash keys | counts
  const uint32_t p = find_problem(off, r.np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t k = blockIdx.x - off[p];
  if (k >= r.ntr) return;
  BTrial& t = pr.bt[k];
  const uint32_t md = t.mode;
  if (!is_amm(md)) return;
  const uint32_t tid = threadIdx.x;
  const int which = md == kModeAmmProposed ? 1 : 0;
  if (which) {
    double sv = 0.0, qv = 0.0;
    for (uint32_t blk = tid; blk < pr.nvblk; blk += kSampleThreads) {
      const double* bq = pr.BQ + ((uint64_t)blk * kBT + k) * 2;
      sv += bq[0];
      qv += bq[1];
    }
    sv = wave_sum(sv);
    qv = wave_sum(qv);
    if ((tid & 63) == 0) {
      red[0][tid >> 6] = sv;
      red[1][tid >> 6] = qv;
    }
    __syncthreads();
    if (tid == 0) {
      double ts = 0.0, tq = 0.0;
      for (uint32_t w = 0; w < kSampleThreads / 64; ++w) {
        ts += red[0][w];
        tq += red[1][w];
      }
      t.s = ts;
      t.q = tq;
    }
  }
  const uint32_t n = pr.n, c = pr.c;
  const uint32_t npad = pr.ntiles * kBTile;
  const bool dense = npad <= kDenseCols;
  uint16_t* __restrict__ row = pr.cnt + (uint64_t)k * npad;
  const uint64_t* cdf = which ? pr.cdf1 : pr.cdf0;
  coarse_fill(cs, cdf, n, tid, kSampleThreads);
  uint32_t* hkey = lc;
  uint32_t* hcnt = lc + kHash;
  if (dense) {
    for (uint32_t e = tid; e < npad / 2; e += kSampleThreads) lc[e] = 0u;
  } else {
    for (uint32_t e = tid; e < 2 * kHash; e += kSampleThreads) lc[e] = 0u;
    // the row is zero-filled before any pass adds to it (npad: multiple of 64)
    uint4* r4 = reinterpret_cast<uint4*>(row);
    for (uint32_t e = tid; e < npad / 8; e += kSampleThreads) r4[e] = make_uint4(0u, 0u, 0u, 0u);
  }
  const uint64_t total = pr.st->total[which];
  const uint64_t last_pos = pr.st->last_pos[which];
  const uint32_t k0 = (uint32_t)pr.seed, k1 = (uint32_t)(pr.seed >> 32);
  __syncthreads();  // the coarse table, the cleared counts, the zeroed row
  for (uint32_t base = 0; base < c; base += kSlotsPerPass) {  // uniform over the workgroup
    const uint32_t pair = base / 2 + tid;
    const uint32_t slot = 2 * pair;
    if (slot < c) {
      uint32_t q[4];
      philox4x32_10(pair, it, r.trial0 + k, kStreamSample, k0, k1, q);
      const uint32_t j0 = bcdf_search(cdf, cs, n, draw_target(q[0], q[1], total), last_pos);
      const bool two = slot + 1 < c;
      const uint32_t j1 = two ? bcdf_search(cdf, cs, n, draw_target(q[2], q[3], total), last_pos) : 0u;
      if (dense) {
        atomicAdd(&lc[j0 >> 1], 1u << ((j0 & 1u) << 4));
        if (two) atomicAdd(&lc[j1 >> 1], 1u << ((j1 & 1u) << 4));
      } else {
        hash_count(hkey, hcnt, j0);
        if (two) hash_count(hkey, hcnt, j1);
      }
    }
    if (!dense) {
      __syncthreads();
      // this pass's distinct columns added to the row (one writer per
      // column: this workgroup), the table cleared for the next pass
      for (uint32_t e = tid; e < kHash; e += kSampleThreads) {
        const uint32_t key = hkey[e];
        if (key) {
          uint16_t* cp = row + (key - 1u);
          *cp = (uint16_t)(*cp + hcnt[e]);
          hkey[e] = 0u;
          hcnt[e] = 0u;
        }
      }
      __syncthreads();
    }
  }
  if (dense) {
    __syncthreads();
    const uint4* l4 = reinterpret_cast<const uint4*>(lc);
    uint4* r4 = reinterpret_cast<uint4*>(row);
    for (uint32_t e = tid; e < npad / 8; e += kSampleThreads) r4[e] = l4[e];
  }
}

// ---------------------------------------------------------------------------
// Union of the trials' samples, tile by tile of 64 columns: column j is
// present when a sampled trial drew it (its row's count is nonzero) or a
// trial takes an exact step.  The present columns of a tile are compacted
// in column order into list[tile*64 ...] and counted in tcount[tile].  The
// per-trial weights are formed later from the same draw counts
// (k_bweight), where the products need them.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bprep(const BProb* __restrict__ P, const uint32_t* off,
                                               BRun r) {
  __shared__ uint32_t exact_any;
  const uint32_t p = find_problem(off, r.np, blockIdx.x);
  const BProb& pr = P[p];
  if (threadIdx.x == 0) exact_any = 0u;
  __syncthreads();
  if (threadIdx.x < r.ntr) {
    const uint32_t md = pr.bt[threadIdx.x].mode;
    if (md == kModeExactMemo || md == kModeExactNoMemo) exact_any = 1u;
  }
  __syncthreads();
  const bool all_present = exact_any != 0u;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nwaves = (off[p + 1] - off[p]) * 4;
  const uint32_t npad = pr.ntiles * kBTile;
  const uint32_t ngroups = (pr.ntiles + 7) / 8;
  // A wave takes 8 tiles (512 columns): lane l holds columns 8 l .. 8 l + 7
  // of the group (one 16-byte load per trial row); the 8 lanes of a tile
  // add up their presence counts (prefix over the lane octet) and write the
  // tile's compacted list.  Wave-uniform loop over groups w, w + nwaves, ...
  for (uint32_t grp = (blockIdx.x - off[p]) * 4 + (threadIdx.x >> 6); grp < ngroups; grp += nwaves) {
    const uint32_t tile = grp * 8 + (lane >> 3);
    const uint32_t j0 = grp * 512 + lane * 8;  // < npad whenever tile < ntiles
    uint32_t bits = 0u;
    if (tile < pr.ntiles) {
      if (all_present) {
        bits = 0xFFu;
      } else {
        // every live trial samples (no trial is exact): its row is current
        uint4 v[kBT];
#pragma unroll
        for (uint32_t k = 0; k < kBT; ++k)
          v[k] = k < r.ntr ? *reinterpret_cast<const uint4*>(pr.cnt + (uint64_t)k * npad + j0) : make_uint4(0u, 0u, 0u, 0u);
        uint4 o = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (uint32_t k = 0; k < kBT; ++k) {
          o.x |= v[k].x;
          o.y |= v[k].y;
          o.z |= v[k].z;
          o.w |= v[k].w;
        }
        const uint32_t w4[4] = {o.x, o.y, o.z, o.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          bits |= ((w4[q] & 0xFFFFu) != 0u ? 1u : 0u) << (2 * q);
          bits |= ((w4[q] >> 16) != 0u ? 1u : 0u) << (2 * q + 1);
        }
      }
      // columns at or beyond n are never present
      if (j0 + 8 > pr.n) bits &= j0 >= pr.n ? 0u : (1u << (pr.n - j0)) - 1u;
    }
    const uint32_t cntl = (uint32_t)__popc(bits);
    // inclusive prefix of the counts over the lane octet of this tile
    uint32_t incl = cntl;
#pragma unroll
    for (uint32_t o = 1; o < 8; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o, 8);
      if ((lane & 7u) >= o) incl += v;
    }
    if (tile < pr.ntiles) {
      uint32_t pos = incl - cntl;
      uint32_t* out = pr.list + (uint64_t)tile * kBTile;
#pragma unroll
      for (uint32_t b2 = 0; b2 < 8; ++b2)
        if ((bits >> b2) & 1u) out[pos++] = j0 + b2;
      if ((lane & 7u) == 7u) pr.tcount[tile] = incl;
    }
  }
}

// ---------------------------------------------------------------------------
// 16 bytes of column `col` from row `row` (rows at or beyond d read 0).
// ---------------------------------------------------------------------------
template <typename T, bool NT>
__device__ __forceinline__ u32x4 load_a16(const T* col, uint32_t row, uint32_t d) {
  constexpr uint32_t V = Elem<T>::VEC;
  if (row + V <= d) return load16<NT>(col + row);
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
// Union prefix (one 1024-thread workgroup per problem): tpre[t] = number of
// present columns in tiles < t; tpre[ntiles] = |U|.  Compacted slot s of
// the union is slot s - tpre[t] of tile t = max{t : tpre[t] <= s}.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_bscan(const BProb* __restrict__ P) {
  __shared__ uint32_t sc[2][1024];
  const BProb& pr = P[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  const uint32_t per = (pr.ntiles + 1023) / 1024;
  const uint32_t b = tid * per;
  uint32_t sum = 0;
  for (uint32_t i = 0; i < per; ++i)
    if (b + i < pr.ntiles) sum += pr.tcount[b + i];
  sc[0][tid] = sum;
  __syncthreads();
  int cur = 0;
  for (uint32_t o = 1; o < 1024; o <<= 1) {
    const uint32_t v = tid >= o ? sc[cur][tid] + sc[cur][tid - o] : sc[cur][tid];
    sc[cur ^ 1][tid] = v;
    cur ^= 1;
    __syncthreads();
  }
  uint32_t run = sc[cur][tid] - sum;  // exclusive prefix
  for (uint32_t i = 0; i < per; ++i) {
    if (b + i < pr.ntiles) {
      pr.tpre[b + i] = run;
      run += pr.tcount[b + i];
    }
  }
  if (tid == 1023) pr.tpre[pr.ntiles] = sc[cur][1023];
}

// ---------------------------------------------------------------------------
// Compacted union: tile t's present columns move to union slots tpre[t] ...
// (the tiles of k_bprep's launch: waves take tiles w, w + waves, ...).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bcompact(const BProb* __restrict__ P, const uint32_t* off,
                                                  uint32_t np) {
  const uint32_t p = find_problem(off, np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nwaves = (off[p + 1] - off[p]) * 4;
  // the tiles of k_bprep's launch shape: wave w takes tiles w, w + nwaves, ...
  for (uint32_t tile = (blockIdx.x - off[p]) * 4 + (threadIdx.x >> 6); tile < pr.ntiles; tile += nwaves) {
    const uint32_t cnt = pr.tcount[tile];
    const uint32_t base = pr.tpre[tile];
    if (lane < cnt) pr.colc[base + lane] = pr.list[(uint64_t)tile * kBTile + lane];
  }
}

// ---------------------------------------------------------------------------
// The three passes above in one launch, one 1024-thread workgroup per problem
// (for problems of at most kUnionMaxTiles tiles whose count rows are small:
// a lone workgroup reads them all).  Wave w takes the 512-column groups w,
// w + 16, ...: lane l finds the presence of columns 8 l .. 8 l + 7 of the
// group as k_bprep does and keeps the bits in LDS, the lane octet of a tile
// adds up its count; the tile counts are scanned into tpre (k_bscan); then
// each lane writes its present columns to their union slots tpre[t] +
// (position within the tile), the order k_bprep + k_bcompact produce.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_bunion(const BProb* __restrict__ P, BRun r) {
  __shared__ uint32_t exact_any;
  __shared__ uint8_t bits_l[kUnionMaxTiles * 8];  // presence of 8 columns per byte
  __shared__ uint32_t tc[kUnionMaxTiles];         // tile counts, then their exclusive prefix
  __shared__ uint32_t wsum[16];
  const BProb& pr = P[blockIdx.x];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) exact_any = 0u;
  __syncthreads();
  if (tid < r.ntr) {
    const uint32_t md = pr.bt[tid].mode;
    if (md == kModeExactMemo || md == kModeExactNoMemo) exact_any = 1u;
  }
  __syncthreads();
  const bool all_present = exact_any != 0u;
  const uint32_t nt = pr.ntiles;
  const uint32_t npad = nt * kBTile;
  const uint32_t ngroups = (nt + 7) / 8;
  for (uint32_t grp = wave; grp < ngroups; grp += 16) {
    const uint32_t tile = grp * 8 + (lane >> 3);
    const uint32_t j0 = grp * 512 + lane * 8;
    uint32_t bits = 0u;
    if (tile < nt) {
      if (all_present) {
        bits = 0xFFu;
      } else {
        uint4 o = make_uint4(0u, 0u, 0u, 0u);
        for (uint32_t k = 0; k < r.ntr; ++k) {
          const uint4 v = *reinterpret_cast<const uint4*>(pr.cnt + (uint64_t)k * npad + j0);
          o.x |= v.x;
          o.y |= v.y;
          o.z |= v.z;
          o.w |= v.w;
        }
        const uint32_t w4[4] = {o.x, o.y, o.z, o.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          bits |= ((w4[q] & 0xFFFFu) != 0u ? 1u : 0u) << (2 * q);
          bits |= ((w4[q] >> 16) != 0u ? 1u : 0u) << (2 * q + 1);
        }
      }
      if (j0 + 8 > pr.n) bits &= j0 >= pr.n ? 0u : (1u << (pr.n - j0)) - 1u;
      bits_l[grp * 64 + lane] = (uint8_t)bits;
    }
    uint32_t incl = (uint32_t)__popc(bits);
#pragma unroll
    for (uint32_t o = 1; o < 8; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o, 8);
      if ((lane & 7u) >= o) incl += v;
    }
    if (tile < nt && (lane & 7u) == 7u) tc[tile] = incl;
  }
  __syncthreads();
  // exclusive prefix of the tile counts (nt <= 1024: a tile per thread)
  const uint32_t v = tid < nt ? tc[tid] : 0u;
  uint32_t incl = v;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t base = 0, total = 0;
  for (uint32_t w = 0; w < 16; ++w) {
    base += w < wave ? wsum[w] : 0u;
    total += wsum[w];
  }
  if (tid < nt) {
    tc[tid] = base + incl - v;
    pr.tpre[tid] = base + incl - v;
  }
  if (tid == 0) pr.tpre[nt] = total;
  __syncthreads();
  for (uint32_t grp = wave; grp < ngroups; grp += 16) {
    const uint32_t tile = grp * 8 + (lane >> 3);
    const uint32_t bits = tile < nt ? (uint32_t)bits_l[grp * 64 + lane] : 0u;
    const uint32_t cntl = (uint32_t)__popc(bits);
    uint32_t in8 = cntl;
#pragma unroll
    for (uint32_t o = 1; o < 8; o <<= 1) {
      const uint32_t t = __shfl_up(in8, o, 8);
      if ((lane & 7u) >= o) in8 += t;
    }
    if (tile < nt) {
      uint32_t pos = tc[tile] + in8 - cntl;
      const uint32_t j0 = grp * 512 + lane * 8;
#pragma unroll
      for (uint32_t b2 = 0; b2 < 8; ++b2)
        if ((bits >> b2) & 1u) pr.colc[pos++] = j0 + b2;
    }
  }
}

// ---------------------------------------------------------------------------
// LDS tile of C columns x RB rows staged from HBM with long contiguous runs:
// each column's RB rows are read by RB*sizeof(T)/16 consecutive lanes (512
// or 1024 contiguous bytes), prefetched into registers one round ahead and
// stored to LDS at column pitch RB*sizeof(T) + PAD bytes.
// ---------------------------------------------------------------------------
template <typename T, uint32_t C, uint32_t RB, bool NT, uint32_t PAD = 16>
struct TileStage {
  static constexpr uint32_t kColBytes = RB * (uint32_t)sizeof(T);
  static constexpr uint32_t kLanesPerCol = kColBytes / 16;
  static constexpr uint32_t kColsPerPass = 256 / kLanesPerCol;
  static constexpr uint32_t kLoads = C / kColsPerPass;
  static constexpr uint32_t kPitch = kColBytes + PAD;
  static constexpr uint32_t kBytes = C * kPitch;
  static constexpr uint32_t kElems16 = 16 / (uint32_t)sizeof(T);
  u32x4 r[kLoads];

  // cols: C column indices in LDS (0xFFFFFFFF: absent column, reads 0).
  // load() issues every load unconditionally, from an address that is valid
  // whatever the bounds (row 0 of column 0 for an absent column, row 0 of
  // its column for a lane past d); store() then zeroes absent columns and
  // reloads the lanes past or straddling d element by element, when the
  // round's data is consumed anyway.  (A per-load branch in load() made a
  // partial round wait for each load in turn: twice the time of k_baxpy at
  // d not a multiple of its rows; fixing the lanes up in load() made every
  // round wait for its loads before its MFMAs: 13% slower at full rounds.)
  uint32_t lrow = 0;   // this lane's first row of the loaded round
  uint32_t amask = 0;  // absent columns of the loaded round, bit q
  __device__ __forceinline__ void load(const T* __restrict__ A, uint64_t ld, const uint32_t* cols,
                                       uint32_t row0, uint32_t d) {
    const uint32_t sub = threadIdx.x % kLanesPerCol, c0 = threadIdx.x / kLanesPerCol;
    lrow = row0 + sub * kElems16;
    const bool full = lrow + kElems16 <= d;
    amask = 0u;
#pragma unroll
    for (uint32_t q = 0; q < kLoads; ++q) {
      const uint32_t jq = cols[q * kColsPerPass + c0];
      amask |= (jq == 0xFFFFFFFFu ? 1u : 0u) << q;
      const uint32_t j = jq != 0xFFFFFFFFu ? jq : 0u;
      r[q] = load16<NT>(A + (uint64_t)j * ld + (full ? lrow : 0u));
    }
  }
  // cols: the same column indices the matching load() read
  __device__ __forceinline__ void store(char* lds, const T* __restrict__ A, uint64_t ld, const uint32_t* cols,
                                        uint32_t d) {
    const uint32_t sub = threadIdx.x % kLanesPerCol, c0 = threadIdx.x / kLanesPerCol;
    const bool full = lrow + kElems16 <= d;
#pragma unroll
    for (uint32_t q = 0; q < kLoads; ++q) {
      u32x4 v = r[q];
      if ((amask >> q) & 1u) {
        v = u32x4{0u, 0u, 0u, 0u};
      } else if (!full) {
        v = load_a16<T, NT>(A + (uint64_t)cols[q * kColsPerPass + c0] * ld, lrow, d);  // zeros at and beyond d
      }
      *reinterpret_cast<u32x4*>(lds + (q * kColsPerPass + c0) * kPitch + sub * 16) = v;
    }
  }
};

// ---------------------------------------------------------------------------
// Contraction 1 (previousAMM_kernel1 / proposedAMM_kernel1 for 16 trials):
// P1[ks][s][k] = a_{col(s)}[rows of slice ks]^T dx_k for the 64 compacted
// union slots s of virtual tile vt.  Per round the workgroup stages the 64
// columns x RB rows in LDS (RB = 256 BF16 / 128 FP32, 512 contiguous bytes
// per column); wave w multiplies its quarter of the round's rows for all
// four groups of 16 columns: BF16 v_mfma_f32_16x16x32_bf16 with dx = hi +
// lo (two MFMAs per 32 rows), FP32 v_mfma_f32_16x16x4_f32 (four per 16
// rows).  A-operand lane (i, g): column 16G + i, rows 8g.. (BF16) / 4g..
// (FP32) of the step; B-operand lane (i, g): the same rows of dx_i;
// accumulator lane (i, g) element e: slot 16G + 4g + e, trial i.
// ---------------------------------------------------------------------------
template <typename T, bool NT>
__global__ __launch_bounds__(256) void k_bdot(const BProb* __restrict__ P, const uint32_t* off,
                                              uint32_t np, uint32_t skip_gram) {
  constexpr bool BF = Elem<T>::VEC == 8;
  constexpr uint32_t RB = BF ? 256 : 128;
  using Stage = TileStage<T, 64, RB, NT>;
  __shared__ __align__(16) char lds[Stage::kBytes];
  __shared__ uint32_t cols[64];
  const uint32_t p = find_problem(off, np, blockIdx.x);
  const BProb& pr = P[p];
  if (skip_gram && gram_step(pr)) return;
  const uint32_t local = blockIdx.x - off[p];
  // the grid covers the largest vtiles x kslices any union size can need
  // (< dot_blocks + ntiles); item = slice-major over this product's
  // virtual tiles, so the active items are the first blocks, spread evenly
  // over the XCDs (blocks are dealt to the XCDs round-robin)
  const UPlan up = union_plan(pr, pr.tpre[pr.ntiles]);
  if (local >= up.vtiles * up.kslices) return;
  const uint32_t ks = local / up.vtiles, vt = local % up.vtiles;
  const uint32_t U = up.U;
  const uint32_t s0 = vt * kBTile;
  const uint32_t cnt = min(kBTile, U - s0);
  if (threadIdx.x < kBTile) cols[threadIdx.x] = threadIdx.x < cnt ? pr.colc[s0 + threadIdx.x] : 0xFFFFFFFFu;
  __syncthreads();
  const uint32_t ngroups = (cnt + 15) / 16;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t li = lane & 15, lg = lane >> 4;
  const uint32_t r0 = ks * up.kslice_rows;
  const uint32_t r1 = min(r0 + up.kslice_rows, pr.dstage);
  const T* A = reinterpret_cast<const T*>(pr.A);
  f32x4_t acc[4];
#pragma unroll
  for (int G = 0; G < 4; ++G) acc[G] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  // the next round's A tile is prefetched right after this round's store
  // to LDS, this wave's dx rows of the next round right after the MFMAs
  // have used this round's (no register copy: 4 workgroups per CU fit)
  Stage st;
  u32x4 xa[2], xb[2];  // BF16: dx hi, lo of K-steps 0, 1; FP32: dx of K-steps 0, 1 (xa)
  auto load_dx = [&](uint32_t row0) {
#pragma unroll
    for (uint32_t s = 0; s < 2; ++s) {
      // rows of trial li in the trial-interleaved layout: the wave's 64
      // lanes read 1 KB of consecutive bytes
      if constexpr (BF) {
        const uint64_t o = dx_offset<32>(li, row0 + wave * 64 + s * 32 + 8 * lg);
        xa[s] = *reinterpret_cast<const u32x4*>(pr.DXh + o);
        xb[s] = *reinterpret_cast<const u32x4*>(pr.DXl + o);
      } else {
        xa[s] = *reinterpret_cast<const u32x4*>(pr.DX + dx_offset<16>(li, row0 + wave * 32 + s * 16 + 4 * lg));
      }
    }
  };
  if (r0 < r1) {
    st.load(A, pr.ld, cols, r0, pr.d);
    load_dx(r0);
  }
  for (uint32_t row0 = r0; row0 < r1; row0 += RB) {
    st.store(lds, A, pr.ld, cols, pr.d);
    __syncthreads();
    const bool more = row0 + RB < r1;
    if (more) st.load(A, pr.ld, cols, row0 + RB, pr.d);
    if constexpr (BF) {
#pragma unroll
      for (uint32_t s = 0; s < 2; ++s) {
        const uint32_t rr = wave * 64 + s * 32 + 8 * lg;  // row within the round
        const bf16x8_t bh = __builtin_bit_cast(bf16x8_t, xa[s]);
        const bf16x8_t bl = __builtin_bit_cast(bf16x8_t, xb[s]);
#pragma unroll
        for (uint32_t G = 0; G < 4; ++G) {
          if (G < ngroups) {
            const bf16x8_t a = __builtin_bit_cast(
                bf16x8_t, *reinterpret_cast<const u32x4*>(lds + (G * 16 + li) * Stage::kPitch + rr * 2));
            acc[G] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh, acc[G], 0, 0, 0);
            acc[G] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl, acc[G], 0, 0, 0);
          }
        }
      }
    } else {
#pragma unroll
      for (uint32_t s = 0; s < 2; ++s) {
        const uint32_t rr = wave * 32 + s * 16 + 4 * lg;
        const u32x4 b = xa[s];
#pragma unroll
        for (uint32_t G = 0; G < 4; ++G) {
          if (G < ngroups) {
            const u32x4 a = *reinterpret_cast<const u32x4*>(lds + (G * 16 + li) * Stage::kPitch + rr * 4);
#pragma unroll
            for (int u = 0; u < 4; ++u)
              acc[G] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a[u]), __uint_as_float(b[u]),
                                                            acc[G], 0, 0, 0);
          }
        }
      }
    }
    if (more) load_dx(row0 + RB);
    __syncthreads();
  }
  // cross-wave sum in LDS (reusing the tile buffer), one coalesced write
  float* red = reinterpret_cast<float*>(lds);  // [4][64][17]
  static_assert(4 * 64 * 17 * 4 <= Stage::kBytes, "reduction buffer");
#pragma unroll
  for (int G = 0; G < 4; ++G) {
#pragma unroll
    for (int e = 0; e < 4; ++e) red[(wave * 64 + G * 16 + 4 * lg + e) * 17 + li] = acc[G][e];
  }
  __syncthreads();
  float* out = pr.P1 + ((uint64_t)ks * pr.ntiles * kBTile + s0) * kBT;
  for (uint32_t i = threadIdx.x; i < cnt * kBT; i += 256) {
    const uint32_t slot = i / kBT, k = i % kBT;
    out[i] = (red[slot * 17 + k] + red[(64 + slot) * 17 + k]) +
             (red[(128 + slot) * 17 + k] + red[(192 + slot) * 17 + k]);
  }
}

// ---------------------------------------------------------------------------
// Weights of the current product, once per union slot (previousAMM_kernel1 /
// proposedAMM_kernel1 epilogue, core_previousAMM.cuh:4-175,
// core_proposedAMM.cuh:17-213): Tc[s][k] = w_{s,k} (sum_ks P1[ks][s][k] -
// alpha_j s_k), s_k = 1^T dx_k for a proposed trial, else 0, with w formed
// from the draw counts as update_weights (common.h:42-53) and k_sample form
// it: 1 on an exact step, mult total / (c q_j) for a draw count mult > 0 on
// a sampled step (q_j = cdf_j - cdf_{j-1} of the step's CDF), else 0.  (Forming t in
// k_bdot's epilogue where it has one row slice measured slower: its reads
// and writes compete with the stream of A.)  With out_w,
// Tc[s][k] = w_{s,k} for every problem instead (the union sample and its
// weights, for amm_batch_sample).  A lane quad per union slot, four trials
// per lane (16-byte accesses of the partials and t; the counts from the
// trials' rows, adjacent columns of a row in adjacent quads); the problem's
// blocks stride over the union, 64 slots per workgroup and pass.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bweight(const BProb* __restrict__ P, const uint32_t* off,
                                                 BRun r, uint32_t proposed_any, uint32_t out_w,
                                                 uint32_t skip_gram) {
  __shared__ double s_sk[kBT];
  const uint32_t p = find_problem(off, r.np, blockIdx.x);
  const BProb& pr = P[p];
  if (skip_gram && gram_step(pr)) return;
  const uint32_t local = blockIdx.x - off[p];
  const uint32_t nblk = off[p + 1] - off[p];
  const UPlan up = union_plan(pr, pr.tpre[pr.ntiles]);
  const uint64_t slab = (uint64_t)pr.ntiles * kBTile * kBT;
  const uint32_t* __restrict__ colc = pr.colc;
  const uint16_t* __restrict__ cnt = pr.cnt;
  const uint32_t npad = pr.ntiles * kBTile;
  const float* __restrict__ P1 = pr.P1;
  float* __restrict__ Tc = pr.Tc;
  const uint32_t ntr = r.ntr, ks = up.kslices;
  // per-trial step kinds as bit masks (uniform over the workgroup); s_k
  // (0 unless a proposed product) in LDS
  uint32_t exact_m = 0u, prop_m = 0u;
#pragma unroll
  for (uint32_t k = 0; k < kBT; ++k) {
    const uint32_t md = k < ntr ? pr.bt[k].mode : kModeExactMemo;
    exact_m |= (md == kModeExactMemo || md == kModeExactNoMemo ? 1u : 0u) << k;
    prop_m |= (md == kModeAmmProposed ? 1u : 0u) << k;
  }
  const uint32_t live_m = ntr >= 32 ? 0xFFFFFFFFu : (1u << ntr) - 1u;
  const uint32_t need0 = live_m & ~exact_m & ~prop_m, need1 = live_m & ~exact_m & prop_m;
  const bool use_sk = !out_w && proposed_any && (live_m & prop_m) != 0u;
  if (threadIdx.x < kBT)
    s_sk[threadIdx.x] = (use_sk && ((live_m & prop_m) >> threadIdx.x) & 1u) ? pr.bt[threadIdx.x].s : 0.0;
  __syncthreads();
  const double tot0 = need0 && pr.st ? pr.st->total[0] : 0.0;
  const double tot1 = need1 && pr.st ? pr.st->total[1] : 0.0;
  const double cd = (double)pr.c;
  const uint32_t nq = (ntr + 3) / 4;  // 4-trial groups holding live trials
  // lane quad per slot: lane g of the quad takes trials 4 g .. 4 g + 3, so
  // the quad's 16-byte accesses of the counts, dots and t are one 64-byte
  // run; the column of the next slot is loaded one iteration ahead (one
  // dependent latency per slot instead of two)
  const uint32_t g = threadIdx.x & 3u;
  const bool gl = g < nq;
  const uint32_t stride = nblk * 64u;
  uint32_t sl = local * 64u + (threadIdx.x >> 2);
  uint32_t jn = sl < up.U ? colc[sl] : 0u;
  for (; sl < up.U; sl += stride) {
    const uint32_t j = jn;
    if (sl + stride < up.U) jn = colc[sl + stride];
    const uint64_t base = (uint64_t)sl * kBT + 4u * g;
    f32x4_t pv = {0.f, 0.f, 0.f, 0.f};  // the first slice's dots
    if (!out_w) pv = *reinterpret_cast<const f32x4_t*>(P1 + base);
    // the counts of trials 4 g .. 4 g + 3 (rows of exact or unused trials
    // are not read: their weights do not use them)
    uint32_t m[4];
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
      const uint32_t k = 4u * g + e;
      m[e] = (gl && ((live_m & ~exact_m) >> k) & 1u) ? (uint32_t)cnt[(uint64_t)k * npad + j] : 0u;
    }
    // q_j of each CDF a sampled trial uses
    uint64_t q0 = 0ull, q1 = 0ull;
    if (need0 && pr.cdf0) q0 = pr.cdf0[j] - (j ? pr.cdf0[j - 1] : 0ull);
    if (need1 && pr.cdf1) q1 = pr.cdf1[j] - (j ? pr.cdf1[j - 1] : 0ull);
    const double aj = use_sk ? (double)pr.alpha[j] : 0.0;
    double acc[4];
    if (!out_w) {
      // p = sum over the row slices of the dots, in slice order
#pragma unroll
      for (uint32_t e = 0; e < 4; ++e) acc[e] = 0.0 + (double)pv[e];
      for (uint32_t q = 1; q < ks; ++q) {
        const f32x4_t v = *reinterpret_cast<const f32x4_t*>(P1 + q * slab + base);
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) acc[e] += (double)v[e];
      }
    }
    f32x4_t t;
#pragma unroll
    for (uint32_t e = 0; e < 4; ++e) {
      const uint32_t k = 4u * g + e;
      const uint32_t mult = m[e];
      // weight: 1 on an exact step, mult total / (c q_j) on a sampled one
      float w = 0.0f;
      if ((exact_m >> k) & 1u) {
        w = 1.0f;
      } else if (mult) {
        const bool pp = (prop_m >> k) & 1u;
        const double num = (double)mult * (pp ? tot1 : tot0);
        const double den = cd * (double)(pp ? q1 : q0);
        w = den > 0.0 ? (float)(num / den) : 0.0f;
      }
      if (!((live_m >> k) & 1u)) w = 0.0f;
      if (!out_w) {
        // t = w (p - alpha_j s_k)
        const double skk = s_sk[k];
        const double ajs = skk != 0.0 ? aj : 0.0;
        w = ((live_m >> k) & 1u) ? (float)((double)w * (acc[e] - ajs * skk)) : 0.0f;
      }
      t[e] = w;
    }
    *reinterpret_cast<f32x4_t*>(Tc + base) = t;
  }
}

// ---------------------------------------------------------------------------
// Contraction 2 (previousAMM_kernel2 / proposedAMM_kernel2 for 16 trials):
// P2[ch][k][rows] = sum_{s in chunk ch} a_{col(s)}[rows] t_{s,k} with
//   t = W (sum_ks P1[ks] - alpha_j s_k)       (s_k = 0 unless proposed)
// Chunk ch = compacted slots [ch cpc, (ch + 1) cpc) of the union (chunks
// beyond |U| exit).  Workgroup = (row block of 256 PASSES rows, chunk).  Per
// round: C columns of the chunk staged in LDS (contiguous runs of 512 / 1024
// bytes per column) and their t in LDS.
//   BF16: C = 32 = one v_mfma_f32_16x16x32_bf16 K step; A-operand lane
//     (i, g) = row i, columns g, g + 4, ..., g + 28 read transposed from LDS
//     (K slot 8g + m = column g + 4m, t in the same order); t = hi + lo
//     (two MFMAs); accumulator lane (i, g) element e = row 4g + e, trial i.
//   FP32: C = 16, v_mfma_f32_16x16x4_f32 on 4 columns: A-operand lane (i, g)
//     = column 4q + g, rows 4i + u (element u of a 16-byte LDS read);
//     accumulator u lane (i, g) element e = row 16g + 4e + u, trial i.
// The proposed scalar sum_j alpha_j t_j of the chunk goes to P2s (row block
// 0 only).
// ---------------------------------------------------------------------------
template <typename T, uint32_t PASSES, bool NT>
__global__ __launch_bounds__(256) void k_baxpy(const BProb* __restrict__ P, const uint32_t* off,
                                               uint32_t np, uint32_t proposed_any, uint32_t skip_gram) {
  constexpr bool BF = Elem<T>::VEC == 8;
  constexpr uint32_t C = BF ? 32 : 16;
  constexpr uint32_t RB = 256 * PASSES;
  constexpr uint32_t WR = 64 * PASSES;  // rows per wave
  // BF16: column pitch = 8 dwords mod 64 banks, and lane group g of the
  // A operand reads columns g, g + 4, ..., g + 28 (K slot 8 g + m holds
  // column g + 4 m; t is stored in the same order): the four groups then
  // read disjoint banks, where columns 8g .. 8g + 7 at a 16-byte pad had
  // groups 0 and 2 (1 and 3) on the same banks
  using Stage = TileStage<T, C, RB, NT, BF ? 32u : 16u>;
  __shared__ __align__(16) char lds[Stage::kBytes];
  __shared__ uint32_t cols[2][C];                // column indices, double-buffered by round
  __shared__ __align__(16) uint16_t th[kBT][C];  // BF16 t, high part [trial][column]
  __shared__ __align__(16) uint16_t tlo[kBT][C];
  __shared__ float tf[C][kBT];                   // FP32 t [column][trial]
  __shared__ double cred[256];
  const uint32_t p = find_problem(off, np, blockIdx.x);
  const BProb& pr = P[p];
  if (skip_gram && gram_step(pr)) return;
  const uint32_t local = blockIdx.x - off[p];
  // chunk-major order: active chunks form one contiguous range of blocks
  const uint32_t ch = local / pr.rblocks, rb = local % pr.rblocks;
  const UPlan up = union_plan(pr, pr.tpre[pr.ntiles]);
  const uint32_t U = up.U;
  if (ch >= up.active) return;
  const uint32_t sb = ch * up.cpc;
  const uint32_t se = min(sb + up.cpc, U);
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t li = lane & 15, lg = lane >> 4;
  const uint32_t R0 = rb * RB;
  const T* A = reinterpret_cast<const T*>(pr.A);
  const uint32_t d = pr.d;
  double corr = 0.0;  // proposed: sum_j alpha_j t_j of this thread's trial (threadIdx.x & 15)
  constexpr uint32_t NACC = BF ? 4 * PASSES : 4 * PASSES;
  f32x4_t acc[NACC];
#pragma unroll
  for (uint32_t m = 0; m < NACC; ++m) acc[m] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  // Every load of round r + 1 (A tile, t, and the column indices of round
  // r + 2) is issued right after round r's first barrier, so it overlaps
  // round r's MFMAs; one barrier pair per round.
  constexpr uint32_t NTE = C * kBT / 256;  // t entries per thread
  const float* Tc = pr.Tc;
  if (threadIdx.x < C) cols[0][threadIdx.x] = (sb + threadIdx.x < se) ? pr.colc[sb + threadIdx.x] : 0xFFFFFFFFu;
  uint32_t colnext = (threadIdx.x < C && sb + C + threadIdx.x < se) ? pr.colc[sb + C + threadIdx.x] : 0xFFFFFFFFu;
  __syncthreads();
  Stage st;
  st.load(A, pr.ld, cols[0], R0, d);
  float tr[NTE];
#pragma unroll
  for (uint32_t q = 0; q < NTE; ++q) {
    const uint32_t i = threadIdx.x + 256 * q, c = i >> 4;
    tr[q] = (sb + c < se) ? Tc[(uint64_t)(sb + c) * kBT + (i & 15)] : 0.0f;
  }
  uint32_t buf = 0;
  for (uint32_t s = sb; s < se; s += C, buf ^= 1u) {
    // this round's t into LDS (BF16: hi + lo), its A tile, next round's columns
#pragma unroll
    for (uint32_t q = 0; q < NTE; ++q) {
      const uint32_t i = threadIdx.x + 256 * q, c = i >> 4, k = i & 15;
      const float tt = tr[q];
      if (proposed_any && rb == 0 && s + c < se) corr += (double)pr.alpha[cols[buf][c]] * (double)tt;
      if constexpr (BF) {
        const uint32_t slot = 8 * (c & 3) + (c >> 2);  // K slot of column c
        const uint16_t h = bf16_rne(tt);
        th[k][slot] = h;
        tlo[k][slot] = bf16_rne(tt - bf16_bits_to_f32(h));
      } else {
        tf[c][k] = tt;
      }
    }
    st.store(lds, A, pr.ld, cols[buf], d);
    if (threadIdx.x < C) cols[buf ^ 1u][threadIdx.x] = colnext;
    __syncthreads();
    const uint32_t sn = s + C;
    if (sn < se) {
      st.load(A, pr.ld, cols[buf ^ 1u], R0, d);
#pragma unroll
      for (uint32_t q = 0; q < NTE; ++q) {
        const uint32_t i = threadIdx.x + 256 * q, c = i >> 4;
        tr[q] = (sn + c < se) ? Tc[(uint64_t)(sn + c) * kBT + (i & 15)] : 0.0f;
      }
      colnext = (threadIdx.x < C && sn + C + threadIdx.x < se) ? pr.colc[sn + C + threadIdx.x] : 0xFFFFFFFFu;
    }
    if constexpr (BF) {
      const bf16x8_t bh = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4*>(&th[li][8 * lg]));
      const bf16x8_t bl = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4*>(&tlo[li][8 * lg]));
      const uint16_t* l16 = reinterpret_cast<const uint16_t*>(lds);
#pragma unroll
      for (uint32_t b = 0; b < 4 * PASSES; ++b) {
        const uint32_t row = wave * WR + b * 16 + li;
        uint32_t w4[4];
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
          const uint32_t c0 = lg + 8 * e;  // K slots 8 lg + 2e, + 1: columns lg + 8e, + 4
          const uint32_t lo16 = l16[(c0 * Stage::kPitch) / 2 + row];
          const uint32_t hi16 = l16[((c0 + 4) * Stage::kPitch) / 2 + row];
          w4[e] = lo16 | (hi16 << 16);
        }
        const bf16x8_t a = __builtin_bit_cast(bf16x8_t, u32x4{w4[0], w4[1], w4[2], w4[3]});
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh, acc[b], 0, 0, 0);
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bl, acc[b], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (uint32_t q = 0; q < C / 4; ++q) {
        const float bv = tf[4 * q + lg][li];
#pragma unroll
        for (uint32_t m = 0; m < PASSES; ++m) {
          const u32x4 a = *reinterpret_cast<const u32x4*>(lds + (4 * q + lg) * Stage::kPitch +
                                                           (wave * WR + m * 64 + 4 * li) * 4);
#pragma unroll
          for (uint32_t u = 0; u < 4; ++u)
            acc[m * 4 + u] =
                __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(a[u]), bv, acc[m * 4 + u], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  float* out = pr.P2 + ((uint64_t)ch * kBT + li) * pr.p2pitch;
  if constexpr (BF) {
#pragma unroll
    for (uint32_t b = 0; b < 4 * PASSES; ++b) {
      const uint32_t row = R0 + wave * WR + b * 16 + 4 * lg;
      *reinterpret_cast<float4*>(out + row) = make_float4(acc[b][0], acc[b][1], acc[b][2], acc[b][3]);
    }
  } else {
#pragma unroll
    for (uint32_t m = 0; m < PASSES; ++m) {
#pragma unroll
      for (uint32_t e = 0; e < 4; ++e) {
        const uint32_t row = R0 + wave * WR + m * 64 + 16 * lg + 4 * e;
        *reinterpret_cast<float4*>(out + row) =
            make_float4(acc[m * 4 + 0][e], acc[m * 4 + 1][e], acc[m * 4 + 2][e], acc[m * 4 + 3][e]);
      }
    }
  }
  if (proposed_any && rb == 0) {
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
// The product of a one-trial batch in one read of A, for every step kind.
// It follows k_gram's column-resident form; the exact step is the two
// cublasSgemv calls of powerMethod_by_exact.cuh:46-62, and a sampled step is
// previousAMM_kernel1 / proposedAMM_kernel1 with its weights
// (core_previousAMM.cuh:4-175, core_proposedAMM.cuh:17-213, common.h:42-53).
// Workgroup (problem p, chunk ch) walks the union slots
// [U ch / G, U (ch + 1) / G) of its problem (G = gram_chunks; every column
// on an exact step).  It holds each column in registers and forms
// c_j = a_j^T dx (FP64 sum of exact FP32 products) with dx in LDS:
// x - x_memo on a sampled step, x on an exact one.  The weight comes from
// the draw count as k_bweight forms it: t_j = 1 * c_j on an exact step,
// t_j = w_j (c_j - alpha_j s) on a sampled one (s = 0 unless proposed), and
// t_j a_j goes into a register partial of the workgroup's rows.  The partial
// goes to P2[ch][trial 0] and, proposed, sum alpha_j t_j to P2s[ch][0], both
// for k_bfinish.  The next column (and its weight's operands) streams in
// while the workgroup reduces the current one.  ANY = false: the host knows
// the step is exact (static schedule), and the instance without the weight
// path keeps fewer registers (7 against 5 workgroups per CU at 4096 BF16
// rows).
// ---------------------------------------------------------------------------
template <typename T, int L, bool ANY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_bgram(
    const BProb* __restrict__ P, const uint32_t* off, uint32_t np) {
  constexpr int V = Elem<T>::VEC;
  constexpr uint32_t kRows = 4u * L * 64u * V;
  __shared__ float xs[kRows];
  __shared__ double part[2][4];
  const uint32_t p = find_problem(off, np, blockIdx.x);
  const BProb& pr = P[p];
  if (!gram_step(pr)) return;
  const uint32_t md = pr.bt[0].mode;
  const bool exact = !ANY || md == kModeExactMemo || md == kModeExactNoMemo;
  const bool prop = ANY && md == kModeAmmProposed;
  const uint64_t ch = blockIdx.x - off[p];
  const uint64_t G = pr.gram_chunks;
  const uint64_t U = exact ? pr.n : pr.tpre[pr.ntiles];
  const uint64_t s1 = U * (ch + 1) / G;
  const uint32_t d = pr.d;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t q = resident_rows<V>(d);  // <= L * 64 * V (host: gram_L)
  const uint32_t rb = w * q;
  const uint32_t re = rb + q < d ? rb + q : d;
  const T* A = reinterpret_cast<const T*>(pr.A);
  const uint32_t* __restrict__ colc = pr.colc;
  // the weight's operands of column j: draw count, q_j of the step's CDF, alpha_j
  const uint64_t* __restrict__ cdf = prop ? pr.cdf1 : pr.cdf0;
  const double tot = exact ? 0.0 : pr.st->total[prop ? 1 : 0];
  const double sk = prop ? pr.bt[0].s : 0.0;
  struct Wops {
    uint32_t mult;
    uint64_t qj;
    float aj;
  };
  auto wops = [&](uint32_t j) {
    Wops o{1u, 1ull, 0.0f};
    if (!exact) {
      o.mult = pr.cnt[j];
      o.qj = cdf[j] - (j ? cdf[j - 1] : 0ull);
      if (prop) o.aj = pr.alpha[j];
    }
    return o;
  };
  ColRegs<T, L> cr;
  uint64_t sl = U * ch / G;
  uint32_t j = 0;
  Wops wo{};
  if (sl < s1) {
    j = exact ? (uint32_t)sl : colc[sl];
    cr.load(A + (uint64_t)j * pr.ld, rb, re, lane);
    wo = wops(j);
  }
  uint32_t jn = (sl + 1 < s1) ? (exact ? (uint32_t)(sl + 1) : colc[sl + 1]) : 0u;
  const bool memo = ANY && is_amm(md);
  for (uint32_t r = threadIdx.x; r < kRows; r += 256)
    xs[r] = r < d ? pr.X[r] - (memo ? pr.XM[r] : 0.0f) : 0.0f;  // dx as k_bnorm stages it
  float yp[L * V];
#pragma unroll
  for (int e = 0; e < L * V; ++e) yp[e] = 0.0f;
  double corr = 0.0;  // proposed: sum alpha_j t_j (thread 0)
  __syncthreads();
  for (int par = 0; sl < s1; ++sl, par ^= 1) {
    u32x4 cur[L];
#pragma unroll
    for (int l = 0; l < L; ++l) cur[l] = cr.raw[l];
    const Wops wc = wo;
    if (sl + 1 < s1) {  // the next column and its weight's operands stream in meanwhile
      cr.load(A + (uint64_t)jn * pr.ld, rb, re, lane);
      wo = wops(jn);
      jn = (sl + 2 < s1) ? (exact ? (uint32_t)(sl + 2) : colc[sl + 2]) : 0u;
    }
    double c = 0.0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      float v[V];
      ColRegs<T, L>::unpack(cur[l], v);
      const float* xr = xs + rb + lane * V + l * ColRegs<T, L>::step;
#pragma unroll
      for (int k = 0; k < V; k += 4) {
        const float4 xv = *reinterpret_cast<const float4*>(xr + k);
        c += (double)v[k] * (double)xv.x;
        c += (double)v[k + 1] * (double)xv.y;
        c += (double)v[k + 2] * (double)xv.z;
        c += (double)v[k + 3] * (double)xv.w;
      }
    }
    c = wave_sum(c);
    if (lane == 0) part[par][w] = c;
    __syncthreads();
    const double cj = (part[par][0] + part[par][1]) + (part[par][2] + part[par][3]);
    float tj;
    if (exact) {
      tj = (float)cj;
    } else {
      // weight mult total / (c q_j) as k_bweight forms it, then t = w (c_j - alpha_j s)
      const double den = (double)pr.c * (double)wc.qj;
      const float wj = wc.mult && den > 0.0 ? (float)((double)wc.mult * tot / den) : 0.0f;
      const double ajs = sk != 0.0 ? (double)wc.aj : 0.0;
      tj = (float)((double)wj * (cj - ajs * sk));
      if (prop) corr += (double)wc.aj * (double)tj;
    }
    resident_axpy<T, L>(yp, tj, cur);
  }
  resident_store<T, L>(yp, pr.P2 + ch * kBT * pr.p2pitch, rb, q, lane, pr.dstage);
  if (prop && threadIdx.x == 0) pr.P2s[ch * kBT] = corr;
}

// ---------------------------------------------------------------------------
// Finish the products (dotProduct_and_generate_x, common.h:89-154, and the
// proposed rank-2 part, core_proposedAMM.cuh:198-212): y_k = base_k +
// sum_ch P2[ch][k] with base 0 (exact), y_memo (previous) or y_memo + s v + q
// - sum_ch P2s[ch][k] (proposed); block partials of y^T x, y^T y, y^T v*;
// exact steps with memo snapshot [x, y] -> [x_memo, y_memo].
// Workgroups stride over the problem's 64-row slabs (a workgroup's start is
// a chain of dependent loads that one slab would not amortise); block
// partials stay per slab.  Thread (k, 4 rows).  With G > 1, G wave groups
// of 256 threads split the chunk partials into G contiguous ranges (more
// loads in flight where a problem has many chunks); group 0 adds the ranges
// in group order through LDS and finishes.
// ---------------------------------------------------------------------------
template <uint32_t G>
__global__ __launch_bounds__(256 * G) void k_bfinish(const BProb* __restrict__ P, const uint32_t* off, BRun r) {
  __shared__ double red[G > 1 ? G - 1 : 1][256][5];
  const uint32_t p = find_problem(off, r.np, blockIdx.x);
  const BProb& pr = P[p];
  const uint32_t local = blockIdx.x - off[p];
  const uint32_t nblk = off[p + 1] - off[p];
  const uint32_t grp = threadIdx.x >> 8, lt = threadIdx.x & 255;
  // lane group lt >> 4 takes trial lt >> 4 of a slab, or, with one trial,
  // trial 0 of the pass's slab lt >> 4 (16 slabs per pass)
  const bool one = r.ntr == 1;
  const uint32_t k = one ? 0u : lt >> 4;
  const uint32_t sub = one ? lt >> 4 : 0u, spw = one ? 16u : 1u;
  const bool live = k < r.ntr;
  const uint32_t md = live ? pr.bt[k].mode : kModeExactNoMemo;
  const bool memo_base = is_amm(md);
  const bool prop = md == kModeAmmProposed;
  // chunk partials to add: k_bgram's chunks for the problems it takes, else
  // the second contraction's active chunks
  const uint32_t active = r.gram && gram_step(pr) ? pr.gram_chunks : union_plan(pr, pr.tpre[pr.ntiles]).active;
  const uint32_t c0 = (uint32_t)(((uint64_t)active * grp) / G);
  const uint32_t c1 = (uint32_t)(((uint64_t)active * (grp + 1)) / G);
  const float* __restrict__ P2 = pr.P2;
  const uint64_t pitch = pr.p2pitch;
  const uint32_t d = pr.d, nvblk = pr.nvblk;
  static_assert(kBVecRows == 64, "one 64-row slab per item");
  // the proposed correction sum_ch P2s[ch][k] is the same for every slab
  double corr = 0.0;
  if (live && prop) {
    for (uint32_t ch = c0; ch < c1; ++ch) corr += pr.P2s[(uint64_t)ch * kBT + k];
  }
  double sx = 0.0, qx = 0.0;
  if (prop) {
    sx = pr.bt[k].s;
    qx = pr.bt[k].q;
  }
  for (uint32_t b0 = local * spw; b0 < nvblk; b0 += nblk * spw) {
    const uint32_t blk = b0 + sub;
    const bool in = blk < nvblk;  // uniform over the lane group
    const uint32_t row = blk * kBVecRows + (lt & 15) * 4;
    const uint64_t o = (uint64_t)k * pr.dstage + row;
    // loads that do not depend on the partials first
    float4 ym4 = make_float4(0.f, 0.f, 0.f, 0.f), x4 = make_float4(0.f, 0.f, 0.f, 0.f);
    if (grp == 0 && live && in) {
      if (memo_base) ym4 = *reinterpret_cast<const float4*>(pr.YM + o);
      x4 = *reinterpret_cast<const float4*>(pr.X + o);
    }
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (live && in) {
      // batches of 4 chunk partials, loads first (a loop with a bound per
      // element gets one load in flight at a time); sums in chunk order
      uint32_t ch = c0;
      for (; ch + 4 <= c1; ch += 4) {
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          q[u] = *reinterpret_cast<const float4*>(P2 + ((uint64_t)(ch + u) * kBT + k) * pitch + row);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          a[0] += q[u].x;
          a[1] += q[u].y;
          a[2] += q[u].z;
          a[3] += q[u].w;
        }
      }
      for (; ch < c1; ++ch) {
        const float4 q = *reinterpret_cast<const float4*>(P2 + ((uint64_t)ch * kBT + k) * pitch + row);
        a[0] += q.x;
        a[1] += q.y;
        a[2] += q.z;
        a[3] += q.w;
      }
    }
    double cr = corr;
    if constexpr (G > 1) {
      if (grp > 0) {
        double* sp = red[grp - 1][lt];
        sp[0] = a[0];
        sp[1] = a[1];
        sp[2] = a[2];
        sp[3] = a[3];
        sp[4] = corr;
      }
      __syncthreads();
      if (grp == 0) {
        for (uint32_t g = 0; g + 1 < G; ++g) {
          const double* sp = red[g][lt];
          a[0] += sp[0];
          a[1] += sp[1];
          a[2] += sp[2];
          a[3] += sp[3];
          cr += sp[4];
        }
      }
      __syncthreads();  // red is rewritten by the next slab
      if (grp > 0) continue;
    }
    double yx = 0.0, yy = 0.0, yw = 0.0;
    if (!in) {
      // past the problem's last slab (one trial): nothing to write
    } else if (!live) {  // unused trial slot: zeros, no reads
      *reinterpret_cast<float4*>(pr.Y + o) = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      const float ymv[4] = {ym4.x, ym4.y, ym4.z, ym4.w};
      const float xv[4] = {x4.x, x4.y, x4.z, x4.w};
      float yv[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double acc = memo_base ? (double)ymv[e] : 0.0;
        if (prop && row + e < d) acc += sx * (double)pr.v[row + e] + qx - cr;
        acc += a[e];
        yv[e] = row + e < d ? (float)acc : 0.0f;
        const double yd = yv[e];
        yx += yd * (double)xv[e];
        yy += yd * yd;
        if (pr.eigvec && row + e < d) yw += yd * pr.eigvec[row + e];
      }
      *reinterpret_cast<float4*>(pr.Y + o) = make_float4(yv[0], yv[1], yv[2], yv[3]);
      if (md == kModeExactMemo) {
        *reinterpret_cast<float4*>(pr.XM + o) = x4;
        *reinterpret_cast<float4*>(pr.YM + o) = make_float4(yv[0], yv[1], yv[2], yv[3]);
      }
    }
    yx = sum16(yx);
    yy = sum16(yy);
    yw = sum16(yw);
    if ((lt & 15) == 0 && in) {
      double* b = pr.BP + ((uint64_t)blk * kBT + k) * 3;
      b[0] = yx;
      b[1] = yy;
      b[2] = yw;
    }
  }
}

// ---------------------------------------------------------------------------
// sigma_max of every (problem, trial) after a run: set every trial to an
// exact step (k_bset_mode) so that k_bnorm stages dx = x and the dense union
// gives Tc[j][k] = a_j^T x_k; then sigma_k = ||Tc[:, k]|| (FP64, fixed
// order), and for trial 0 u = x_0, v = Tc[:, 0] / sigma_0 (k_bsigma, one
// workgroup per problem; thread (k, sub) sums columns sub, sub + 16, ...).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_bset_mode(const BProb* __restrict__ P, uint32_t mode) {
  if (threadIdx.x < kBT) P[blockIdx.x].bt[threadIdx.x].mode = mode;
}

__global__ __launch_bounds__(256) void k_bsigma(const BProb* __restrict__ P, BRun r,
                                                double* __restrict__ sigma,
                                                float* const* __restrict__ u_out,
                                                float* const* __restrict__ v_out) {
  __shared__ double s0;
  const BProb& pr = P[blockIdx.x];
  const uint32_t k = threadIdx.x >> 4, sub = threadIdx.x & 15;
  double acc = 0.0;
#pragma unroll 8
  for (uint32_t j = sub; j < pr.n; j += 16) {  // eight 1 KB rows of Tc in flight
    const double t = pr.Tc[(uint64_t)j * kBT + k];
    acc += t * t;
  }
  acc = sum16(acc);
  if (sub == 0 && k < r.ntr) {
    const double sg = sqrt(acc);
    sigma[(uint64_t)blockIdx.x * r.ntr + k] = sg;
    if (k == 0) s0 = sg;
  }
  __syncthreads();
  float* u = u_out ? u_out[blockIdx.x] : nullptr;
  float* v = v_out ? v_out[blockIdx.x] : nullptr;
  const double sg = s0;
  if (v)
    for (uint32_t j = threadIdx.x; j < pr.n; j += 256)
      v[j] = sg > 0.0 ? (float)((double)pr.Tc[(uint64_t)j * kBT] / sg) : 0.0f;
  if (u)
    for (uint32_t i = threadIdx.x; i < pr.d; i += 256) u[i] = pr.X[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Grouped setup (the single-matrix statistics and CDF passes of every
// problem in one launch each; same device functions, same bits).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bcolstats(const BSetup* __restrict__ S, const uint32_t* off,
                                                   uint32_t np, uint32_t skip_fusable) {
  const uint32_t p = find_problem(off, np, blockIdx.x);
  const BSetup& q = S[p];
  if (skip_fusable && q.fusable) return;
  const uint64_t b = blockIdx.x - off[p];
  if (short_column_dt(q.dtype, q.d)) {  // a wave per column, four per workgroup (as k_colstats)
    const uint64_t j = b * 4 + (threadIdx.x >> 6);
    if (j >= q.n) return;
    if (q.dtype == 1)
      colstats_wave<__hip_bfloat16>(reinterpret_cast<const __hip_bfloat16*>(q.A) + j * q.ld, q.d, j, q.alpha,
                                    q.nrm2, q.cnrm2);
    else
      colstats_wave<float>(reinterpret_cast<const float*>(q.A) + j * q.ld, q.d, j, q.alpha, q.nrm2, q.cnrm2);
    return;
  }
  if (q.dtype == 1)
    colstats_column<__hip_bfloat16>(reinterpret_cast<const __hip_bfloat16*>(q.A) + b * q.ld, q.d, b, q.alpha,
                                    q.nrm2, q.cnrm2);
  else
    colstats_column<float>(reinterpret_cast<const float*>(q.A) + b * q.ld, q.d, b, q.alpha, q.nrm2, q.cnrm2);
}

__global__ __launch_bounds__(256) void k_bcdf_max(const BSetup* __restrict__ S, const uint32_t* off,
                                                  uint32_t np, int which) {
  const uint32_t p = find_problem(off, np, blockIdx.x);
  const BSetup& q = S[p];
  cdf_max_tile(which ? q.cnrm2 : q.nrm2, q.n, q.tile_max, blockIdx.x - off[p]);
}

__global__ __launch_bounds__(256) void k_bcdf_scan(const BSetup* __restrict__ S, const uint32_t* off,
                                                   uint32_t np, int which) {
  const uint32_t p = find_problem(off, np, blockIdx.x);
  const BSetup& q = S[p];
  cdf_scan_tile(which ? q.cnrm2 : q.nrm2, q.n, q.cdf[which], q.st, which ? q.alpha : nullptr, q.tile_max,
                q.ntiles, q.tile_sum, q.tile_lp, q.tile_a2, blockIdx.x - off[p]);
}

__global__ __launch_bounds__(256) void k_bcdf_offsets(const BSetup* __restrict__ S, const uint32_t* off,
                                                      uint32_t np, int which) {
  const uint32_t p = find_problem(off, np, blockIdx.x);
  const BSetup& q = S[p];
  cdf_offsets_tile(q.cdf[which], q.n, q.tile_sum, q.tile_lp, q.tile_a2, q.st, which, which, blockIdx.x - off[p],
                   q.ntiles);
}

// v = z - ||alpha||^2 (generate_v, core_proposedAMM.cuh:4-15) for the
// problems whose z came from the statistics pass; 256 rows per block.
__global__ __launch_bounds__(256) void k_bzv(const BSetup* __restrict__ S, const uint32_t* off, uint32_t np) {
  const uint32_t p = find_problem(off, np, blockIdx.x);
  const BSetup& q = S[p];
  if (!q.fusable) return;
  const uint64_t r = (uint64_t)(blockIdx.x - off[p]) * 256 + threadIdx.x;
  if (r >= q.dpad) return;
  q.v[r] = r < q.d ? (float)((double)q.z[r] - q.st->alpha2) : 0.0f;
}

hipError_t launch_bcolstats(const BSetup* S, BGroup g, uint32_t np, uint32_t skip_fusable, hipStream_t s) {
  if (!g.total) return hipSuccess;
  k_bcolstats<<<g.total, 256, 0, s>>>(S, g.off, np, skip_fusable);
  return hipGetLastError();
}

hipError_t launch_bcdf(const BSetup* S, BGroup g, uint32_t np, int which, hipStream_t s) {
  if (!g.total) return hipSuccess;
  k_bcdf_max<<<g.total, 256, 0, s>>>(S, g.off, np, which);
  k_bcdf_scan<<<g.total, 256, 0, s>>>(S, g.off, np, which);
  k_bcdf_offsets<<<g.total, 256, 0, s>>>(S, g.off, np, which);
  return hipGetLastError();
}

hipError_t launch_bzv(const BSetup* S, BGroup g, uint32_t np, hipStream_t s) {
  if (!g.total) return hipSuccess;
  k_bzv<<<g.total, 256, 0, s>>>(S, g.off, np);
  return hipGetLastError();
}

hipError_t launch_bbegin(const BProb* P, BGroup g, BRun r, hipStream_t s) {
  if (g.total) k_bbegin<<<g.total, 256, 0, s>>>(P, g.off, r);
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
  if (g.total) k_bsample<<<g.total, kSampleThreads, 0, s>>>(P, g.off, r, it);
  return hipGetLastError();
}

hipError_t launch_bprep(const BProb* P, BGroup g, BRun r, hipStream_t s) {
  if (g.total) k_bprep<<<g.total, 256, 0, s>>>(P, g.off, r);
  return hipGetLastError();
}

hipError_t launch_bscan(const BProb* P, uint32_t np, hipStream_t s) {
  k_bscan<<<np, 1024, 0, s>>>(P);
  return hipGetLastError();
}

hipError_t launch_bcompact(const BProb* P, BGroup g, uint32_t np, hipStream_t s) {
  if (g.total) k_bcompact<<<g.total, 256, 0, s>>>(P, g.off, np);
  return hipGetLastError();
}

hipError_t launch_bunion(const BProb* P, BGroup g, BRun r, bool fused, hipStream_t s) {
  if (fused) {
    k_bunion<<<r.np, 1024, 0, s>>>(P, r);
    return hipGetLastError();
  }
  hipError_t e = launch_bprep(P, g, r, s);
  if (e == hipSuccess) e = launch_bscan(P, r.np, s);
  if (e == hipSuccess) e = launch_bcompact(P, g, r.np, s);
  return e;
}

hipError_t launch_bweight(const BProb* P, BGroup g, BRun r, uint32_t proposed_any, uint32_t out_w, hipStream_t s,
                          bool skip_gram) {
  if (g.total) k_bweight<<<g.total, 256, 0, s>>>(P, g.off, r, proposed_any, out_w, skip_gram ? 1u : 0u);
  return hipGetLastError();
}

template <bool NT>
static void bdot_launch(int dtype, const BProb* P, BGroup g, uint32_t np, uint32_t skip, hipStream_t s) {
  if (dtype == 1)
    k_bdot<__hip_bfloat16, NT><<<g.total, 256, 0, s>>>(P, g.off, np, skip);
  else
    k_bdot<float, NT><<<g.total, 256, 0, s>>>(P, g.off, np, skip);
}

hipError_t launch_bdot(int dtype, bool nt, const BProb* P, BGroup g, uint32_t np, hipStream_t s,
                       bool skip_gram) {
  if (!g.total) return hipSuccess;
  if (nt)
    bdot_launch<true>(dtype, P, g, np, skip_gram ? 1u : 0u, s);
  else
    bdot_launch<false>(dtype, P, g, np, skip_gram ? 1u : 0u, s);
  return hipGetLastError();
}

template <typename T, bool ANY>
static void bgram_launch(int L, const BProb* P, BGroup g, uint32_t np, hipStream_t s) {
  switch (L) {
    case 1: k_bgram<T, 1, ANY><<<g.total, 256, 0, s>>>(P, g.off, np); break;
    case 2: k_bgram<T, 2, ANY><<<g.total, 256, 0, s>>>(P, g.off, np); break;
    case 4: k_bgram<T, 4, ANY><<<g.total, 256, 0, s>>>(P, g.off, np); break;
    default: k_bgram<T, 8, ANY><<<g.total, 256, 0, s>>>(P, g.off, np); break;
  }
}

hipError_t launch_bgram(int dtype, int L, bool exact_step, const BProb* P, BGroup g, uint32_t np, hipStream_t s) {
  if (!g.total) return hipSuccess;
  if (L != 1 && L != 2 && L != 4 && L != 8) return hipErrorInvalidValue;
  if (dtype == 1) {
    if (exact_step) bgram_launch<__hip_bfloat16, false>(L, P, g, np, s);
    else bgram_launch<__hip_bfloat16, true>(L, P, g, np, s);
  } else {
    if (exact_step) bgram_launch<float, false>(L, P, g, np, s);
    else bgram_launch<float, true>(L, P, g, np, s);
  }
  return hipGetLastError();
}

template <bool NT>
static void baxpy_launch(int dtype, uint32_t passes, const BProb* P, BGroup g, uint32_t np,
                         uint32_t proposed_any, uint32_t skip, hipStream_t s) {
  if (dtype == 1) {
    if (passes == 1)
      k_baxpy<__hip_bfloat16, 1, NT><<<g.total, 256, 0, s>>>(P, g.off, np, proposed_any, skip);
    else
      k_baxpy<__hip_bfloat16, 2, NT><<<g.total, 256, 0, s>>>(P, g.off, np, proposed_any, skip);
  } else {
    if (passes == 1)
      k_baxpy<float, 1, NT><<<g.total, 256, 0, s>>>(P, g.off, np, proposed_any, skip);
    else
      k_baxpy<float, 2, NT><<<g.total, 256, 0, s>>>(P, g.off, np, proposed_any, skip);
  }
}

hipError_t launch_baxpy(int dtype, uint32_t passes, bool nt, const BProb* P, BGroup g, uint32_t np,
                        uint32_t proposed_any, hipStream_t s, bool skip_gram) {
  if (!g.total) return hipSuccess;
  if (nt)
    baxpy_launch<true>(dtype, passes, P, g, np, proposed_any, skip_gram ? 1u : 0u, s);
  else
    baxpy_launch<false>(dtype, passes, P, g, np, proposed_any, skip_gram ? 1u : 0u, s);
  return hipGetLastError();
}

hipError_t launch_bset_mode(const BProb* P, uint32_t np, uint32_t mode, hipStream_t s) {
  k_bset_mode<<<np, 64, 0, s>>>(P, mode);
  return hipGetLastError();
}

hipError_t launch_bsigma(const BProb* P, BRun r, double* sigma, float* const* u, float* const* v,
                         hipStream_t s) {
  k_bsigma<<<r.np, 256, 0, s>>>(P, r, sigma, u, v);
  return hipGetLastError();
}

hipError_t launch_bfinish(const BProb* P, BGroup g, BRun r, uint32_t groups, hipStream_t s) {
  if (!g.total) return hipSuccess;
  if (groups > 1)
    k_bfinish<4><<<g.total, 1024, 0, s>>>(P, g.off, r);
  else
    k_bfinish<1><<<g.total, 256, 0, s>>>(P, g.off, r);
  return hipGetLastError();
}

}  // namespace amm
