// amm_batch.hpp -- batched AMM power method: many matrices, and up to 16
// trials (sampling seeds) per matrix stacked as one right-hand-side block, so
// that both contractions of a product are skinny GEMMs on the matrix cores.
//
// Per iteration and problem, the 16 trials' samples are merged into one
// union of columns U (tiles of 64 columns compacted in place); then
//   P = A_U^T DX            (contraction 1, MFMA, DX = [x_k - x_memo_k])
//   T = W o (P - alpha s^T) (per-trial weights; proposed mean shift)
//   Y = A_U T               (contraction 2, MFMA, split over column chunks)
// Every launch covers all problems ("grouped" launches: a workgroup finds
// its problem in a workgroup -> problem map that follows the per-problem
// workgroup offsets).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "amm_device.hpp"

namespace amm {

constexpr uint32_t kBT = 16;      // trials per right-hand-side block
constexpr uint32_t kBTile = 64;   // columns per compaction tile
constexpr uint32_t kBVecRows = 64;  // rows per block of the vector kernels
constexpr uint32_t kBStageAlign = 256;  // staged vectors are padded to this many rows
constexpr uint32_t kBSampleSlots = 2048;  // draws (slots) per pass of a sampler workgroup
constexpr uint32_t kDenseCols = 32768;    // columns whose 16-bit counts the sampler holds in LDS (64 KB)
constexpr uint32_t kMaxBatchC = 65535;    // draws per trial and iteration (the counts are 16-bit)
constexpr uint32_t kUnionMaxTiles = 1024;  // tiles per problem of the one-workgroup union pass

// Per (problem, trial) scalar state.
struct alignas(16) BTrial {
  double scale;         // 1 / ||y|| of the latest product
  double cos_prev;      // y^T x / ||y|| of the latest iteration
  double s, q;          // proposed: 1^T dx, z^T dx for the current product
  uint64_t last_exact;  // last exact iteration (proposed gate)
  uint32_t mode;        // StepMode of the current iteration
  uint32_t pad;
};

// Device-resident problem descriptor (all pointers are device addresses).
struct alignas(16) BProb {
  const void* A;        // d x n column-major, leading dimension ld
  uint64_t ld;
  uint64_t seed;        // Philox key
  uint32_t d, n;
  uint32_t dstage;      // rows of each staged vector (multiple of 256, >= d)
  uint32_t c;           // samples per trial and iteration
  int32_t dtype;        // 0 FP32, 1 BF16
  uint32_t ntiles;      // ceil(n / 64)
  uint32_t kslices_max; // contraction-1 row slices the partial buffer holds (<= 16)
  uint32_t dot_blocks;  // contraction-1 workgroups to aim for (sets the slices per |U|)
  uint32_t passes;      // contraction-2 row block = 256 passes rows
  uint32_t rblocks;     // contraction-2 row blocks
  uint32_t chunks;      // contraction-2 column chunks of the union (at most)
  uint32_t p2pitch;     // rows of one P2 partial (rblocks * rows per block)
  uint32_t nvblk;       // vector-kernel blocks = dstage / 64
  uint32_t gram_chunks; // one-read products (k_bgram): chunks of union slots, 0 = off (one-trial batches)
  float* X;             // [16][dstage] iterate
  float* Y;             // [16][dstage] product
  float* XM;            // [16][dstage] x_memo
  float* YM;            // [16][dstage] y_memo
  float* DX;            // [dstage / 16][16][16] FP32 staging of x - x_memo (FP32 A), dx_offset<16>
  uint16_t* DXh;        // [dstage / 32][16][32] BF16 high part (BF16 A), dx_offset<32>
  uint16_t* DXl;        // [dstage / 32][16][32] BF16 low part
  uint16_t* cnt;        // [16][ntiles*64] draw counts, a row per trial (rewritten whole when it samples)
  uint32_t* list;       // [ntiles*64] compacted column indices, per tile
  uint32_t* tcount;     // [ntiles] columns present per tile
  uint32_t* tpre;       // [ntiles + 1] prefix of tcount; tpre[ntiles] = |U|
  uint32_t* colc;       // [n] union columns in compacted order
  float* Tc;            // [n][16] t = W (A_U^T dx - alpha s) of the current product
  float* P1;            // [kslices_max][ntiles*64][16], compacted slots
  float* P2;            // [chunks][16][p2pitch]
  double* P2s;          // [chunks][16] proposed: sum alpha_j t_j per chunk
  double* BP;           // [nvblk][16][3] block partials of y^T x, y^T y, y^T v*
  double* BQ;           // [nvblk][16][2] block partials of 1^T dx, z^T dx
  DevState* st;         // column statistics: totals, last positive column, ||alpha||^2
  const uint64_t* cdf0; // plain CDF
  const uint64_t* cdf1; // centred CDF
  const float* alpha;   // column means
  const float* z;       // A alpha
  const float* v;       // z - ||alpha||^2
  const double* eigvec; // reference eigenvector (may be null)
  BTrial* bt;           // [16]
  double* eig;          // [16][step]
  double* err;          // [16][step]
  uint32_t* modes;      // [16][step] step kind used at each iteration
  uint32_t* pad_ptr;
};

// Staged dx (the operand of contraction 1) is stored trial-interleaved by
// groups of G rows: rows [G g, G g + G) of trial k are the G consecutive
// elements at (g * 16 + k) * G, so one K step of a wave (G rows of all 16
// trials: 1 KB) is one contiguous read.  G = 32 for the BF16 parts, 16 for
// FP32.
template <uint32_t G>
__host__ __device__ inline uint64_t dx_offset(uint32_t k, uint32_t row) {
  return ((uint64_t)(row / G) * kBT + k) * G + row % G;
}

// Scalars shared by every launch of a run.
struct BRun {
  uint32_t np;          // problems
  uint32_t gram;        // a run whose products may go through k_bgram (k_bfinish sums its partials)
  uint32_t ntr;         // trials per problem (<= 16)
  uint32_t trial0;      // Philox counter word 2 of trial 0
  uint32_t kernel;      // AMM_KERNEL_* (1 previous / 2 proposed drive adaptive steps)
  uint32_t step;
  uint32_t adaptive;
  uint64_t period;
  double thr;
  const uint32_t* modes_static;  // [step] static schedule
  int64_t* ts;          // [step] device time stamps
  int64_t* t_start;     // wall clock at the run's first kernel
};

enum BTailOp : uint32_t { kBTailFinish = 0, kBTailScale = 1 };

// Launch plan of one product from the size of its union (every kernel
// derives it identically from |U| = tpre[ntiles]):
//   contraction 1: ceil(|U| / 64) virtual tiles x kslices row slices of
//     kslice_rows rows, kslices ~ dot_blocks / tiles (at most kslices_max);
//   contraction 2: the union in `active` chunks of cpc slots (multiple of
//     64), at most `chunks` of them.
constexpr uint32_t kMinChunk = 256;

struct UPlan {
  uint32_t U, vtiles, kslices, kslice_rows, cpc, active;
};

__host__ __device__ inline UPlan union_plan(const BProb& p, uint32_t U) {
  UPlan q;
  q.U = U;
  q.vtiles = (U + 63) / 64;
  uint32_t ks = q.vtiles ? (p.dot_blocks + q.vtiles - 1) / q.vtiles : 1;
  ks = ks < 1 ? 1 : (ks > p.kslices_max ? p.kslices_max : ks);
  const uint32_t rows = (p.dstage + ks - 1) / ks;
  q.kslice_rows = (rows + 255) / 256 * 256;
  q.kslices = (p.dstage + q.kslice_rows - 1) / q.kslice_rows;
  // at least kMinChunk columns per chunk: a chunk's partial (16 trials x
  // 4 bytes per row) stays within 1/8 of the bytes it reads per row
  const uint32_t per = (U + p.chunks - 1) / p.chunks;
  q.cpc = per < kMinChunk ? kMinChunk : (per + 63) / 64 * 64;
  q.active = (U + q.cpc - 1) / q.cpc;
  return q;
}

// grouped launch of one kernel: per-problem item offsets [np + 1]
struct BGroup {
  const uint32_t* off;
  uint32_t total;
};

// Per-problem setup descriptor of the grouped statistics / CDF launches.
struct BSetup {
  const void* A;
  uint64_t ld;
  uint32_t d, n;
  int32_t dtype;
  uint32_t fusable;  // the statistics pass can form z = A alpha (register-held columns)
  float* alpha;
  double* nrm2;
  double* cnrm2;
  uint64_t* cdf[2];
  DevState* st;
  float* z;
  float* v;
  uint64_t dpad;
  uint64_t* tile_max;  // this problem's CDF tile scratch (ntiles each)
  uint64_t* tile_sum;
  uint64_t* tile_lp;
  double* tile_a2;
  uint32_t ntiles;
  uint32_t pad;
};

// Grouped setup: column statistics of every problem (one workgroup per
// column; problems with fusable set are skipped when skip_fusable), the
// CDF of weights `which` (0: nrm2, 1: cnrm2 with ||alpha||^2) of every
// problem, and v = z - ||alpha||^2 of the fusable problems.
hipError_t launch_bcolstats(const BSetup* S, BGroup g, uint32_t np, uint32_t skip_fusable, hipStream_t s);
hipError_t launch_bcdf(const BSetup* S, BGroup g, uint32_t np, int which, hipStream_t s);
hipError_t launch_bzv(const BSetup* S, BGroup g, uint32_t np, hipStream_t s);

hipError_t launch_bbegin(const BProb* P, BGroup g, BRun r, hipStream_t s);
hipError_t launch_btail(const BProb* P, BRun r, uint32_t op, uint32_t it, hipStream_t s);
// items: nvblk per problem, or ceil(nvblk / 16) with one trial (16 slabs each)
hipError_t launch_bnorm(const BProb* P, BGroup g, BRun r, hipStream_t s);
// items: ntr per problem (one workgroup per trial)
hipError_t launch_bsample(const BProb* P, BGroup g, BRun r, uint32_t it, hipStream_t s);
hipError_t launch_bprep(const BProb* P, BGroup g, BRun r, hipStream_t s);
hipError_t launch_bscan(const BProb* P, uint32_t np, hipStream_t s);
// items: ceil(ntiles / 4) per problem (the k_bprep group)
hipError_t launch_bcompact(const BProb* P, BGroup g, uint32_t np, hipStream_t s);
// the union of a product: k_bprep, k_bscan, k_bcompact (g: the k_bprep
// group), or with fused k_bunion, one workgroup per problem (every problem
// has at most kUnionMaxTiles tiles)
hipError_t launch_bunion(const BProb* P, BGroup g, BRun r, bool fused, hipStream_t s);
// items: any count per problem (the blocks stride over the union slots);
// t of every union slot, or (out_w) the bare weights
// skip_gram (runs of a batch with k_bgram): the problems k_bgram takes are
// left alone by the weight pass and both contractions
hipError_t launch_bweight(const BProb* P, BGroup g, BRun r, uint32_t proposed_any, uint32_t out_w, hipStream_t s,
                          bool skip_gram = false);
// dtype-uniform groups: all problems of a group share the element type
// nt: A read with non-temporal loads (streams that no later pass reuses)
hipError_t launch_bdot(int dtype, bool nt, const BProb* P, BGroup g, uint32_t np, hipStream_t s,
                       bool skip_gram = false);
// passes: 1 or 2 (a group shares it)
hipError_t launch_baxpy(int dtype, uint32_t passes, bool nt, const BProb* P, BGroup g, uint32_t np,
                        uint32_t proposed_any, hipStream_t s, bool skip_gram = false);
// The products of a one-trial batch in one read of A (k_bgram): items
// gram_chunks per problem of element type dtype; L: 16-byte loads per lane
// that hold the group's longest column (1, 2, 4 or 8); exact_step: every
// trial takes an exact step this iteration (known on the host), else the
// step kind is read on the device
hipError_t launch_bgram(int dtype, int L, bool exact_step, const BProb* P, BGroup g, uint32_t np, hipStream_t s);
// 16-byte loads per lane of k_bgram's register-held column of d rows: 0
// beyond 8 (16384 BF16 / 8192 FP32 rows), where its registers would spill
__host__ __device__ inline int gram_L(int dtype, uint32_t d) {
  const uint32_t step = 64u * (dtype == 1 ? 8u : 4u);
  const uint32_t need = ((d + 3) / 4 + step - 1) / step;
  int L = 1;
  while ((uint32_t)L < need) L *= 2;
  return L <= 8 ? L : 0;
}
// groups: 1 (256 threads per workgroup) or 4 (1024: four ranges of chunks)
hipError_t launch_bfinish(const BProb* P, BGroup g, BRun r, uint32_t groups, hipStream_t s);
hipError_t launch_bset_mode(const BProb* P, uint32_t np, uint32_t mode, hipStream_t s);
// sigma: device [np][ntr]; u, v: device arrays of np device pointers (or null)
hipError_t launch_bsigma(const BProb* P, BRun r, double* sigma, float* const* u, float* const* v,
                         hipStream_t s);

}  // namespace amm
