// amm_launch.hpp -- host-side launchers of the AMM kernels (amm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amm {

struct DevState;

// wave-steps (64 lanes x 16 bytes) per trip of the sampled dot pass; the
// plan splits rows so that one wave covers one trip
constexpr int kListDotSteps = 4;

// The tail of an iteration carried by the next sampled dot pass (static
// schedules, any step followed by a sampled step): every wave
// reduces the finish partials BP to 1/||y|| itself (finish_sums_wave, the
// k_tail order) and forms x = y / ||y|| for its rows on the fly; the first
// workgroup of each row split writes x (and, with z, its part r of s = 1^T dx
// and q = z^T dx) and the first workgroup writes the iteration's scalars --
// what k_tail(kTailFinish, norm) writes, without its launch.
struct GatherTail {
  const double* BP;  // finish block partials of iteration it (y^T x, y^T y, y^T v*)
  uint32_t nb;
  DevState* st;
  const float* y;    // the product of iteration it
  float* x;          // written: y / ||y||
  const float* z;    // proposed: z (s, q parts per row split), else null
  double* eig;
  double* err;
  int64_t* ts;       // time stamp of iteration it (null: none)
  uint32_t it;
  uint32_t exact;    // iteration it was an exact step (last_exact = it)
  uint32_t x0;       // the raw x0 instead: BP holds its per-block sums of squares
                     // (1 double per block); only x, s, q and 1/||x0|| are written
};

struct GatherDotArgs {
  const void* A;
  uint64_t ld;
  uint32_t d;
  const uint32_t* idx;        // null: dense (column e for entry e)
  const uint32_t* seg_count;  // per-segment unique counts (list mode)
  uint32_t cap;               // list capacity (c) or n
  const float* x;
  const float* xm;            // null: no memo
  uint32_t rows_per_split;
  uint32_t R;
  float* P1;                  // [R][cap]
  const uint32_t* mode;       // null: always run
  uint32_t it;
  uint32_t gate_mask;
  bool nt;                    // list mode: non-temporal loads of A
  const GatherTail* tail;     // null: x is given; else the previous iteration's tail (list mode)
};

struct GatherAxpyArgs {
  const void* A;
  uint64_t ld;
  uint32_t d;
  uint64_t dpad;
  const uint32_t* idx;
  const float* wgt;
  const uint32_t* seg_count;
  uint32_t cap;
  uint32_t cs;                // entries per chunk (power of two in list mode)
  uint32_t S;                 // number of chunks
  const float* P1;
  uint32_t R;
  const float* tdirect;       // t_e = tdirect[col] when non-null
  const float* alpha;         // proposed mean shift when non-null
  const DevState* st;
  float* P2;                  // [S][dpad]
  double* P2s;                // [S]
  const uint32_t* mode;
  uint32_t it;
  uint32_t gate_mask;
  bool nt;                    // list mode: non-temporal loads of A
};

struct FinishArgs {
  const float* P2;
  uint32_t S_exact;  // chunks of the exact product
  uint32_t S_amm;    // chunks of the sampled product
  uint64_t dpad;
  uint32_t d;
  const float* x;
  float* y;
  float* xm;
  float* ym;
  const float* v;
  const double* P2s;
  const double* eigvec;
  double* BP;
  const DevState* st;
  const uint32_t* mode;
  uint32_t it;
  uint32_t gate_mask;  // nonzero: run only when bit mode[it] is set
};

// Second contraction of a sampled product with the finish fused in
// (k_axpy_finish): y over every row and the block partials of y^T x, y^T y,
// y^T v* in the layout of k_finish (finish_blocks(dpad) blocks).  List mode
// only; worthwhile where the row blocks fill the device
// (axpy_finish_applies).
struct AxpyFinishArgs {
  const void* A;
  uint64_t ld;
  uint32_t d;
  uint64_t dpad;
  const uint32_t* idx;
  const float* wgt;
  const uint32_t* seg_count;
  uint32_t cap;
  const float* P1;
  uint32_t R;
  const float* alpha;  // proposed mean shift when non-null
  const DevState* st;
  const float* x;
  float* y;
  const float* ym;     // y_memo (the base of every sampled product)
  const float* v;      // z - ||alpha||^2 (proposed)
  const double* eigvec;
  double* BP;
  const uint32_t* mode;
  uint32_t it;
  uint32_t gate_mask;
  bool nt;
};
constexpr uint64_t kAfMinBlocks = 128;  // row blocks of 64 rows below which the chunked form is used
bool axpy_finish_applies(uint64_t dpad);
hipError_t launch_axpy_finish(int dtype, const AxpyFinishArgs& a, hipStream_t s);

// single-block tail of a grid reduction (k_tail)
enum TailOp : uint32_t {
  kTailFinish = 0,
  kTailScale = 1,
  kTailSQ = 2,
  kTailNorm = 3,
  kTailNorm2 = 4,  // sigma = sqrt(sum of one partial per block)
};

struct TailArgs {
  uint32_t op;            // 0 finish, 1 scale of x0, 2 s and q
  const double* BP;
  uint32_t nb;            // blocks whose partials are in BP
  DevState* st;
  int norm;               // also write x = y / ||y|| (and s, q when z is given)
  const float* y;
  float* x;
  const float* xm;
  const float* z;
  uint32_t d;
  uint64_t dpad;
  uint32_t* mode;
  uint32_t it;
  double* eig;
  double* err;
  int64_t* ts;
  int stamp;              // write the iteration's time stamp
  uint32_t adapt_kernel;  // 0: static schedule; 1 previous / 2 proposed: decide next step on device
  uint64_t period;
  double thr;
  uint32_t step;
};

// blocks of one k_finish launch (each writes 3 doubles of block partials)
uint64_t finish_blocks(uint64_t dpad);
uint64_t x0_blocks(uint64_t dpad);         // 1 double per block
uint64_t normalize_blocks(uint64_t dpad);  // 2 doubles per block
uint64_t dots_blocks(uint64_t dpad);       // 2 doubles per block
hipError_t launch_tail(const TailArgs& a, hipStream_t s);
// tiles of the CDF scan (tile scratch: one uint64 sum, uint64 last column,
// double ||alpha||^2 partial per tile)
uint64_t cdf_tiles(uint64_t n);

// Column statistics; where the pass runs as workgroups over runs of columns
// (colstats_max_blocks(dtype, d, n) > 0) it also writes each workgroup's
// largest plain and centred weight to wmax[2][that count] (wmax may be null)
// for launch_cdf, which then skips its own maximum pass.
constexpr uint32_t kColstatsMaxBlocks = 1024;
uint32_t colstats_max_blocks(int dtype, uint32_t d, uint64_t n);
hipError_t launch_colstats(int dtype, const void* A, uint64_t ld, uint32_t d, uint64_t n,
                           float* alpha, double* nrm2, double* cnrm2, DevState* st, uint64_t* wmax,
                           hipStream_t s);
// Column statistics and the z = A alpha partials
// ZP[colstats_z_blocks(dtype, d, n)][dpad] in the same read of A (finish
// with launch_sum_rows).  Needs colstats_fuses_z(dtype, d).
hipError_t launch_colstats_z(int dtype, const void* A, uint64_t ld, uint32_t d, uint64_t n, uint64_t dpad,
                             float* alpha, double* nrm2, double* cnrm2, float* ZP, DevState* st,
                             hipStream_t s);
uint32_t colstats_z_blocks(int dtype, uint32_t d, uint64_t n);
bool colstats_fuses_z(int dtype, uint32_t d);
// The exact product y = A (A^T x) in one read of A: partials
// P2[gram_blocks(dtype, d, n)][dpad] for the finish kernel.  Needs
// gram_fits(dtype, d); skipped unless the gate admits mode[it].
hipError_t launch_gram(int dtype, const void* A, uint64_t ld, uint32_t d, uint64_t n, uint64_t dpad,
                       const float* x, float* P2, const uint32_t* gate_mode, uint32_t it, uint32_t gate_mask,
                       hipStream_t s);
uint32_t gram_blocks(int dtype, uint32_t d, uint64_t n);
bool gram_fits(int dtype, uint32_t d);
// wmax / nwmax: the statistics pass's per-workgroup maxima (null / 0: the
// CDF's own maximum pass)
hipError_t launch_cdf(const double* w, uint64_t n, uint64_t* cdf, DevState* st, int which,
                      const float* alpha, uint64_t* tile_max, uint64_t* tile_sum, uint64_t* tile_lp,
                      double* tile_a2, const uint64_t* wmax, uint32_t nwmax, hipStream_t s);
hipError_t launch_sample(const uint64_t* cdf, uint64_t n, const DevState* st, int which, uint32_t it0, uint32_t nit,
                         uint32_t buf_it0, uint32_t c, uint32_t* idx, float* wgt,
                         uint32_t* seg_count, hipStream_t s);
hipError_t launch_dots(const float* x, const float* xm, const float* z, uint64_t dpad, double* BP,
                       hipStream_t s);
// run prologue with x0: raw x0 into y, its per-block sums of squares into
// BP (x0_blocks(dpad) of them, for k_tail's kTailScale)
hipError_t launch_begin(float* xm, float* ym, uint64_t dpad, DevState* st, uint32_t* mode,
                        uint32_t mode0, float* y, uint32_t d, double* BP, hipStream_t s);
hipError_t launch_gather_dot(int dtype, const GatherDotArgs& a, hipStream_t s);
hipError_t launch_gather_axpy(int dtype, const GatherAxpyArgs& a, hipStream_t s);
hipError_t launch_finish(const FinishArgs& a, hipStream_t s);
// plain streaming read of `bytes` (a multiple of 16 is read) from the
// 16-byte aligned A; sink: one writable word
hipError_t launch_read(const void* A, uint64_t bytes, uint32_t* sink, hipStream_t s);

hipError_t launch_normalize(const float* y, float* x, uint32_t d, uint64_t dpad, const DevState* st,
                            const float* xm, const float* z, double* BP, hipStream_t s);
hipError_t launch_sum_rows(const float* P2, uint32_t S, uint64_t dpad, uint32_t d, float* z,
                           float* v, const DevState* st, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t val, hipStream_t s);
uint64_t reduce_p1_blocks(uint64_t n);  // 1 double per block
hipError_t launch_reduce_p1(const float* P1, uint32_t R, uint64_t n, float* t, double* BP,
                            hipStream_t s);
hipError_t launch_unit(const float* t, uint64_t n, const DevState* st, float* v, hipStream_t s);
hipError_t launch_rank1(int dtype, void* out, uint32_t d, uint64_t n, uint64_t ld, float g,
                        const float* a, const float* b, hipStream_t s);
// count matrices of a strided stack in one launch (count < 65536): matrix k
// at out + k ostride (elements), g_k = gs[k] when gs is given (device), else g.
hipError_t launch_rank1_batched(int dtype, void* out, uint32_t d, uint64_t n, uint64_t ld, uint64_t count,
                                uint64_t ostride, float g, const float* gs, const float* a, uint64_t astride,
                                const float* b, uint64_t bstride, hipStream_t s);
hipError_t launch_generate_symmetric(float* A, uint64_t m, uint64_t ld, int dist, float p1, float p2,
                                    uint64_t seed, uint32_t* delta_bits, hipStream_t s);
hipError_t launch_f32_to_bf16(const float* src, uint16_t* dst, uint64_t count, hipStream_t s);

}  // namespace amm
