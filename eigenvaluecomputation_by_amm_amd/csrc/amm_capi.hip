// amm_capi.hip -- C ABI of the AMM power method (include/amm.h): context,
// buffer ownership, launch planning, the power-iteration schedules and the
// hipGraph that carries a whole run.
//
// Schedules restate powerMethod_by_exact.cuh:4-100,
// powerMethod_by_previous.cuh:100-210, powerMethod_by_proposed.cuh:105-235 and
// powerMethod_by_proposed_v2.cuh:199-377.  Unlike the reference there is no
// host round trip inside a run: samples of every iteration are drawn up
// front (the Philox stream is keyed by iteration), the Rayleigh quotient,
// normalisation and the adaptive variance-reduction decision stay on the
// device, per-iteration time stamps come from the device wall clock, and the
// whole run is one hipGraph launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "amm.h"
#include "amm_device.hpp"
#include "amm_launch.hpp"

namespace amm_host {
std::recursive_mutex& device_mutex();
}

using amm::DevState;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(AMM_E_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

// Host copies run on the object's own stream and then wait for it.  A plain
// hipMemcpy runs on the legacy null stream, which may not take a dependency
// on a blocking stream that another thread of this process is capturing into
// a graph; the failed copy would then abort that capture too.
hipError_t copy_sync(hipStream_t s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}


template <typename T>
int grow(T*& p, size_t& cap, size_t need) {
  if (need <= cap && p) return AMM_OK;
  std::lock_guard<std::recursive_mutex> lk(amm_host::device_mutex());
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  const size_t n = need ? need : 1;
  if (hipMalloc((void**)&p, n * sizeof(T)) != hipSuccess) {
    p = nullptr;
    return fail(AMM_E_NOMEM, "device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed");
  }
  if (hipMemset(p, 0, n * sizeof(T)) != hipSuccess) return fail(AMM_E_HIP, "hipMemset failed");
  cap = n;
  return AMM_OK;
}

uint32_t next_pow2(uint32_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

constexpr uint32_t kTargetBlocks = 1024;  // >> 256 CUs
constexpr uint32_t kMaxChunk = 2048;

struct DotPlan {
  uint32_t R, rows_per_split;
};
struct AxpyPlan {
  uint32_t cs, S;
};

}  // namespace

// error reporting shared with the batched engine (amm_batch_capi.hip)
namespace amm_host {
int fail(int code, const std::string& msg) { return ::fail(code, msg); }

// Held while any context or batch of this process captures a graph and
// while one allocates, frees or clears device memory with the synchronous
// calls (hipMalloc / hipFree / hipMemset), so that those never fall inside
// another thread's capture.
std::recursive_mutex& device_mutex() {
  static std::recursive_mutex m;
  return m;
}
}  // namespace amm_host

struct amm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  const void* A = nullptr;
  int dtype = AMM_DTYPE_F32;
  uint64_t d = 0, n = 0, ld = 0, dpad = 0;
  int V = 4;
  int wall_khz = 100000;

  float* alpha = nullptr;
  double* nrm2 = nullptr;
  double* cnrm2 = nullptr;
  uint64_t* cdf[2] = {nullptr, nullptr};
  float *z = nullptr, *v = nullptr, *x = nullptr, *y = nullptr, *xm = nullptr, *ym = nullptr;
  double* eigvec = nullptr;
  bool has_eigvec = false;
  DevState* st = nullptr;
  double* BP = nullptr;
  uint64_t* tile_max = nullptr;
  // per-workgroup maxima of the weights from the statistics pass ([2][wmax_n],
  // plain then centred), valid while wmax_n > 0; the CDF then skips its
  // maximum pass
  uint64_t* wmax = nullptr;
  uint32_t wmax_n = 0;
  uint64_t* tile_sum = nullptr;
  uint64_t* tile_lp = nullptr;
  double* tile_a2 = nullptr;

  uint32_t* mode = nullptr;
  size_t mode_cap = 0;
  double* eig = nullptr;
  size_t eig_cap = 0;
  double* err = nullptr;
  size_t err_cap = 0;
  int64_t* ts = nullptr;
  size_t ts_cap = 0;
  uint32_t* idx = nullptr;
  size_t idx_cap = 0;
  float* wgt = nullptr;
  size_t wgt_cap = 0;
  uint32_t* segc = nullptr;
  size_t segc_cap = 0;
  float* P1 = nullptr;
  size_t P1_cap = 0;
  float* P2 = nullptr;
  size_t P2_cap = 0;
  double* P2s = nullptr;
  size_t P2s_cap = 0;
  float* tvec = nullptr;  // A^T u of sigma_max
  size_t tvec_cap = 0;
  double* tbp = nullptr;  // block partials of ||A^T u||^2
  size_t tbp_cap = 0;
  float* ZP = nullptr;  // per-workgroup partials of z = A alpha from the statistics pass
  size_t ZP_cap = 0;

  bool stats_valid = false;

  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_valid = false;
  amm_params graph_key{};

  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  uint64_t last_step = 0;
  std::vector<uint32_t> host_modes;
  // host staging of small uploads (kept alive until the copies complete)
  uint64_t h_seed = 0;
  uint32_t h_trial = 0;
  uint32_t h_mode = 0;

  // launch plans
  DotPlan dot_dense{}, dot_list{};
  uint32_t gram_S = 0;  // partials of the one-pass exact product (0: two-pass form)
  AxpyPlan axpy_dense{}, axpy_list{};
  uint32_t list_c = 0;
  // non-temporal loads in the sampled passes: bit 0 dot, bit 1 axpy
  // (environment AMM_LIST_NT overrides).  Off by default: sampled columns are
  // read twice per iteration and heavy ones recur across iterations, and on
  // the 16384 x 65536 workload plain loads were faster for both passes
  // (dot 7.7 vs 10.7 us, axpy 8.1 vs 9.0 us per launch).
  uint32_t list_nt = 0;
};

namespace {

uint32_t rowblocks(const amm_ctx* c) { return (uint32_t)((c->dpad + 256u * c->V - 1) / (256u * c->V)); }

DotPlan plan_dot(const amm_ctx* c, uint64_t entries, int cpw) {
  const uint64_t groups = (entries + 4 * cpw - 1) / (4 * cpw);
  const uint32_t step = 64u * c->V * 2;
  uint64_t R = groups >= kTargetBlocks ? 1 : (kTargetBlocks + groups - 1) / groups;
  const uint64_t maxR = std::max<uint64_t>(1, (c->d + step - 1) / step);
  R = std::min(R, maxR);
  uint64_t rps = (c->d + R - 1) / R;
  rps = (rps + step - 1) / step * step;
  R = (c->d + rps - 1) / rps;
  return DotPlan{(uint32_t)R, (uint32_t)rps};
}

// Sampled dot pass: rows split so that each wave covers one trip of
// kListDotSteps wave-steps, every load of it in flight at once (a sampled
// pass is a few tens of MB: latency, not bandwidth, sets its time).
DotPlan plan_dot_list(const amm_ctx* c) {
  const uint64_t trip = 64u * c->V * amm::kListDotSteps;
  const uint64_t R = std::max<uint64_t>(1, (c->d + trip - 1) / trip);
  return DotPlan{(uint32_t)R, (uint32_t)trip};
}

AxpyPlan plan_axpy(const amm_ctx* c, uint64_t entries, bool pow2) {
  const uint32_t rb = rowblocks(c);
  const uint64_t S_t = std::max<uint64_t>(1, (kTargetBlocks + rb - 1) / rb);
  uint64_t cs = std::max<uint64_t>(1, (entries + S_t - 1) / S_t);
  if (pow2) cs = next_pow2((uint32_t)std::min<uint64_t>(cs, kMaxChunk));
  cs = std::min<uint64_t>(cs, kMaxChunk);
  const uint64_t S = (entries + cs - 1) / cs;
  return AxpyPlan{(uint32_t)cs, (uint32_t)S};
}

int set_device(amm_ctx* c) {
  HIPCHK(hipSetDevice(c->device));
  return AMM_OK;
}

int ensure_plans(amm_ctx* c, uint32_t list_c) {
  c->dot_dense = plan_dot(c, c->n, 4);
  c->axpy_dense = plan_axpy(c, c->n, false);
  c->list_c = list_c;
  if (list_c) {
    c->dot_list = plan_dot_list(c);
    c->axpy_list = plan_axpy(c, list_c, true);
  } else {
    c->dot_list = DotPlan{1, 0};
    c->axpy_list = AxpyPlan{1, 0};
  }
  size_t p1 = (size_t)c->dot_dense.R * c->n;
  if (list_c) p1 = std::max(p1, (size_t)c->dot_list.R * list_c);
  c->gram_S = amm::gram_fits(c->dtype, (uint32_t)c->d) ? amm::gram_blocks(c->dtype, (uint32_t)c->d, c->n) : 0;
  const uint32_t S = std::max({c->axpy_dense.S, c->axpy_list.S, c->gram_S});
  int rc;
  if ((rc = grow(c->P1, c->P1_cap, p1))) return rc;
  if ((rc = grow(c->P2, c->P2_cap, (size_t)S * c->dpad))) return rc;
  if ((rc = grow(c->P2s, c->P2s_cap, (size_t)S))) return rc;
  return AMM_OK;
}

int ensure_step(amm_ctx* c, uint64_t step) {
  int rc;
  if ((rc = grow(c->mode, c->mode_cap, step))) return rc;
  if ((rc = grow(c->eig, c->eig_cap, step))) return rc;
  if ((rc = grow(c->err, c->err_cap, step))) return rc;
  if ((rc = grow(c->ts, c->ts_cap, step))) return rc;
  return AMM_OK;
}

int ensure_lists(amm_ctx* c, uint64_t nit, uint32_t cc) {
  const uint64_t nseg = (cc + amm::kSampleSeg - 1) / amm::kSampleSeg;
  int rc;
  if ((rc = grow(c->idx, c->idx_cap, (size_t)nit * cc))) return rc;
  if ((rc = grow(c->wgt, c->wgt_cap, (size_t)nit * cc))) return rc;
  if ((rc = grow(c->segc, c->segc_cap, (size_t)nit * nseg))) return rc;
  return AMM_OK;
}

// ---- enqueue helpers (all asynchronous on c->stream) ----------------------
int enq_stats(amm_ctx* c) {
  HIPCHK(amm::launch_colstats(c->dtype, c->A, c->ld, (uint32_t)c->d, c->n, c->alpha, c->nrm2,
                              c->cnrm2, c->st, c->wmax, c->stream));
  c->wmax_n = amm::colstats_max_blocks(c->dtype, (uint32_t)c->d, c->n);
  return AMM_OK;
}

// z = A alpha rides in the statistics pass when the columns fit the
// register-held form (d <= 16384 FP32, 32768 BF16) and the partial buffer
// is there; the run then finishes it with enq_z_partials after the
// centred CDF has set ||alpha||^2.
bool z_fusable(const amm_ctx* c) {
  return amm::colstats_fuses_z(c->dtype, (uint32_t)c->d) &&
         c->ZP_cap >= (size_t)amm::colstats_z_blocks(c->dtype, (uint32_t)c->d, c->n) * c->dpad;
}

int enq_stats_z(amm_ctx* c) {
  HIPCHK(amm::launch_colstats_z(c->dtype, c->A, c->ld, (uint32_t)c->d, c->n, c->dpad, c->alpha, c->nrm2,
                                c->cnrm2, c->ZP, c->st, c->stream));
  c->wmax_n = 0;  // this pass leaves no maxima
  return AMM_OK;
}

int enq_z_partials(amm_ctx* c) {
  const uint32_t S = amm::colstats_z_blocks(c->dtype, (uint32_t)c->d, c->n);
  HIPCHK(amm::launch_sum_rows(c->ZP, S, c->dpad, (uint32_t)c->d, c->z, c->v, c->st, c->stream));
  return AMM_OK;
}

int enq_cdf(amm_ctx* c, int which, bool /*unused*/) {
  HIPCHK(amm::launch_cdf(which ? c->cnrm2 : c->nrm2, c->n, c->cdf[which], c->st, which,
                         which ? c->alpha : nullptr, c->tile_max, c->tile_sum, c->tile_lp,
                         c->tile_a2, c->wmax, c->wmax_n, c->stream));
  return AMM_OK;
}

// rows above which the normalisation runs as its own kernel instead of in
// the finish kernel's last block
constexpr uint64_t kFuseNormRows = 65536;

int enq_sample(amm_ctx* c, int which, uint32_t it0, uint32_t nit, uint32_t buf_it0, uint32_t cc) {
  HIPCHK(amm::launch_sample(c->cdf[which], c->n, c->st, which, it0, nit, buf_it0, cc, c->idx, c->wgt,
                            c->segc, c->stream));
  return AMM_OK;
}

// z = A alpha, v = z - ||alpha||^2 (powerMethod_by_proposed.cuh:132-145)
int enq_z(amm_ctx* c) {
  amm::GatherAxpyArgs a{};
  a.A = c->A;
  a.ld = c->ld;
  a.d = (uint32_t)c->d;
  a.dpad = c->dpad;
  a.cap = (uint32_t)c->n;
  a.cs = c->axpy_dense.cs;
  a.S = c->axpy_dense.S;
  a.tdirect = c->alpha;
  a.st = c->st;
  a.P2 = c->P2;
  a.P2s = c->P2s;
  HIPCHK(amm::launch_gather_axpy(c->dtype, a, c->stream));
  HIPCHK(amm::launch_sum_rows(c->P2, a.S, c->dpad, (uint32_t)c->d, c->z, c->v, c->st, c->stream));
  return AMM_OK;
}

// exact product y = A (A^T x): one read of A where the columns fit the
// register-held form, else P1 = A^T x (dense) and P2 = A P1
int enq_exact(amm_ctx* c, uint32_t it, const uint32_t* gate_mode, uint32_t gate_mask) {
  if (c->gram_S) {
    HIPCHK(amm::launch_gram(c->dtype, c->A, c->ld, (uint32_t)c->d, c->n, c->dpad, c->x, c->P2, gate_mode, it,
                            gate_mask, c->stream));
    return AMM_OK;
  }
  amm::GatherDotArgs g{};
  g.A = c->A;
  g.ld = c->ld;
  g.d = (uint32_t)c->d;
  g.cap = (uint32_t)c->n;
  g.x = c->x;
  g.xm = nullptr;
  g.R = c->dot_dense.R;
  g.rows_per_split = c->dot_dense.rows_per_split;
  g.P1 = c->P1;
  g.mode = gate_mode;
  g.it = it;
  g.gate_mask = gate_mask;
  HIPCHK(amm::launch_gather_dot(c->dtype, g, c->stream));
  amm::GatherAxpyArgs a{};
  a.A = c->A;
  a.ld = c->ld;
  a.d = (uint32_t)c->d;
  a.dpad = c->dpad;
  a.cap = (uint32_t)c->n;
  a.cs = c->axpy_dense.cs;
  a.S = c->axpy_dense.S;
  a.P1 = c->P1;
  a.R = c->dot_dense.R;
  a.st = c->st;
  a.P2 = c->P2;
  a.P2s = c->P2s;
  a.mode = gate_mode;
  a.it = it;
  a.gate_mask = gate_mask;
  HIPCHK(amm::launch_gather_axpy(c->dtype, a, c->stream));
  return AMM_OK;
}

// A sampled product's second contraction carries the finish (y and the
// block partials of the Rayleigh quotient) where the row blocks fill the
// device; sanity mode (every column, unit weights) keeps the chunked form.
bool fused_finish(const amm_ctx* c, bool sanity) { return !sanity && amm::axpy_finish_applies(c->dpad); }

// sampled product with iteration `it`'s list at list slot `slot`; with
// fused_finish the product also writes y = c->y and the finish partials
int enq_amm(amm_ctx* c, uint32_t it, uint32_t slot, bool proposed, bool sanity,
            const uint32_t* gate_mode, uint32_t gate_mask, const amm::GatherTail* tail = nullptr) {
  const uint32_t cc = c->list_c;
  const uint64_t nseg = (cc + amm::kSampleSeg - 1) / amm::kSampleSeg;
  const DotPlan dp = sanity ? c->dot_dense : c->dot_list;
  const AxpyPlan ap = sanity ? c->axpy_dense : c->axpy_list;
  amm::GatherDotArgs g{};
  g.A = c->A;
  g.ld = c->ld;
  g.d = (uint32_t)c->d;
  g.idx = sanity ? nullptr : c->idx + (size_t)slot * cc;
  g.seg_count = sanity ? nullptr : c->segc + (size_t)slot * nseg;
  g.cap = sanity ? (uint32_t)c->n : cc;
  g.x = c->x;
  g.xm = c->xm;
  g.R = dp.R;
  g.rows_per_split = dp.rows_per_split;
  g.P1 = c->P1;
  g.mode = gate_mode;
  g.it = it;
  g.gate_mask = gate_mask;
  g.nt = (c->list_nt & 1u) != 0;
  g.tail = tail;
  HIPCHK(amm::launch_gather_dot(c->dtype, g, c->stream));
  if (fused_finish(c, sanity)) {
    amm::AxpyFinishArgs f{};
    f.A = c->A;
    f.ld = c->ld;
    f.d = (uint32_t)c->d;
    f.dpad = c->dpad;
    f.idx = g.idx;
    f.wgt = c->wgt + (size_t)slot * cc;
    f.seg_count = g.seg_count;
    f.cap = cc;
    f.P1 = c->P1;
    f.R = dp.R;
    f.alpha = proposed ? c->alpha : nullptr;
    f.st = c->st;
    f.x = c->x;
    f.y = c->y;
    f.ym = c->ym;
    f.v = c->v;
    f.eigvec = c->has_eigvec ? c->eigvec : nullptr;
    f.BP = c->BP;
    f.mode = gate_mode;
    f.it = it;
    f.gate_mask = gate_mask;
    f.nt = (c->list_nt & 2u) != 0;
    HIPCHK(amm::launch_axpy_finish(c->dtype, f, c->stream));
    return AMM_OK;
  }
  amm::GatherAxpyArgs a{};
  a.nt = (c->list_nt & 2u) != 0;
  a.A = c->A;
  a.ld = c->ld;
  a.d = (uint32_t)c->d;
  a.dpad = c->dpad;
  a.idx = g.idx;
  a.wgt = sanity ? nullptr : c->wgt + (size_t)slot * cc;
  a.seg_count = g.seg_count;
  a.cap = g.cap;
  a.cs = ap.cs;
  a.S = ap.S;
  a.P1 = c->P1;
  a.R = dp.R;
  a.alpha = proposed ? c->alpha : nullptr;
  a.st = c->st;
  a.P2 = c->P2;
  a.P2s = c->P2s;
  a.mode = gate_mode;
  a.it = it;
  a.gate_mask = gate_mask;
  HIPCHK(amm::launch_gather_axpy(c->dtype, a, c->stream));
  return AMM_OK;
}

amm::TailArgs tail_args(amm_ctx* c, uint32_t op, uint32_t nb, uint32_t it, bool norm,
                        const float* zq, bool stamp) {
  amm::TailArgs t{};
  t.op = op;
  t.BP = c->BP;
  t.nb = nb;
  t.st = c->st;
  t.norm = norm ? 1 : 0;
  t.y = c->y;
  t.x = c->x;
  t.xm = c->xm;
  t.z = zq;
  t.d = (uint32_t)c->d;
  t.dpad = c->dpad;
  t.mode = c->mode;
  t.it = it;
  t.eig = c->eig;
  t.err = c->err;
  t.ts = c->ts;
  t.stamp = stamp ? 1 : 0;
  return t;
}

// How the finish kernel runs after a product: always (the product left
// partials), never (a fused sampled product wrote y and the block partials
// itself), or only on exact steps (adaptive runs whose sampled steps fuse).
enum FinishMode { kFinishAlways = 0, kFinishNone = 1, kFinishExactOnly = 2 };

// y = the product's partial sums (+ base), then the tail: lambda_i, err_i,
// scale, step decision; with norm_inline also x = y / ||y|| (+ s, q)
int enq_finish(amm_ctx* c, uint32_t it, bool sanity, float* y_out, const float* ym_base,
               uint32_t adapt_kernel, uint64_t period, double thr, uint32_t step,
               bool norm_inline = false, const float* zq = nullptr, bool stamp = false,
               FinishMode fm = kFinishAlways, bool tail = true) {
  amm::FinishArgs f{};
  f.P2 = c->P2;
  f.S_exact = c->gram_S ? c->gram_S : c->axpy_dense.S;
  f.S_amm = sanity ? c->axpy_dense.S : c->axpy_list.S;
  f.dpad = c->dpad;
  f.d = (uint32_t)c->d;
  f.x = c->x;
  f.y = y_out;
  f.xm = c->xm;
  f.ym = const_cast<float*>(ym_base);
  f.v = c->v;
  f.P2s = c->P2s;
  f.eigvec = c->has_eigvec ? c->eigvec : nullptr;
  f.BP = c->BP;
  f.st = c->st;
  f.mode = c->mode;
  f.it = it;
  f.gate_mask = fm == kFinishExactOnly ? (1u << amm::kModeExactMemo) | (1u << amm::kModeExactNoMemo) : 0u;
  if (fm != kFinishNone) HIPCHK(amm::launch_finish(f, c->stream));
  if (!tail) return AMM_OK;  // the next dot pass carries it
  amm::TailArgs t = tail_args(c, amm::kTailFinish, (uint32_t)amm::finish_blocks(c->dpad), it,
                              norm_inline, zq, stamp);
  t.y = y_out;
  t.adapt_kernel = adapt_kernel;
  t.period = period;
  t.thr = thr;
  t.step = step;
  HIPCHK(amm::launch_tail(t, c->stream));
  return AMM_OK;
}

// x = y / ||y|| with the 1/||y|| already in the state, plus s, q when zq is
// given; `stamp` closes iteration it.  Up to 65536 rows: the tail kernel's
// blocks normalise (and form the parts of s, q); beyond: a grid pass and a
// tail for s, q.
int enq_norm(amm_ctx* c, uint32_t it, const float* zq, bool stamp) {
  if (c->dpad <= kFuseNormRows) {
    HIPCHK(amm::launch_tail(tail_args(c, amm::kTailNorm, 0, it, true, zq, stamp), c->stream));
  } else {
    HIPCHK(amm::launch_normalize(c->y, c->x, (uint32_t)c->d, c->dpad, c->st, c->xm, zq, c->BP,
                                 c->stream));
    const uint32_t nb = zq ? (uint32_t)amm::normalize_blocks(c->dpad) : 0u;
    HIPCHK(amm::launch_tail(tail_args(c, amm::kTailSQ, nb, it, false, nullptr, stamp), c->stream));
  }
  return AMM_OK;
}

// x0 ~ N(0, I), normalised (curandGenerateNormal + normalize<512>)
// x0's normalisation (the raw x0 and its block sums come from the run
// prologue, k_begin; nothing in between touches y or BP)
int enq_x0(amm_ctx* c, const float* zq) {
  const bool small = c->dpad <= kFuseNormRows;
  HIPCHK(amm::launch_tail(tail_args(c, amm::kTailScale, (uint32_t)amm::x0_blocks(c->dpad), 0, small,
                                    zq, false),
                          c->stream));
  if (!small) {
    HIPCHK(amm::launch_normalize(c->y, c->x, (uint32_t)c->d, c->dpad, c->st, c->xm, zq, c->BP,
                                 c->stream));
    if (zq)
      HIPCHK(amm::launch_tail(tail_args(c, amm::kTailSQ, (uint32_t)amm::normalize_blocks(c->dpad), 0,
                                        false, nullptr, false),
                              c->stream));
  }
  return AMM_OK;
}

// s, q of x against x_memo for a stand-alone proposed product
int enq_dots(amm_ctx* c) {
  HIPCHK(amm::launch_dots(c->x, c->xm, c->z, c->dpad, c->BP, c->stream));
  HIPCHK(amm::launch_tail(tail_args(c, amm::kTailSQ, (uint32_t)amm::dots_blocks(c->dpad), 0, false,
                                    nullptr, false),
                          c->stream));
  return AMM_OK;
}

// Step kind of iteration i for a static schedule; mirrors orc_is_exact_step
// (powerMethod_by_previous.cuh:135-138, powerMethod_by_proposed.cuh:156-161,
// powerMethod_by_proposed_v2.cuh:234-326).
std::vector<uint32_t> static_modes(const amm_params& p) {
  std::vector<uint32_t> m(p.step);
  uint64_t last_exact = 0;
  const uint64_t P = p.vr_period;
  const bool vr = !(p.flags & AMM_FLAG_NO_VR);
  for (uint64_t i = 0; i < p.step; ++i) {
    uint32_t md;
    const bool due = vr && P > 0 && (i + 1) % P == 0;
    if (p.kernel == AMM_KERNEL_EXACT) {
      md = amm::kModeExactNoMemo;
    } else if (p.kernel == AMM_KERNEL_PREVIOUS) {
      md = due ? amm::kModeExactMemo : amm::kModeAmmPrevious;
    } else if (p.kernel == AMM_KERNEL_PROPOSED) {
      md = (i > last_exact + 1 && due) ? amm::kModeExactMemo : amm::kModeAmmProposed;
    } else {
      if (i + 1 < P) md = amm::kModeAmmPrevious;
      else if (i + 1 == P) md = amm::kModeExactNoMemo;
      else md = due ? amm::kModeExactMemo : amm::kModeAmmProposed;
    }
    if (md == amm::kModeExactMemo || md == amm::kModeExactNoMemo) last_exact = i;
    m[i] = md;
  }
  return m;
}

bool is_adaptive(const amm_params& p) {
  return p.vr_adaptive_threshold > 0.0 && !(p.flags & AMM_FLAG_NO_VR) &&
         (p.kernel == AMM_KERNEL_PREVIOUS || p.kernel == AMM_KERNEL_PROPOSED);
}

int validate(const amm_ctx* c, const amm_params* p) {
  if (!p) return fail(AMM_E_INVALID, "params is NULL");
  if (p->kernel < AMM_KERNEL_EXACT || p->kernel > AMM_KERNEL_PROPOSED_V2)
    return fail(AMM_E_INVALID, "unknown kernel " + std::to_string(p->kernel));
  if (p->step < 1 || p->step > (1ull << 31)) return fail(AMM_E_INVALID, "step must be in [1, 2^31]");
  if (p->kernel != AMM_KERNEL_EXACT) {
    if (p->c < 1 || p->c > c->n)
      return fail(AMM_E_INVALID, "c must satisfy 1 <= c <= n (host_memory.h:166-167)");
    if ((uint64_t)p->c * p->step > (1ull << 34))
      return fail(AMM_E_INVALID, "c * step exceeds the sample-list limit 2^34");
  }
  if (p->kernel == AMM_KERNEL_PROPOSED_V2 && (p->vr_period < 1 || p->vr_period > p->step))
    return fail(AMM_E_INVALID,
                "proposedAMM_v2 needs 1 <= vr_period <= step (powerMethod_by_proposed_v2.cuh:234)");
  if (!(p->vr_adaptive_threshold >= 0.0)) return fail(AMM_E_INVALID, "vr_adaptive_threshold < 0");
  return AMM_OK;
}

int upload_run_state(amm_ctx* c, uint64_t seed, uint32_t trial) {
  c->h_seed = seed;
  c->h_trial = trial;
  HIPCHK(hipMemcpyAsync((char*)c->st + offsetof(DevState, seed), &c->h_seed, sizeof(uint64_t),
                        hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync((char*)c->st + offsetof(DevState, trial), &c->h_trial, sizeof(uint32_t),
                        hipMemcpyHostToDevice, c->stream));
  return AMM_OK;
}

// All kernels of one run, in stream order.
int enqueue_run(amm_ctx* c, const amm_params& p, const std::vector<uint32_t>& modes) {
  int rc;
  const int K = p.kernel;
  const uint64_t P = p.vr_period;
  const bool adaptive = is_adaptive(p);
  const bool sanity = K != AMM_KERNEL_EXACT && p.sanity && p.c == c->n;
  const uint32_t step = (uint32_t)p.step;
  HIPCHK(amm::launch_begin(c->xm, c->ym, c->dpad, c->st, c->mode, adaptive ? modes[0] : 0, c->y,
                           (uint32_t)c->d, c->BP, c->stream));
  const bool need_plain = K == AMM_KERNEL_PREVIOUS || K == AMM_KERNEL_PROPOSED_V2;
  const bool need_cent = K == AMM_KERNEL_PROPOSED || K == AMM_KERNEL_PROPOSED_V2;
  // the exact power method needs no column statistics; the centred kernels
  // take z = A alpha from the same read of A when it fits
  bool fused_z = false;
  if (K != AMM_KERNEL_EXACT && (!(p.flags & AMM_FLAG_KEEP_STATS) || !c->stats_valid)) {
    fused_z = need_cent && z_fusable(c);
    if ((rc = fused_z ? enq_stats_z(c) : enq_stats(c))) return rc;
  }
  if (need_plain && !sanity)
    if ((rc = enq_cdf(c, 0, false))) return rc;
  if (need_cent)
    if ((rc = enq_cdf(c, 1, false))) return rc;  // also ||alpha||^2
  if (fused_z) {
    if ((rc = enq_z_partials(c))) return rc;
  } else if (K == AMM_KERNEL_PROPOSED) {
    if ((rc = enq_z(c))) return rc;
  }
  if (!sanity && K != AMM_KERNEL_EXACT) {
    if (K == AMM_KERNEL_PREVIOUS) {
      if ((rc = enq_sample(c, 0, 0, step, 0, p.c))) return rc;
    } else if (K == AMM_KERNEL_PROPOSED) {
      if ((rc = enq_sample(c, 1, 0, step, 0, p.c))) return rc;
    } else {
      const uint32_t np = (uint32_t)(P - 1);
      if ((rc = enq_sample(c, 0, 0, np, 0, p.c))) return rc;
      if (step > P)
        if ((rc = enq_sample(c, 1, (uint32_t)P, step - (uint32_t)P, 0, p.c))) return rc;
    }
  }
  const uint32_t gate_exact = 1u << amm::kModeExactMemo;
  const uint32_t gate_amm = (1u << amm::kModeAmmPrevious) | (1u << amm::kModeAmmProposed);
  const bool fuse = fused_finish(c, sanity);
  auto sampled = [&](uint32_t i) {
    return modes[i] == amm::kModeAmmPrevious || modes[i] == amm::kModeAmmProposed;
  };
  // the tail of a step followed by a sampled step (static schedule, sample
  // lists, inline normalisation, d a multiple of the dot pass's row split)
  // rides in the next dot pass, and so does x0's normalisation
  const bool carry_ok = !sanity && !adaptive && c->dpad <= kFuseNormRows &&
                        c->dot_list.R <= amm::kTailMaxBlocks && c->dot_list.rows_per_split &&
                        c->d % c->dot_list.rows_per_split == 0;
  amm::GatherTail carried{};
  bool have_carried = false;
  // x0 (curandGenerateNormal + normalize<512>)
  const float* z0 = K == AMM_KERNEL_PROPOSED ? c->z : nullptr;
  if (carry_ok && sampled(0)) {
    carried.BP = c->BP;
    carried.nb = (uint32_t)amm::x0_blocks(c->dpad);
    carried.st = c->st;
    carried.y = c->y;
    carried.x = c->x;
    carried.z = z0;
    carried.x0 = 1;
    have_carried = true;
  } else if ((rc = enq_x0(c, z0))) {
    return rc;
  }
  for (uint32_t i = 0; i < step; ++i) {
    bool z_after = false;
    FinishMode fm = kFinishAlways;
    if (adaptive) {
      if ((rc = enq_exact(c, i, c->mode, gate_exact))) return rc;
      if ((rc = enq_amm(c, i, i, K == AMM_KERNEL_PROPOSED, sanity, c->mode, gate_amm))) return rc;
      if (fuse) fm = kFinishExactOnly;
    } else {
      const uint32_t md = modes[i];
      if (md == amm::kModeExactMemo || md == amm::kModeExactNoMemo) {
        if ((rc = enq_exact(c, i, nullptr, 0))) return rc;
        z_after = K == AMM_KERNEL_PROPOSED_V2 && i + 1 == P && !fused_z;
      } else {
        if ((rc = enq_amm(c, i, i, md == amm::kModeAmmProposed, sanity, nullptr, 0,
                          have_carried ? &carried : nullptr)))
          return rc;
        have_carried = false;
        if (fuse) fm = kFinishNone;
      }
    }
    const bool want_sq = K == AMM_KERNEL_PROPOSED || (K == AMM_KERNEL_PROPOSED_V2 && i + 1 >= P);
    if (carry_ok && !z_after && i + 1 < step && sampled(i + 1)) {
      // the finish kernel still runs (y and the block partials) unless the
      // second contraction carried it
      if (fm != kFinishNone)
        if ((rc = enq_finish(c, i, sanity, c->y, c->ym, 0u, P, 0.0, step, false, nullptr, false, fm, false)))
          return rc;
      carried.BP = c->BP;
      carried.nb = (uint32_t)amm::finish_blocks(c->dpad);
      carried.st = c->st;
      carried.y = c->y;
      carried.x = c->x;
      carried.z = want_sq ? c->z : nullptr;
      carried.eig = c->eig;
      carried.err = c->err;
      carried.ts = c->ts;
      carried.it = i;
      carried.exact = sampled(i) ? 0u : 1u;
      carried.x0 = 0;
      have_carried = true;
      continue;
    }
    // the normalisation rides in the finish tail unless d is large or z is
    // only formed after this iteration's product (the v2 transition)
    const bool inline_norm = c->dpad <= kFuseNormRows && !z_after;
    if ((rc = enq_finish(c, i, sanity, c->y, c->ym, adaptive ? (uint32_t)K : 0u, P,
                         p.vr_adaptive_threshold, step, inline_norm, want_sq ? c->z : nullptr,
                         inline_norm, fm)))
      return rc;
    if (z_after)
      if ((rc = enq_z(c))) return rc;  // [z, .] = A [alpha, .] at the v2 transition
    if (!inline_norm)
      if ((rc = enq_norm(c, i, want_sq ? c->z : nullptr, true))) return rc;
  }
  return AMM_OK;
}

// Reads and clears the sticky device status after a synchronisation.
int check_status(amm_ctx* c) {
  uint32_t stt = 0;
  HIPCHK(copy_sync(c->stream, &stt, (char*)c->st + offsetof(DevState, status), sizeof(uint32_t),
                   hipMemcpyDeviceToHost));
  if (stt) {
    const uint32_t zero = 0;
    HIPCHK(copy_sync(c->stream, (char*)c->st + offsetof(DevState, status), &zero, sizeof(uint32_t),
                     hipMemcpyHostToDevice));
    if (stt & amm::kStatusNonFinite)
      return fail(AMM_E_INVALID, "the matrix has non-finite entries (NaN or Inf column norms)");
    return fail(AMM_E_STATE, "device status " + std::to_string(stt));
  }
  return AMM_OK;
}

bool same_key(const amm_params& a, const amm_params& b) {
  return a.kernel == b.kernel && a.c == b.c && a.step == b.step && a.vr_period == b.vr_period &&
         a.vr_adaptive_threshold == b.vr_adaptive_threshold && a.sanity == b.sanity &&
         a.flags == b.flags;
}

void drop_graph(amm_ctx* c) {
  if (c->exec) (void)hipGraphExecDestroy(c->exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->exec = nullptr;
  c->graph = nullptr;
  c->graph_valid = false;
}

}  // namespace

// ===========================================================================
extern "C" {

const char* amm_last_error(void) { return g_err.c_str(); }

int amm_abi_version(void) { return AMM_ABI_VERSION; }

int amm_ctx_create(const void* A, int32_t dtype, uint64_t d, uint64_t n, uint64_t ld, void* stream,
                   amm_ctx** out) {
  if (!out) return fail(AMM_E_INVALID, "out is NULL");
  *out = nullptr;
  std::lock_guard<std::recursive_mutex> lk(amm_host::device_mutex());
  if (!A) return fail(AMM_E_INVALID, "A is NULL");
  if (dtype != AMM_DTYPE_F32 && dtype != AMM_DTYPE_BF16) return fail(AMM_E_INVALID, "unknown dtype");
  if (d < 1 || n < 1) return fail(AMM_E_INVALID, "d and n must be >= 1");
  if (d >= (1ull << 31) || n >= (1ull << 31)) return fail(AMM_E_INVALID, "d and n must be < 2^31");
  if (ld < d) return fail(AMM_E_INVALID, "ld must be >= d");
  const int V = dtype == AMM_DTYPE_BF16 ? 8 : 4;
  if (ld % V) return fail(AMM_E_INVALID, "ld must be a multiple of 16 bytes of elements");
  if (((uintptr_t)A) & 15u) return fail(AMM_E_INVALID, "A must be 16-byte aligned");
  amm_ctx* c = new amm_ctx();
  c->A = A;
  c->dtype = dtype;
  c->d = d;
  c->n = n;
  c->ld = ld;
  c->dpad = (d + 7) / 8 * 8;
  c->V = V;
  int rc = AMM_OK;
  auto bail = [&](int r) {
    amm_ctx_destroy(c);
    return r;
  };
  if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(AMM_E_HIP, "hipGetDevice failed"));
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    // blocking stream: ordered against work on the null stream (the caller's
    // uploads of A and of the vectors)
    if (hipStreamCreateWithFlags(&c->stream, hipStreamDefault) != hipSuccess)
      return bail(fail(AMM_E_HIP, "hipStreamCreate failed"));
    c->own_stream = true;
  }
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) == hipSuccess && khz > 0)
    c->wall_khz = khz;
  size_t cap = 0;
  const size_t dp = c->dpad;
  if ((rc = grow(c->alpha, cap, n)) || (cap = 0, rc = grow(c->nrm2, cap, n)) ||
      (cap = 0, rc = grow(c->cnrm2, cap, n)) || (cap = 0, rc = grow(c->cdf[0], cap, n)) ||
      (cap = 0, rc = grow(c->cdf[1], cap, n)) || (cap = 0, rc = grow(c->z, cap, dp)) ||
      (cap = 0, rc = grow(c->v, cap, dp)) || (cap = 0, rc = grow(c->x, cap, dp)) ||
      (cap = 0, rc = grow(c->y, cap, dp)) || (cap = 0, rc = grow(c->xm, cap, dp)) ||
      (cap = 0, rc = grow(c->ym, cap, dp)) || (cap = 0, rc = grow(c->eigvec, cap, dp)) ||
      (cap = 0, rc = grow(c->st, cap, 1)) ||
      // block partials: 3 doubles per finish block bound every reduction kernel
      (cap = 0, rc = grow(c->BP, cap, 3 * amm::finish_blocks(dp) + 64)) ||
      (cap = 0, rc = grow(c->tile_max, cap, amm::cdf_tiles(n))) ||
      (cap = 0, rc = grow(c->wmax, cap, 2 * (size_t)amm::kColstatsMaxBlocks)) ||
      (cap = 0, rc = grow(c->tile_sum, cap, amm::cdf_tiles(n))) ||
      (cap = 0, rc = grow(c->tile_lp, cap, amm::cdf_tiles(n))) ||
      (cap = 0, rc = grow(c->tile_a2, cap, amm::cdf_tiles(n))))
    return bail(rc);
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return bail(fail(AMM_E_HIP, "hipEventCreate failed"));
  if (const char* e = std::getenv("AMM_LIST_NT")) c->list_nt = (uint32_t)std::strtoul(e, nullptr, 0) & 3u;
  if ((rc = ensure_plans(c, 0))) return bail(rc);
  *out = c;
  return AMM_OK;
}

int amm_ctx_destroy(amm_ctx* c) {
  if (!c) return AMM_OK;
  std::lock_guard<std::recursive_mutex> lk(amm_host::device_mutex());
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  void* bufs[] = {c->alpha, c->nrm2,  c->cnrm2,    c->cdf[0],  c->cdf[1], c->z,   c->v,
                  c->x,     c->y,     c->xm,       c->ym,      c->eigvec, c->st,  c->BP,
                  c->mode,  c->eig,   c->err,      c->ts,      c->idx,    c->wgt, c->segc,
                  c->P1,    c->P2,    c->P2s,      c->tile_sum, c->tile_lp, c->tile_a2,
                  c->tile_max, c->wmax, c->tvec, c->tbp, c->ZP};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return AMM_OK;
}

int amm_set_exact_eigenvector(amm_ctx* c, const double* v, uint64_t len) {
  if (!c) return fail(AMM_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = set_device(c))) return rc;
  if (!v) {
    // a captured run baked the eigenvector pointer into its finish kernels
    c->has_eigvec = false;
    drop_graph(c);
    return AMM_OK;
  }
  if (len != c->d) return fail(AMM_E_INVALID, "eigenvector length must equal d");
  HIPCHK(copy_sync(c->stream, c->eigvec, v, sizeof(double) * len, hipMemcpyHostToDevice));
  c->has_eigvec = true;
  drop_graph(c);
  return AMM_OK;
}

int amm_column_stats(amm_ctx* c, float* alpha_h, double* nrm2_h, double* cnrm2_h, double* alpha2_h) {
  if (!c) return fail(AMM_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = set_device(c))) return rc;
  if ((rc = enq_stats(c))) return rc;
  if ((rc = enq_cdf(c, 1, false))) return rc;
  c->stats_valid = true;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (alpha_h) HIPCHK(copy_sync(c->stream, alpha_h, c->alpha, sizeof(float) * c->n, hipMemcpyDeviceToHost));
  if (nrm2_h) HIPCHK(copy_sync(c->stream, nrm2_h, c->nrm2, sizeof(double) * c->n, hipMemcpyDeviceToHost));
  if (cnrm2_h) HIPCHK(copy_sync(c->stream, cnrm2_h, c->cnrm2, sizeof(double) * c->n, hipMemcpyDeviceToHost));
  if (alpha2_h)
    HIPCHK(copy_sync(c->stream, alpha2_h, (char*)c->st + offsetof(DevState, alpha2), sizeof(double),
                     hipMemcpyDeviceToHost));
  return AMM_OK;
}

int amm_build_cdf(amm_ctx* c, int32_t centred, uint64_t* cdf_h, uint64_t* total_h, uint64_t* last_h) {
  if (!c) return fail(AMM_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = set_device(c))) return rc;
  const int which = centred ? 1 : 0;
  if (!c->stats_valid) {
    if ((rc = enq_stats(c))) return rc;
    c->stats_valid = true;
  }
  if ((rc = enq_cdf(c, which, false))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  if ((rc = check_status(c))) return rc;
  if (cdf_h) HIPCHK(copy_sync(c->stream, cdf_h, c->cdf[which], sizeof(uint64_t) * c->n, hipMemcpyDeviceToHost));
  if (total_h)
    HIPCHK(copy_sync(c->stream, total_h, (char*)c->st + offsetof(DevState, total) + sizeof(uint64_t) * which,
                     sizeof(uint64_t), hipMemcpyDeviceToHost));
  if (last_h)
    HIPCHK(copy_sync(c->stream, last_h, (char*)c->st + offsetof(DevState, last_pos) + sizeof(uint64_t) * which,
                     sizeof(uint64_t), hipMemcpyDeviceToHost));
  return AMM_OK;
}

int amm_sample(amm_ctx* c, int32_t centred, uint64_t seed, uint32_t iter, uint32_t trial, uint32_t cc,
               uint32_t* idx_h, float* wgt_h, uint32_t* segc_h) {
  if (!c) return fail(AMM_E_INVALID, "ctx is NULL");
  if (cc < 1 || cc > c->n) return fail(AMM_E_INVALID, "c must satisfy 1 <= c <= n");
  int rc;
  if ((rc = set_device(c))) return rc;
  drop_graph(c);
  if ((rc = ensure_lists(c, 1, cc))) return rc;
  if (!c->stats_valid) {
    if ((rc = enq_stats(c))) return rc;
    c->stats_valid = true;
  }
  const int which = centred ? 1 : 0;
  if ((rc = enq_cdf(c, which, false))) return rc;
  if ((rc = upload_run_state(c, seed, trial))) return rc;
  if ((rc = enq_sample(c, which, iter, 1, iter, cc))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  if ((rc = check_status(c))) return rc;
  const size_t nseg = (cc + amm::kSampleSeg - 1) / amm::kSampleSeg;
  if (idx_h) HIPCHK(copy_sync(c->stream, idx_h, c->idx, sizeof(uint32_t) * cc, hipMemcpyDeviceToHost));
  if (wgt_h) HIPCHK(copy_sync(c->stream, wgt_h, c->wgt, sizeof(float) * cc, hipMemcpyDeviceToHost));
  if (segc_h) HIPCHK(copy_sync(c->stream, segc_h, c->segc, sizeof(uint32_t) * nseg, hipMemcpyDeviceToHost));
  return AMM_OK;
}

int amm_square_of_A(amm_ctx* c, int32_t kernel, int32_t sanity, uint32_t cc, uint64_t seed, uint32_t iter,
                    uint32_t trial, const float* x_dev, const float* xm_dev, const float* ym_dev,
                    float* y_dev) {
  if (!c) return fail(AMM_E_INVALID, "ctx is NULL");
  if (kernel != AMM_KERNEL_PREVIOUS && kernel != AMM_KERNEL_PROPOSED)
    return fail(AMM_E_INVALID, "kernel must be previousAMM or proposedAMM");
  if (cc < 1 || cc > c->n) return fail(AMM_E_INVALID, "c must satisfy 1 <= c <= n");
  if (!x_dev || !y_dev) return fail(AMM_E_INVALID, "x and y must be device pointers");
  int rc;
  if ((rc = set_device(c))) return rc;
  drop_graph(c);
  const bool prop = kernel == AMM_KERNEL_PROPOSED;
  const bool san = sanity && cc == c->n;
  const int which = prop ? 1 : 0;
  if ((rc = ensure_plans(c, cc))) return rc;
  if ((rc = ensure_lists(c, 1, cc))) return rc;
  if ((rc = ensure_step(c, 1))) return rc;
  const size_t db = sizeof(float) * c->d;
  HIPCHK(hipMemcpyAsync(c->x, x_dev, db, hipMemcpyDeviceToDevice, c->stream));
  if (xm_dev) HIPCHK(hipMemcpyAsync(c->xm, xm_dev, db, hipMemcpyDeviceToDevice, c->stream));
  else HIPCHK(hipMemsetAsync(c->xm, 0, sizeof(float) * c->dpad, c->stream));
  if (ym_dev) HIPCHK(hipMemcpyAsync(c->ym, ym_dev, db, hipMemcpyDeviceToDevice, c->stream));
  else HIPCHK(hipMemsetAsync(c->ym, 0, sizeof(float) * c->dpad, c->stream));
  if (!c->stats_valid) {
    if ((rc = enq_stats(c))) return rc;
    c->stats_valid = true;
  }
  if ((rc = enq_cdf(c, 1, false))) return rc;  // ||alpha||^2
  if (!prop && !san)
    if ((rc = enq_cdf(c, 0, false))) return rc;
  if (prop) {
    if ((rc = enq_z(c))) return rc;
    if ((rc = enq_dots(c))) return rc;
  }
  if ((rc = upload_run_state(c, seed, trial))) return rc;
  if (!san)
    if ((rc = enq_sample(c, which, iter, 1, iter, cc))) return rc;
  c->h_mode = prop ? amm::kModeAmmProposed : amm::kModeAmmPrevious;
  HIPCHK(hipMemcpyAsync(c->mode, &c->h_mode, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  if ((rc = enq_amm(c, 0, 0, prop, san, nullptr, 0))) return rc;
  if ((rc = enq_finish(c, 0, san, c->y, c->ym, 0, 0, 0.0, 1, false, nullptr, false,
                       fused_finish(c, san) ? kFinishNone : kFinishAlways)))
    return rc;
  HIPCHK(hipMemcpyAsync(y_dev, c->y, db, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_status(c);
}

int amm_exact_product(amm_ctx* c, const float* x_dev, float* y_dev) {
  if (!c) return fail(AMM_E_INVALID, "ctx is NULL");
  if (!x_dev || !y_dev) return fail(AMM_E_INVALID, "x and y must be device pointers");
  int rc;
  if ((rc = set_device(c))) return rc;
  drop_graph(c);
  if ((rc = ensure_step(c, 1))) return rc;
  const size_t db = sizeof(float) * c->d;
  HIPCHK(hipMemcpyAsync(c->x, x_dev, db, hipMemcpyDeviceToDevice, c->stream));
  c->h_mode = amm::kModeExactNoMemo;
  HIPCHK(hipMemcpyAsync(c->mode, &c->h_mode, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  if ((rc = enq_exact(c, 0, nullptr, 0))) return rc;
  if ((rc = enq_finish(c, 0, false, c->y, c->ym, 0, 0, 0.0, 1))) return rc;
  HIPCHK(hipMemcpyAsync(y_dev, c->y, db, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return AMM_OK;
}

int amm_power_method(amm_ctx* c, const amm_params* p, double* time_ms, double* eig, double* err) {
  if (!c) return fail(AMM_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = validate(c, p))) return rc;
  if ((rc = set_device(c))) return rc;
  const bool use_graph = !(p->flags & AMM_FLAG_NO_GRAPH);
  const uint32_t cc = p->kernel == AMM_KERNEL_EXACT ? 0 : p->c;
  const bool rebuild = !c->graph_valid || !same_key(c->graph_key, *p) || c->list_c != cc;
  // buffers (re)sized and the graph captured under the device lock: no
  // other thread of this process allocates or frees meanwhile
  std::unique_lock<std::recursive_mutex> lk(amm_host::device_mutex(), std::defer_lock);
  if (rebuild) lk.lock();
  if (rebuild) {
    drop_graph(c);
    if ((rc = ensure_plans(c, cc))) return rc;
    if ((rc = ensure_step(c, p->step))) return rc;
    if (cc)
      if ((rc = ensure_lists(c, p->step, cc))) return rc;
    if ((p->kernel == AMM_KERNEL_PROPOSED || p->kernel == AMM_KERNEL_PROPOSED_V2) &&
        amm::colstats_fuses_z(c->dtype, (uint32_t)c->d))
      if ((rc = grow(c->ZP, c->ZP_cap,
                     (size_t)amm::colstats_z_blocks(c->dtype, (uint32_t)c->d, c->n) * c->dpad)))
        return rc;
    c->host_modes = static_modes(*p);
    if (is_adaptive(*p)) {
      // iteration 0: prev_cos = 0, last_exact = 0
      const uint64_t P = p->vr_period;
      const bool due0 = P > 0 && 1 % P == 0;
      const bool ex0 = p->kernel == AMM_KERNEL_PREVIOUS ? due0 : false;
      c->host_modes[0] = ex0 ? amm::kModeExactMemo
                             : (p->kernel == AMM_KERNEL_PROPOSED ? amm::kModeAmmProposed
                                                                  : amm::kModeAmmPrevious);
    }
  }
  HIPCHK(hipMemcpyAsync(c->mode, c->host_modes.data(), sizeof(uint32_t) * p->step,
                        hipMemcpyHostToDevice, c->stream));
  if ((rc = upload_run_state(c, p->seed, p->trial))) return rc;
  if (use_graph && rebuild) {
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_run(c, *p, c->host_modes);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess) return fail(AMM_E_HIP, std::string("capture failed: ") + hipGetErrorString(ec));
    c->graph = g;
    HIPCHK(hipGraphInstantiate(&c->exec, g, nullptr, nullptr, 0));
    c->graph_valid = true;
    c->graph_key = *p;
  }
  if (lk.owns_lock()) lk.unlock();
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  if (use_graph) {
    HIPCHK(hipGraphLaunch(c->exec, c->stream));
  } else {
    if ((rc = enqueue_run(c, *p, c->host_modes))) return rc;
  }
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  c->last_step = p->step;
  if (p->kernel != AMM_KERNEL_EXACT) c->stats_valid = true;
  if ((rc = check_status(c))) return rc;
  const size_t S = p->step;
  if (eig) HIPCHK(copy_sync(c->stream, eig, c->eig, sizeof(double) * S, hipMemcpyDeviceToHost));
  if (err) HIPCHK(copy_sync(c->stream, err, c->err, sizeof(double) * S, hipMemcpyDeviceToHost));
  if (time_ms) {
    std::vector<int64_t> ts(S);
    int64_t t0 = 0;
    HIPCHK(copy_sync(c->stream, ts.data(), c->ts, sizeof(int64_t) * S, hipMemcpyDeviceToHost));
    HIPCHK(copy_sync(c->stream, &t0, (char*)c->st + offsetof(DevState, t_start), sizeof(int64_t),
                     hipMemcpyDeviceToHost));
    for (size_t i = 0; i < S; ++i) time_ms[i] = (double)(ts[i] - t0) / (double)c->wall_khz;
  }
  return AMM_OK;
}

int amm_last_modes(amm_ctx* c, uint32_t* modes_h, uint64_t len) {
  if (!c || !modes_h) return fail(AMM_E_INVALID, "NULL argument");
  if (len > c->mode_cap) return fail(AMM_E_INVALID, "len exceeds the last run's step");
  int rc;
  if ((rc = set_device(c))) return rc;
  HIPCHK(copy_sync(c->stream, modes_h, c->mode, sizeof(uint32_t) * len, hipMemcpyDeviceToHost));
  return AMM_OK;
}

int amm_get_vector(amm_ctx* c, int32_t which, float* host, uint64_t len) {
  if (!c || !host) return fail(AMM_E_INVALID, "NULL argument");
  if (len != c->d) return fail(AMM_E_INVALID, "len must equal d");
  int rc;
  if ((rc = set_device(c))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(copy_sync(c->stream, host, which ? c->y : c->x, sizeof(float) * len, hipMemcpyDeviceToHost));
  return AMM_OK;
}

int amm_last_run_ms(amm_ctx* c, double* ms) {
  if (!c || !ms) return fail(AMM_E_INVALID, "NULL argument");
  *ms = c->last_ms;
  return AMM_OK;
}

int amm_generate_symmetric(void* A_dev, uint64_t m, uint64_t ld, int32_t dist, float p1, float p2,
                           uint64_t seed, void* stream, float* delta_out) {
  if (!A_dev) return fail(AMM_E_INVALID, "A is NULL");
  if (m < 1 || m >= (1ull << 31) || ld < m) return fail(AMM_E_INVALID, "need 1 <= m < 2^31, ld >= m");
  if (dist < AMM_DIST_GAUSSIAN || dist > AMM_DIST_UNIFORM)
    return fail(AMM_E_INVALID, "unknown distribution " + std::to_string(dist));
  hipStream_t s = (hipStream_t)stream;
  uint32_t* dbits = nullptr;
  if (hipMalloc((void**)&dbits, sizeof(uint32_t)) != hipSuccess)
    return fail(AMM_E_NOMEM, "device allocation failed");
  uint32_t h = 0;
  hipError_t e = amm::launch_generate_symmetric((float*)A_dev, m, ld, dist, p1, p2, seed, dbits, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&h, dbits, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(dbits);
  if (e != hipSuccess) return fail(AMM_E_HIP, std::string("generate failed: ") + hipGetErrorString(e));
  float delta;
  std::memcpy(&delta, &h, sizeof(float));
  if (delta_out) *delta_out = delta;
  return AMM_OK;
}

int amm_sigma_max(amm_ctx* c, const amm_params* p, double* sigma, float* u_dev, float* v_dev) {
  if (!c) return fail(AMM_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = amm_power_method(c, p, nullptr, nullptr, nullptr))) return rc;
  if ((rc = grow(c->tvec, c->tvec_cap, c->n)) ||
      (rc = grow(c->tbp, c->tbp_cap, amm::reduce_p1_blocks(c->n))))
    return rc;
  // t = A^T u for the final iterate u = x (one exact pass)
  amm::GatherDotArgs g{};
  g.A = c->A;
  g.ld = c->ld;
  g.d = (uint32_t)c->d;
  g.cap = (uint32_t)c->n;
  g.x = c->x;
  g.R = c->dot_dense.R;
  g.rows_per_split = c->dot_dense.rows_per_split;
  g.P1 = c->P1;
  HIPCHK(amm::launch_gather_dot(c->dtype, g, c->stream));
  HIPCHK(amm::launch_reduce_p1(c->P1, g.R, c->n, c->tvec, c->tbp, c->stream));
  amm::TailArgs t =
      tail_args(c, amm::kTailNorm2, (uint32_t)amm::reduce_p1_blocks(c->n), 0, false, nullptr, false);
  t.BP = c->tbp;
  HIPCHK(amm::launch_tail(t, c->stream));
  if (v_dev) HIPCHK(amm::launch_unit(c->tvec, c->n, c->st, v_dev, c->stream));
  if (u_dev)
    HIPCHK(hipMemcpyAsync(u_dev, c->x, sizeof(float) * c->d, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (sigma)
    HIPCHK(copy_sync(c->stream, sigma, (char*)c->st + offsetof(DevState, sigma), sizeof(double),
                     hipMemcpyDeviceToHost));
  return AMM_OK;
}

int amm_rank1_update(void* out, int32_t dtype, uint64_t d, uint64_t n, uint64_t ld, float g,
                     const float* u_dev, const float* v_dev, void* stream) {
  if (!out || !u_dev || !v_dev) return fail(AMM_E_INVALID, "NULL argument");
  if (dtype != AMM_DTYPE_F32 && dtype != AMM_DTYPE_BF16) return fail(AMM_E_INVALID, "unknown dtype");
  if (d < 1 || n < 1 || ld < d || d >= (1ull << 31) || n >= (1ull << 31))
    return fail(AMM_E_INVALID, "need 1 <= d, n < 2^31 and ld >= d");
  HIPCHK(amm::launch_rank1(dtype, out, (uint32_t)d, n, ld, g, u_dev, v_dev, (hipStream_t)stream));
  return AMM_OK;
}

int amm_rank1_update_batched(void* out, int32_t dtype, uint64_t d, uint64_t n, uint64_t ld, uint64_t count,
                             uint64_t out_stride, const float* g_dev, const float* u_dev, uint64_t u_stride,
                             const float* v_dev, uint64_t v_stride, void* stream) {
  if (!out || !g_dev || !u_dev || !v_dev) return fail(AMM_E_INVALID, "NULL argument");
  if (dtype != AMM_DTYPE_F32 && dtype != AMM_DTYPE_BF16) return fail(AMM_E_INVALID, "unknown dtype");
  if (d < 1 || n < 1 || ld < d || d >= (1ull << 31) || n >= (1ull << 31))
    return fail(AMM_E_INVALID, "need 1 <= d, n < 2^31 and ld >= d");
  if (count < 1 || count >= 65536) return fail(AMM_E_INVALID, "count must be in [1, 65535]");
  if (count > 1 && (out_stride < (n - 1) * ld + d || u_stride < d || v_stride < n))
    return fail(AMM_E_INVALID, "strides must separate the matrices and vectors of the stack");
  HIPCHK(amm::launch_rank1_batched(dtype, out, (uint32_t)d, n, ld, count, out_stride, 0.0f, g_dev, u_dev,
                                   u_stride, v_dev, v_stride, (hipStream_t)stream));
  return AMM_OK;
}

int amm_f32_to_bf16(const float* src, void* dst, uint64_t count, void* stream) {
  if (!src || !dst) return fail(AMM_E_INVALID, "NULL argument");
  HIPCHK(amm::launch_f32_to_bf16(src, (uint16_t*)dst, count, (hipStream_t)stream));
  return AMM_OK;
}

int amm_time_column_stats(amm_ctx* c, uint32_t iters, double* ms) {
  if (!c || !ms) return fail(AMM_E_INVALID, "NULL argument");
  if (iters < 1 || iters > 100000) return fail(AMM_E_INVALID, "iters must be in [1, 100000]");
  int rc;
  if ((rc = set_device(c))) return rc;
  if ((rc = enq_stats(c))) return rc;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  for (uint32_t i = 0; i < iters; ++i)
    if ((rc = enq_stats(c))) return rc;
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float t = 0.0f;
  HIPCHK(hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = t / iters;
  c->stats_valid = true;
  return AMM_OK;
}

int amm_time_pass(amm_ctx* c, int32_t pass, uint32_t iters, double* ms) {
  if (!c || !ms) return fail(AMM_E_INVALID, "NULL argument");
  if (iters < 1 || iters > 100000) return fail(AMM_E_INVALID, "iters must be in [1, 100000]");
  if (pass != AMM_PASS_STATS_Z && pass != AMM_PASS_EXACT && pass != AMM_PASS_READ)
    return fail(AMM_E_INVALID, "unknown pass");
  int rc;
  if ((rc = set_device(c))) return rc;
  drop_graph(c);
  if (pass == AMM_PASS_READ) {
    // the matrix's bytes up to the end of its last column's d rows
    const uint64_t es = c->dtype == AMM_DTYPE_BF16 ? 2 : 4;
    const uint64_t bytes = ((c->n - 1) * c->ld + c->d) * es / 16 * 16;
    uint32_t* sink = reinterpret_cast<uint32_t*>(c->BP);
    HIPCHK(amm::launch_read(c->A, bytes, sink, c->stream));  // warm-up
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (uint32_t i = 0; i < iters; ++i) HIPCHK(amm::launch_read(c->A, bytes, sink, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = t / iters;
    return AMM_OK;
  }
  if ((rc = ensure_plans(c, c->list_c))) return rc;
  if (pass == AMM_PASS_STATS_Z) {
    if (!amm::colstats_fuses_z(c->dtype, (uint32_t)c->d))
      return fail(AMM_E_INVALID, "z = A alpha is not fused with the statistics at this d");
    if ((rc = grow(c->ZP, c->ZP_cap,
                   (size_t)amm::colstats_z_blocks(c->dtype, (uint32_t)c->d, c->n) * c->dpad)))
      return rc;
  }
  auto one = [&]() -> int {
    if (pass == AMM_PASS_EXACT) return enq_exact(c, 0, nullptr, 0);
    int r = enq_stats_z(c);
    return r ? r : enq_z_partials(c);
  };
  if ((rc = one())) return rc;  // warm-up (and ||alpha||^2 = 0 is harmless for v)
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  for (uint32_t i = 0; i < iters; ++i)
    if ((rc = one())) return rc;
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float t = 0.0f;
  HIPCHK(hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = t / iters;
  if (pass == AMM_PASS_STATS_Z) c->stats_valid = true;
  return AMM_OK;
}

int amm_time_product(amm_ctx* c, int32_t kernel, uint32_t cc, uint64_t seed, uint32_t iters,
                     double* ms_dot, double* ms_axpy, uint64_t* columns) {
  if (!c) return fail(AMM_E_INVALID, "ctx is NULL");
  if (kernel < AMM_KERNEL_EXACT || kernel > AMM_KERNEL_PROPOSED)
    return fail(AMM_E_INVALID, "kernel must be exact, previousAMM or proposedAMM");
  const bool exact = kernel == AMM_KERNEL_EXACT;
  if (!exact && (cc < 1 || cc > c->n)) return fail(AMM_E_INVALID, "c must satisfy 1 <= c <= n");
  if (iters < 1 || iters > 100000) return fail(AMM_E_INVALID, "iters must be in [1, 100000]");
  int rc;
  if ((rc = set_device(c))) return rc;
  drop_graph(c);
  const bool prop = kernel == AMM_KERNEL_PROPOSED;
  if ((rc = ensure_plans(c, exact ? 0 : cc))) return rc;
  if (!exact && (rc = ensure_lists(c, 1, cc))) return rc;
  if ((rc = ensure_step(c, 1))) return rc;
  if (!c->stats_valid) {
    if ((rc = enq_stats(c))) return rc;
    c->stats_valid = true;
  }
  if ((rc = enq_cdf(c, 1, false))) return rc;
  if (!exact) {
    if (!prop && (rc = enq_cdf(c, 0, false))) return rc;
    if ((rc = upload_run_state(c, seed, 0))) return rc;
    if ((rc = enq_sample(c, prop ? 1 : 0, 0, 1, 0, cc))) return rc;
  }
  const int dtype = c->dtype;
  // the two launches of one product, as enq_exact / enq_amm issue them
  amm::GatherDotArgs g{};
  g.A = c->A;
  g.ld = c->ld;
  g.d = (uint32_t)c->d;
  g.idx = exact ? nullptr : c->idx;
  g.seg_count = exact ? nullptr : c->segc;
  g.cap = exact ? (uint32_t)c->n : cc;
  g.x = c->x;
  g.xm = exact ? nullptr : c->xm;
  g.R = exact ? c->dot_dense.R : c->dot_list.R;
  g.rows_per_split = exact ? c->dot_dense.rows_per_split : c->dot_list.rows_per_split;
  g.P1 = c->P1;
  amm::GatherAxpyArgs a{};
  a.A = c->A;
  a.ld = c->ld;
  a.d = (uint32_t)c->d;
  a.dpad = c->dpad;
  a.idx = g.idx;
  a.wgt = exact ? nullptr : c->wgt;
  a.seg_count = g.seg_count;
  a.cap = g.cap;
  a.cs = exact ? c->axpy_dense.cs : c->axpy_list.cs;
  a.S = exact ? c->axpy_dense.S : c->axpy_list.S;
  a.P1 = c->P1;
  a.R = g.R;
  a.alpha = prop ? c->alpha : nullptr;
  a.st = c->st;
  a.P2 = c->P2;
  a.P2s = c->P2s;
  g.nt = (c->list_nt & 1u) != 0;
  a.nt = (c->list_nt & 2u) != 0;
  // a sampled product's second contraction as a run launches it: with the
  // finish fused in where that form applies
  const bool fused = !exact && fused_finish(c, false);
  amm::AxpyFinishArgs f{};
  if (fused) {
    f.A = c->A;
    f.ld = c->ld;
    f.d = (uint32_t)c->d;
    f.dpad = c->dpad;
    f.idx = c->idx;
    f.wgt = c->wgt;
    f.seg_count = c->segc;
    f.cap = cc;
    f.P1 = c->P1;
    f.R = g.R;
    f.alpha = a.alpha;
    f.st = c->st;
    f.x = c->x;
    f.y = c->y;
    f.ym = c->ym;
    f.v = c->v;
    f.BP = c->BP;
    f.nt = a.nt;
  }
  auto second = [&]() {
    return fused ? amm::launch_axpy_finish(dtype, f, c->stream) : amm::launch_gather_axpy(dtype, a, c->stream);
  };
  HIPCHK(amm::launch_gather_dot(dtype, g, c->stream));
  HIPCHK(second());
  hipEvent_t e2 = nullptr;
  HIPCHK(hipEventCreate(&e2));
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  for (uint32_t i = 0; i < iters; ++i) HIPCHK(amm::launch_gather_dot(dtype, g, c->stream));
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  for (uint32_t i = 0; i < iters; ++i) HIPCHK(second());
  HIPCHK(hipEventRecord(e2, c->stream));
  const hipError_t se = hipEventSynchronize(e2);
  float t1 = 0.0f, t2 = 0.0f;
  hipError_t q1 = hipEventElapsedTime(&t1, c->ev0, c->ev1);
  hipError_t q2 = hipEventElapsedTime(&t2, c->ev1, e2);
  (void)hipEventDestroy(e2);
  if (se != hipSuccess || q1 != hipSuccess || q2 != hipSuccess)
    return fail(AMM_E_HIP, "event timing failed");
  if (ms_dot) *ms_dot = t1 / iters;
  if (ms_axpy) *ms_axpy = t2 / iters;
  if (columns) {
    if (exact) {
      *columns = c->n;
    } else {
      const size_t nseg = (cc + amm::kSampleSeg - 1) / amm::kSampleSeg;
      std::vector<uint32_t> sc(nseg);
      HIPCHK(copy_sync(c->stream, sc.data(), c->segc, sizeof(uint32_t) * nseg, hipMemcpyDeviceToHost));
      uint64_t u = 0;
      for (uint32_t v : sc) u += v;
      *columns = u;
    }
  }
  return check_status(c);
}

}  // extern "C"
