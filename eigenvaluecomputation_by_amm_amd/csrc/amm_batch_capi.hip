// amm_batch_capi.hip -- C ABI of the batched AMM power method
// (include/amm.h, amm_batch_*): many matrices of uneven sizes, up to 16
// trials each, every launch grouped over all problems and the whole run --
// column statistics, CDFs, z = A alpha, x0 and all iterations -- captured
// in one hipGraph.  Tile shapes come from a launch-planning heuristic or
// from amm_batch_autotune, which times candidate shapes on the live
// problems.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "amm.h"
#include "amm_batch.hpp"
#include "amm_device.hpp"
#include "amm_launch.hpp"
#include "amm_setup.hpp"

namespace amm_host {
int fail(int code, const std::string& msg);
std::recursive_mutex& device_mutex();  // amm_capi.hip: capture vs. allocation
}

using amm::BProb;
using amm::BTrial;
using amm::DevState;
using amm_host::fail;

#define BHIPCHK(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(AMM_E_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

// Host copies run on the object's own stream and then wait for it.  A plain
// hipMemcpy runs on the legacy null stream, which may not take a dependency
// on a blocking stream that another thread of this process is capturing into
// a graph; the failed copy would then abort that capture too.
static hipError_t copy_sync(hipStream_t s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}


namespace {

constexpr uint32_t kTargetItems = 2048;      // workgroups the first contraction aims for
constexpr uint32_t kTargetAxpyItems = 512;   // and the second: fewer chunks, fewer partials
constexpr uint32_t kWeightBlocks = 4096;     // weight workgroups at most (about twice resident)
constexpr uint32_t kFinishSlabs = 4096;      // finish workgroups of 256 threads at most
constexpr uint32_t kMaxSlices = 16;      // contraction-1 row slices per problem
constexpr uint32_t kZChunk = 2048;       // columns per chunk of z = A alpha
constexpr uint64_t kNtMinBytes = 512ull << 20;
constexpr uint64_t kUnionFusedBytes = 256u << 10;
constexpr uint32_t kGramItems = 1024;  // k_bgram workgroups to aim for (a share per problem by d x n)
// columns shorter than this keep the two contractions: a k_bgram workgroup
// has one column in flight, too little to stream short ones
constexpr uint64_t kGramMinColBytes = 8192;  // count rows per problem up to which one workgroup forms its union  // matrix bytes beyond which the contractions stream A non-temporally

uint64_t rup(uint64_t a, uint64_t b) { return (a + b - 1) / b * b; }
uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// One device allocation carved into 256-byte aligned pieces.
struct Pool {
  size_t bytes = 0;
  template <typename T>
  size_t take(size_t count) {
    const size_t o = bytes;
    bytes = rup(bytes + std::max<size_t>(count, 1) * sizeof(T), 256);
    return o;
  }
};

struct Tuning {
  uint32_t dot_blocks, passes, chunks;
};

// per-problem buffers of the single-matrix setup kernels
struct SetupBufs {
  float* alpha;
  double* nrm2;
  double* cnrm2;
  uint64_t* cdf0;
  uint64_t* cdf1;
  float* z;
  float* v;
  DevState* st;
  double* eigvec;
  uint64_t dpad8;
};

// rows of one contraction-2 row block
uint32_t block_rows(uint32_t passes) { return 256u * passes; }

}  // namespace

struct amm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  uint32_t np = 0, ntr = 0;
  std::vector<amm_problem> probs;
  std::vector<BProb> hp;
  std::vector<SetupBufs> sb;
  std::vector<Tuning> tune;
  BProb* dP = nullptr;
  char* spool = nullptr;  // static per-problem buffers
  char* tpool = nullptr;  // tuning-dependent partials (P1, P2, P2s)
  size_t tpool_bytes = 0;
  char* rpool = nullptr;  // per-run outputs (eig, err, modes, ts, schedule)
  size_t rpool_bytes = 0;
  double* dsig = nullptr;  // sigma_max outputs [np][ntr] and u, v pointer table [2 np]
  float** dvec = nullptr;
  uint64_t rstep = 0;
  // shared setup scratch
  uint64_t* tile_max = nullptr;
  uint64_t* tile_sum = nullptr;
  uint64_t* tile_lp = nullptr;
  double* tile_a2 = nullptr;
  float* zP2 = nullptr;
  double* zP2s = nullptr;
  // the contractions read A non-temporally when every problem has d >= 4096
  // rows (measured: +4-6% on 16384 x 65536 and 64 x 4096 x 4096, -4% on a
  // mix of small matrices, whose second contraction finds columns in cache)
  bool nt_loads = false;
  bool union_fused = false;  // the union pass as one workgroup per problem (k_bunion)
  bool gram = false;         // one-trial batch: products in one read of A (k_bgram)
  bool gram_all = false;     // every problem's products are k_bgram's (no contraction launches)
  int gram_Ls[2] = {0, 0};   // k_bgram's loads per lane for the FP32 / BF16 group
  amm::BGroup g_gram[2]{};
  // grouped setup: descriptors (device) and their launch groups
  amm::BSetup* dS = nullptr;
  std::vector<uint32_t> tiles;  // CDF tiles per problem
  amm::BGroup g_cols{}, g_tiles{}, g_zv{};
  // grouped-launch item offsets
  uint32_t* offs = nullptr;
  size_t offs_cap = 0;
  amm::BGroup g_vec{}, g_norm{}, g_fin{}, g_sample{}, g_prep{}, g_weight{}, g_dot[2]{}, g_axpy[2][2]{};
  uint32_t fin_groups = 1;  // wave groups of the finish (4 where a problem has many chunks)
  bool plan_valid = false;
  int plan_kernel = -1;
  // run state
  uint32_t* modes_static = nullptr;
  int64_t* ts = nullptr;
  int64_t* t_start = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_valid = false;
  bool graph_setup = false;
  amm_params key{};
  bool stats_valid = false;
  int stats_kernel = -1;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int wall_khz = 100000;
  uint64_t last_step = 0;
  bool ran = false;
};

namespace {

void drop_graph(amm_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_valid = false;
}

// Default shapes: each problem gets a share of the workgroups of both
// contractions (kTargetItems, kTargetAxpyItems) proportional to its columns
// x rows; the slices of contraction 1 and the chunk width of contraction 2
// then follow each product's union size on the device (union_plan).  The
// second contraction's target is the lower one: every chunk adds a partial
// of 16 trials x 4 bytes per row that the finish pass reads back, and
// about two workgroups per CU already keep HBM busy (measured on 64 x
// 4096 x 4096 and 4096 x 40000 BF16 runs, 8-16% faster than 2048).
Tuning default_tuning_one(const BProb& p, double share, double ks_mult = 1.0, double ch_mult = 1.0,
                          uint32_t passes = 0) {
  const double want = std::max(1.0, kTargetItems * share);
  const double want_ax = std::max(1.0, kTargetAxpyItems * share);
  if (!passes) passes = p.d >= 8192 ? 2u : 1u;
  const uint32_t rblocks = (uint32_t)cdiv(p.d, block_rows(passes));
  const uint32_t dot_blocks = (uint32_t)std::max(1.0, std::round(want * ks_mult));
  const double ch = std::max(1.0, std::round(want_ax / rblocks * ch_mult));
  const uint32_t chunks = (uint32_t)std::min<double>((double)cdiv(p.n, 64), ch);
  return Tuning{dot_blocks, passes, chunks};
}

void default_tuning(amm_batch* b) {
  double tot = 0.0;
  for (const BProb& p : b->hp) tot += (double)p.ntiles * (double)p.d;
  for (uint32_t i = 0; i < b->np; ++i) {
    const BProb& p = b->hp[i];
    b->tune[i] = default_tuning_one(p, tot > 0 ? (double)p.ntiles * (double)p.d / tot : 1.0);
  }
}

int check_tuning(const BProb& p, const Tuning& t) {
  if (t.passes < 1 || t.passes > 2) return fail(AMM_E_INVALID, "passes must be 1 or 2");
  if (t.dot_blocks < 1) return fail(AMM_E_INVALID, "dot_blocks must be >= 1");
  if (t.chunks < 1 || t.chunks > cdiv(p.n, 64))
    return fail(AMM_E_INVALID, "chunks must be in [1, ceil(n / 64)]");
  return AMM_OK;
}

// Descriptors from the tuning, (re)allocation of the partial buffers and
// the grouped-launch offsets.
int upload_plan(amm_batch* b, int kernel) {
  std::lock_guard<std::recursive_mutex> lk(amm_host::device_mutex());
  const uint32_t np = b->np;
  BHIPCHK(hipStreamSynchronize(b->stream));
  Pool pl;
  std::vector<size_t> o1(np), o2(np), o2s(np);
  // A one-trial batch takes every product of its problems with columns of
  // kGramMinColBytes or more that fit k_bgram's registers in one read of A:
  // each such problem gets a share of kGramItems workgroups by its d x n,
  // each a contiguous run of union slots whose partial k_bfinish adds up
  // (like the second contraction's chunks).  The other problems keep the two
  // contractions.
  auto gram_ok = [&](const BProb& p) {
    return b->ntr == 1 && amm::gram_L(p.dtype, p.d) != 0 &&
           (uint64_t)p.d * (p.dtype == 1 ? 2u : 4u) >= kGramMinColBytes;
  };
  b->gram = false;
  b->gram_Ls[0] = b->gram_Ls[1] = 0;
  double gram_tot = 0.0;
  for (const BProb& p : b->hp) {
    if (!gram_ok(p)) continue;
    b->gram = true;
    b->gram_Ls[p.dtype == 1 ? 1 : 0] = std::max(b->gram_Ls[p.dtype == 1 ? 1 : 0], amm::gram_L(p.dtype, p.d));
    gram_tot += (double)p.n * (double)p.d;
  }
  b->gram_all = b->gram;
  for (BProb& p : b->hp) {
    const double want = std::round(kGramItems * (double)p.n * (double)p.d / std::max(gram_tot, 1.0));
    p.gram_chunks = gram_ok(p) ? (uint32_t)std::min<double>((double)p.n, std::max(1.0, want)) : 0u;
    b->gram_all = b->gram_all && p.gram_chunks != 0;
  }
  for (uint32_t i = 0; i < np; ++i) {
    BProb& p = b->hp[i];
    const Tuning& t = b->tune[i];
    p.kslices_max = (uint32_t)std::min<uint64_t>(kMaxSlices, p.dstage / 256);
    p.dot_blocks = t.dot_blocks;
    p.passes = t.passes;
    p.rblocks = (uint32_t)cdiv(p.d, block_rows(t.passes));
    p.p2pitch = p.rblocks * block_rows(t.passes);
    p.chunks = t.chunks;
    o1[i] = pl.take<float>((size_t)p.kslices_max * p.ntiles * amm::kBTile * amm::kBT);
    o2[i] = pl.take<float>((size_t)std::max(p.chunks, p.gram_chunks) * amm::kBT * p.p2pitch);
    o2s[i] = pl.take<double>((size_t)std::max(p.chunks, p.gram_chunks) * amm::kBT);
  }
  if (pl.bytes > b->tpool_bytes) {
    if (b->tpool) (void)hipFree(b->tpool);
    b->tpool = nullptr;
    b->tpool_bytes = 0;
    if (hipMalloc((void**)&b->tpool, pl.bytes) != hipSuccess)
      return fail(AMM_E_NOMEM, "device allocation of " + std::to_string(pl.bytes) + " bytes failed");
    BHIPCHK(hipMemset(b->tpool, 0, pl.bytes));
    b->tpool_bytes = pl.bytes;
  }
  for (uint32_t i = 0; i < np; ++i) {
    b->hp[i].P1 = reinterpret_cast<float*>(b->tpool + o1[i]);
    b->hp[i].P2 = reinterpret_cast<float*>(b->tpool + o2[i]);
    b->hp[i].P2s = reinterpret_cast<double*>(b->tpool + o2s[i]);
  }
  std::vector<uint32_t> all;
  std::vector<std::pair<amm::BGroup*, size_t>> fix;
  // a group: offsets off[0 .. np] of each problem's first workgroup, then
  // the problem of every workgroup (find_problem reads it in one load).
  // With `order`, the problems' workgroups follow each other in that order
  // (off[p] is still problem p's first workgroup, off[np] the total); only
  // kernels that read off[p] alone, not off[p + 1] - off[p], take such a group.
  auto add_group = [&](amm::BGroup& g, auto count_fn, const std::vector<uint32_t>* order = nullptr) {
    const size_t base = all.size();
    all.resize(base + np + 1);
    uint64_t acc = 0;
    for (uint32_t k = 0; k < np; ++k) {
      const uint32_t i = order ? (*order)[k] : k;
      all[base + i] = (uint32_t)std::min<uint64_t>(acc, 0xFFFFFFFFu);
      acc += count_fn(b->hp[i]);
    }
    if (acc >= (1ull << 31)) return false;
    all[base + np] = (uint32_t)acc;
    for (uint32_t k = 0; k < np; ++k) {
      const uint32_t i = order ? (*order)[k] : k;
      all.insert(all.end(), (size_t)count_fn(b->hp[i]), i);
    }
    g.total = (uint32_t)acc;
    fix.emplace_back(&g, base);
    return true;
  };
  // Contraction workgroups are long-lived and their sizes differ by problem
  // (a contraction-2 workgroup of a problem with one chunk reads all n
  // columns of its row block): dispatched in problem order, the largest
  // could start last and run on alone at the end of the launch.  Both
  // contraction groups therefore go largest workgroup first, sized at the
  // dense union (every column present).
  auto largest_first = [&](auto bytes_fn) {
    std::vector<uint32_t> ord(np);
    for (uint32_t i = 0; i < np; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(),
                     [&](uint32_t x, uint32_t y) { return bytes_fn(b->hp[x]) > bytes_fn(b->hp[y]); });
    return ord;
  };
  const std::vector<uint32_t> dot_order = largest_first([](const BProb& p) {
    const amm::UPlan up = amm::union_plan(p, p.n);
    return (uint64_t)amm::kBTile * up.kslice_rows * (p.dtype == 1 ? 2u : 4u);
  });
  const std::vector<uint32_t> gram_order = largest_first([](const BProb& p) {
    return p.gram_chunks ? (uint64_t)p.n * p.d * (p.dtype == 1 ? 2u : 4u) / p.gram_chunks : 0u;
  });
  const std::vector<uint32_t> axpy_order = largest_first([](const BProb& p) {
    const amm::UPlan up = amm::union_plan(p, p.n);
    return (uint64_t)std::min(up.cpc, p.n) * block_rows(p.passes) * (p.dtype == 1 ? 2u : 4u);
  });
  bool ok = true;
  // statistics: a workgroup per column, or per four short columns (a wave each)
  ok &= add_group(b->g_cols, [](const BProb& p) {
    return amm::short_column_dt(p.dtype, p.d) ? (p.n + 3) / 4 : p.n;
  });
  ok &= add_group(b->g_tiles, [](const BProb& p) { return (uint32_t)amm::cdf_tiles(p.n); });
  ok &= add_group(b->g_zv, [](const BProb& p) { return (uint32_t)cdiv(rup(p.d, 8), 256); });
  ok &= add_group(b->g_vec, [](const BProb& p) { return p.nvblk; });
  ok &= add_group(b->g_norm, [&](const BProb& p) { return b->ntr == 1 ? (p.nvblk + 15) / 16 : p.nvblk; });
  // Sampler: a workgroup per (problem, trial), drawing all c slots
  ok &= add_group(b->g_sample, [&](const BProb&) { return kernel == AMM_KERNEL_EXACT ? 0u : b->ntr; });
  // Compaction: a wave per tile; weights: a lane quad per column (the
  // union is at most n slots, a workgroup takes a tile of 64 per pass).  Across many problems both counts shrink by a
  // common factor (each wave / thread then loops over several items): a
  // workgroup's start is a chain of dependent loads (problem, descriptor,
  // union size) that few items per workgroup would not amortise.
  uint64_t tot_tiles = 0;
  for (const BProb& p : b->hp) tot_tiles += p.ntiles;
  const uint64_t f_prep = std::max<uint64_t>(1, cdiv(tot_tiles, 16384));  // <= 4096 workgroups
  const uint64_t f_weight = std::max<uint64_t>(1, cdiv(tot_tiles, kWeightBlocks));
  ok &= add_group(b->g_prep, [f_prep](const BProb& p) { return (uint32_t)cdiv(p.ntiles, 4 * f_prep); });
  // Finish: a workgroup per 64-row slab on small batches; across many
  // problems workgroups take several slabs (about twice resident in all).
  // Four wave groups per workgroup split the chunk partials where some
  // problem has more than 8 chunks.
  uint64_t tot_vblk = 0;
  uint32_t max_chunks = 0;
  for (const BProb& p : b->hp) {
    tot_vblk += p.nvblk;
    max_chunks = std::max(max_chunks, std::max(p.chunks, p.gram_chunks));
  }
  b->fin_groups = max_chunks > 8 ? 4u : 1u;
  // one trial: a workgroup pass takes 16 slabs (its 16 lane groups)
  const uint64_t spw = b->ntr == 1 ? 16u : 1u;
  const uint64_t f_fin = std::max<uint64_t>(1, cdiv(cdiv(tot_vblk, spw), kFinishSlabs / b->fin_groups));
  ok &= add_group(b->g_fin, [f_fin, spw](const BProb& p) { return (uint32_t)cdiv(cdiv(p.nvblk, spw), f_fin); });
  ok &= add_group(b->g_weight, [f_weight](const BProb& p) { return (uint32_t)cdiv(p.ntiles, f_weight); });
  for (int dt = 0; dt < 2; ++dt)
    ok &= add_group(
        b->g_dot[dt],
        [dt](const BProb& p) {
          // vtiles * kslices(|U|) < dot_blocks + vtiles for every union size
          return p.dtype == dt ? std::min(p.ntiles * p.kslices_max, p.dot_blocks + p.ntiles) : 0u;
        },
        &dot_order);
  for (int dt = 0; dt < 2; ++dt)
    for (int pa = 0; pa < 2; ++pa)
      ok &= add_group(
          b->g_axpy[dt][pa],
          [dt, pa](const BProb& p) {
            return (p.dtype == dt && p.passes == (uint32_t)pa + 1) ? p.rblocks * p.chunks : 0u;
          },
          &axpy_order);
  for (int dt = 0; dt < 2; ++dt)
    ok &= add_group(
        b->g_gram[dt], [dt](const BProb& p) { return p.dtype == dt ? p.gram_chunks : 0u; }, &gram_order);
  if (!ok) return fail(AMM_E_INVALID, "a grouped launch exceeds 2^31 workgroups");
  if (all.size() > b->offs_cap) {
    if (b->offs) (void)hipFree(b->offs);
    b->offs = nullptr;
    b->offs_cap = 0;
    BHIPCHK(hipMalloc((void**)&b->offs, all.size() * sizeof(uint32_t)));
    b->offs_cap = all.size();
  }
  BHIPCHK(copy_sync(b->stream, b->offs, all.data(), all.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  for (auto& f : fix) f.first->off = b->offs + f.second;
  BHIPCHK(copy_sync(b->stream, b->dP, b->hp.data(), np * sizeof(BProb), hipMemcpyHostToDevice));
  b->plan_valid = true;
  b->plan_kernel = kernel;
  drop_graph(b);
  return AMM_OK;
}

// per-run outputs: eig, err, modes [16][step] per problem; ts; schedule
int layout_run(amm_batch* b, uint64_t step) {
  if (b->rpool && b->rstep == step) return AMM_OK;
  std::lock_guard<std::recursive_mutex> lk(amm_host::device_mutex());
  Pool pl;
  std::vector<size_t> oe(b->np), oerr(b->np), om(b->np);
  for (uint32_t i = 0; i < b->np; ++i) {
    oe[i] = pl.take<double>(amm::kBT * step);
    oerr[i] = pl.take<double>(amm::kBT * step);
    om[i] = pl.take<uint32_t>(amm::kBT * step);
  }
  const size_t ots = pl.take<int64_t>(step);
  const size_t ot0 = pl.take<int64_t>(1);
  const size_t oms = pl.take<uint32_t>(step);
  if (b->rpool) (void)hipFree(b->rpool);
  b->rpool = nullptr;
  b->rstep = 0;
  if (hipMalloc((void**)&b->rpool, pl.bytes) != hipSuccess)
    return fail(AMM_E_NOMEM, "device allocation of " + std::to_string(pl.bytes) + " bytes failed");
  BHIPCHK(hipMemset(b->rpool, 0, pl.bytes));
  for (uint32_t i = 0; i < b->np; ++i) {
    b->hp[i].eig = reinterpret_cast<double*>(b->rpool + oe[i]);
    b->hp[i].err = reinterpret_cast<double*>(b->rpool + oerr[i]);
    b->hp[i].modes = reinterpret_cast<uint32_t*>(b->rpool + om[i]);
  }
  b->ts = reinterpret_cast<int64_t*>(b->rpool + ots);
  b->t_start = reinterpret_cast<int64_t*>(b->rpool + ot0);
  b->modes_static = reinterpret_cast<uint32_t*>(b->rpool + oms);
  b->rstep = step;
  b->rpool_bytes = pl.bytes;
  b->plan_valid = false;  // descriptors changed
  drop_graph(b);
  return AMM_OK;
}

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

amm::BRun run_args(amm_batch* b, const amm_params& p) {
  amm::BRun r{};
  r.np = b->np;
  r.ntr = b->ntr;
  r.trial0 = p.trial;
  r.kernel = (uint32_t)p.kernel;
  r.step = (uint32_t)p.step;
  r.adaptive = is_adaptive(p) ? 1u : 0u;
  r.period = p.vr_period;
  r.thr = p.vr_adaptive_threshold;
  r.modes_static = b->modes_static;
  r.ts = b->ts;
  r.t_start = b->t_start;
  r.gram = b->gram ? 1u : 0u;
  return r;
}

// Column statistics, CDFs and z = A alpha of every problem.  Statistics
// and CDFs are grouped launches over all problems.  For the centred
// kernels, z rides in each problem's own statistics pass where the columns
// fit it (its partials summed before the next problem's pass reuses the
// buffer; v = z - ||alpha||^2 once the grouped centred CDF has set
// ||alpha||^2), else z comes from a separate pass after the CDF.
int enq_setup(amm_batch* b, int K) {
  const bool need_plain = K == AMM_KERNEL_PREVIOUS || K == AMM_KERNEL_PROPOSED_V2;
  const bool need_cent = K == AMM_KERNEL_PROPOSED || K == AMM_KERNEL_PROPOSED_V2;
  bool any_plain = !need_cent;  // with z fused, only the problems it cannot fuse need this pass
  for (uint32_t i = 0; i < b->np && !any_plain; ++i)
    any_plain = !amm::colstats_fuses_z(b->hp[i].dtype, b->hp[i].d);
  if (any_plain) BHIPCHK(amm::launch_bcolstats(b->dS, b->g_cols, b->np, need_cent ? 1u : 0u, b->stream));
  if (need_cent) {
    for (uint32_t i = 0; i < b->np; ++i) {
      const BProb& p = b->hp[i];
      const SetupBufs& s = b->sb[i];
      if (!amm::colstats_fuses_z(p.dtype, p.d)) continue;
      BHIPCHK(amm::launch_colstats_z(p.dtype, p.A, p.ld, p.d, p.n, s.dpad8, s.alpha, s.nrm2, s.cnrm2,
                                     b->zP2, s.st, b->stream));
      BHIPCHK(amm::launch_sum_rows(b->zP2, amm::colstats_z_blocks(p.dtype, p.d, p.n), s.dpad8, p.d, s.z,
                                   nullptr, s.st, b->stream));
    }
  }
  if (need_plain) BHIPCHK(amm::launch_bcdf(b->dS, b->g_tiles, b->np, 0, b->stream));
  if (need_cent) {
    BHIPCHK(amm::launch_bcdf(b->dS, b->g_tiles, b->np, 1, b->stream));
    BHIPCHK(amm::launch_bzv(b->dS, b->g_zv, b->np, b->stream));
    for (uint32_t i = 0; i < b->np; ++i) {
      const BProb& p = b->hp[i];
      const SetupBufs& s = b->sb[i];
      if (amm::colstats_fuses_z(p.dtype, p.d)) continue;
      amm::GatherAxpyArgs a{};
      a.A = p.A;
      a.ld = p.ld;
      a.d = p.d;
      a.dpad = s.dpad8;
      a.cap = p.n;
      a.cs = kZChunk;
      a.S = (uint32_t)cdiv(p.n, kZChunk);
      a.tdirect = s.alpha;
      a.st = s.st;
      a.P2 = b->zP2;
      a.P2s = b->zP2s;
      BHIPCHK(amm::launch_gather_axpy(p.dtype, a, b->stream));
      BHIPCHK(amm::launch_sum_rows(b->zP2, a.S, s.dpad8, p.d, s.z, s.v, s.st, b->stream));
    }
  }
  return AMM_OK;
}

// the iterations' kernels (after x0)
// kind: the iteration's step kind where the schedule is static, else 0
int enq_iteration(amm_batch* b, const amm::BRun& r, int K, uint32_t it, uint32_t kind) {
  const BProb* P = b->dP;
  hipStream_t s = b->stream;
  const bool exact_kind = kind == amm::kModeExactMemo || kind == amm::kModeExactNoMemo;
  if (K != AMM_KERNEL_EXACT) BHIPCHK(amm::launch_bsample(P, b->g_sample, r, it, s));
  BHIPCHK(amm::launch_bunion(P, b->g_prep, r, b->union_fused, s));
  const uint32_t prop = (K == AMM_KERNEL_PROPOSED || K == AMM_KERNEL_PROPOSED_V2) ? 1u : 0u;
  // a one-trial batch's products are k_bgram's where its columns allow (the
  // contractions and the weight pass leave those problems alone)
  if (b->gram)
    for (int dt = 0; dt < 2; ++dt)
      BHIPCHK(amm::launch_bgram(dt, b->gram_Ls[dt], exact_kind, P, b->g_gram[dt], b->np, s));
  if (!b->gram_all) {
    for (int dt = 0; dt < 2; ++dt) BHIPCHK(amm::launch_bdot(dt, b->nt_loads, P, b->g_dot[dt], b->np, s, b->gram));
    BHIPCHK(amm::launch_bweight(P, b->g_weight, r, prop, 0u, s, b->gram));
    for (int dt = 0; dt < 2; ++dt)
      for (int pa = 0; pa < 2; ++pa)
        BHIPCHK(amm::launch_baxpy(dt, (uint32_t)pa + 1, b->nt_loads, P, b->g_axpy[dt][pa], b->np, prop, s, b->gram));
  }
  BHIPCHK(amm::launch_bfinish(P, b->g_fin, r, b->fin_groups, s));
  BHIPCHK(amm::launch_btail(P, r, amm::kBTailFinish, it, s));
  BHIPCHK(amm::launch_bnorm(P, b->g_norm, r, s));
  return AMM_OK;
}

int enqueue_run(amm_batch* b, const amm_params& p, bool setup) {
  int rc;
  const int K = p.kernel;
  const amm::BRun r = run_args(b, p);
  BHIPCHK(amm::launch_bbegin(b->dP, b->g_vec, r, b->stream));
  if (setup)
    if ((rc = enq_setup(b, K))) return rc;
  // x0 and its block sums came with the prologue (the setup leaves Y and BP alone)
  BHIPCHK(amm::launch_btail(b->dP, r, amm::kBTailScale, 0, b->stream));
  BHIPCHK(amm::launch_bnorm(b->dP, b->g_norm, r, b->stream));
  const std::vector<uint32_t> kinds = is_adaptive(p) ? std::vector<uint32_t>(p.step, 0u) : static_modes(p);
  for (uint32_t it = 0; it < (uint32_t)p.step; ++it)
    if ((rc = enq_iteration(b, r, K, it, kinds[it]))) return rc;
  return AMM_OK;
}

bool same_key(const amm_params& a, const amm_params& b) {
  return a.kernel == b.kernel && a.step == b.step && a.vr_period == b.vr_period &&
         a.vr_adaptive_threshold == b.vr_adaptive_threshold && a.trial == b.trial &&
         a.flags == b.flags;
}

int validate(const amm_batch* b, const amm_params* p) {
  if (!p) return fail(AMM_E_INVALID, "params is NULL");
  if (p->kernel < AMM_KERNEL_EXACT || p->kernel > AMM_KERNEL_PROPOSED_V2)
    return fail(AMM_E_INVALID, "unknown kernel " + std::to_string(p->kernel));
  if (p->step < 1 || p->step > (1ull << 24)) return fail(AMM_E_INVALID, "step must be in [1, 2^24]");
  if (p->sanity) return fail(AMM_E_INVALID, "sanity mode is a single-matrix debug mode");
  if (p->kernel == AMM_KERNEL_PROPOSED_V2 && (p->vr_period < 1 || p->vr_period > p->step))
    return fail(AMM_E_INVALID,
                "proposedAMM_v2 needs 1 <= vr_period <= step (powerMethod_by_proposed_v2.cuh:234)");
  if (!(p->vr_adaptive_threshold >= 0.0)) return fail(AMM_E_INVALID, "vr_adaptive_threshold < 0");
  if (p->kernel != AMM_KERNEL_EXACT)
    for (const BProb& q : b->hp)
      if (q.c < 1 || q.c > q.n)
        return fail(AMM_E_INVALID, "every problem needs 1 <= c <= n (host_memory.h:166-167)");
  return AMM_OK;
}

int check_status(amm_batch* b) {
  std::vector<DevState> sts(b->np);
  BHIPCHK(copy_sync(b->stream, sts.data(), b->sb[0].st, b->np * sizeof(DevState), hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < b->np; ++i) {
    const uint32_t stt = sts[i].status;
    if (stt & amm::kStatusNonFinite)
      return fail(AMM_E_INVALID, "problem " + std::to_string(i) +
                                     ": the matrix has non-finite entries (NaN or Inf column norms)");
  }
  return AMM_OK;
}

// Prepare descriptors, schedule and graph for parameters p (no launch).
int prepare(amm_batch* b, const amm_params* p) {
  int rc;
  if ((rc = layout_run(b, p->step))) return rc;
  if (!b->plan_valid || b->plan_kernel != p->kernel)
    if ((rc = upload_plan(b, p->kernel))) return rc;
  const std::vector<uint32_t> modes = static_modes(*p);
  BHIPCHK(hipMemcpyAsync(b->modes_static, modes.data(), modes.size() * sizeof(uint32_t),
                         hipMemcpyHostToDevice, b->stream));
  return AMM_OK;
}

int run(amm_batch* b, const amm_params* p) {
  int rc;
  BHIPCHK(hipSetDevice(b->device));
  if ((rc = validate(b, p))) return rc;
  if ((rc = prepare(b, p))) return rc;
  const bool use_graph = !(p->flags & AMM_FLAG_NO_GRAPH);
  // the setup kernels are part of the run unless statistics are kept
  const bool setup_included =
      p->kernel != AMM_KERNEL_EXACT &&
      (!(p->flags & AMM_FLAG_KEEP_STATS) || !b->stats_valid || b->stats_kernel != p->kernel);
  if (use_graph) {
    if (!b->graph_valid || !same_key(b->key, *p) || b->graph_setup != setup_included) {
    std::lock_guard<std::recursive_mutex> lk(amm_host::device_mutex());
      drop_graph(b);
      BHIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      rc = enqueue_run(b, *p, setup_included);
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(b->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (ce != hipSuccess) return fail(AMM_E_HIP, std::string("graph capture failed: ") + hipGetErrorString(ce));
      b->graph = g;
      BHIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_valid = true;
      b->graph_setup = setup_included;
      b->key = *p;
    }
    BHIPCHK(hipEventRecord(b->ev0, b->stream));
    BHIPCHK(hipGraphLaunch(b->exec, b->stream));
    BHIPCHK(hipEventRecord(b->ev1, b->stream));
  } else {
    BHIPCHK(hipEventRecord(b->ev0, b->stream));
    if ((rc = enqueue_run(b, *p, setup_included))) return rc;
    BHIPCHK(hipEventRecord(b->ev1, b->stream));
  }
  BHIPCHK(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  BHIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  if (p->kernel != AMM_KERNEL_EXACT) {
    if ((rc = check_status(b))) {
      b->stats_valid = false;
      return rc;
    }
    if (setup_included) {
      b->stats_valid = true;
      b->stats_kernel = p->kernel;
    }
  }
  b->last_step = p->step;
  b->ran = true;
  return AMM_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

int amm_batch_create(const amm_problem* probs, uint32_t count, uint32_t trials, void* stream,
                     amm_batch** out) {
  if (!out) return fail(AMM_E_INVALID, "out is NULL");
  *out = nullptr;
  if (!probs || count < 1) return fail(AMM_E_INVALID, "need at least one problem");
  if (trials < 1 || trials > amm::kBT) return fail(AMM_E_INVALID, "trials must be in [1, 16]");
  for (uint32_t i = 0; i < count; ++i) {
    const amm_problem& q = probs[i];
    const std::string w = "problem " + std::to_string(i) + ": ";
    if (!q.A) return fail(AMM_E_INVALID, w + "A is NULL");
    if (q.dtype != AMM_DTYPE_F32 && q.dtype != AMM_DTYPE_BF16) return fail(AMM_E_INVALID, w + "unknown dtype");
    if (q.d < 1 || q.n < 1 || q.d >= (1ull << 31) || q.n >= (1ull << 26))
      return fail(AMM_E_INVALID, w + "need 1 <= d < 2^31 and 1 <= n < 2^26");
    if (q.ld < q.d) return fail(AMM_E_INVALID, w + "ld must be >= d");
    const uint64_t V = q.dtype == AMM_DTYPE_BF16 ? 8 : 4;
    if (q.ld % V) return fail(AMM_E_INVALID, w + "ld must be a multiple of 16 bytes of elements");
    if (((uintptr_t)q.A) & 15u) return fail(AMM_E_INVALID, w + "A must be 16-byte aligned");
    if (q.c > q.n) return fail(AMM_E_INVALID, w + "c must be <= n");
    if (q.c > amm::kMaxBatchC)
      return fail(AMM_E_INVALID, w + "c must be <= 65535 (16-bit draw counts of the batched engine)");
  }
  std::lock_guard<std::recursive_mutex> lk(amm_host::device_mutex());
  amm_batch* b = new amm_batch();
  auto bail = [&](int r) {
    amm_batch_destroy(b);
    return r;
  };
  b->np = count;
  b->ntr = trials;
  b->probs.assign(probs, probs + count);
  if (hipGetDevice(&b->device) != hipSuccess) return bail(fail(AMM_E_HIP, "hipGetDevice failed"));
  if (stream) {
    b->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&b->stream, hipStreamDefault) != hipSuccess)
      return bail(fail(AMM_E_HIP, "hipStreamCreate failed"));
    b->own_stream = true;
  }
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device) == hipSuccess && khz > 0)
    b->wall_khz = khz;
  // static per-problem layout
  Pool pl;
  struct Off {
    size_t X, Y, XM, YM, DX, DXh, DXl, cnt, list, tcount, tpre, colc, Tc, BP, BQ, bt, alpha, nrm2,
        cnrm2, cdf0, cdf1, z, v, eigvec;
  };
  std::vector<Off> o(count);
  b->hp.resize(count);
  b->sb.resize(count);
  b->tune.resize(count);
  uint64_t max_n = 0, max_zs = 0;
  // the problems' device states side by side: one copy reads every status
  const size_t st_all = pl.take<DevState>(count);
  for (uint32_t i = 0; i < count; ++i) {
    const amm_problem& q = probs[i];
    BProb& p = b->hp[i];
    p = BProb{};
    p.A = q.A;
    p.ld = q.ld;
    p.seed = q.seed;
    p.d = (uint32_t)q.d;
    p.n = (uint32_t)q.n;
    p.dstage = (uint32_t)rup(q.d, amm::kBStageAlign);
    p.c = q.c;
    p.dtype = q.dtype;
    p.ntiles = (uint32_t)cdiv(q.n, amm::kBTile);
    p.nvblk = p.dstage / amm::kBVecRows;
    const size_t vec = (size_t)amm::kBT * p.dstage;
    const size_t cols = (size_t)p.ntiles * amm::kBTile;
    const uint64_t dpad8 = rup(q.d, 8);
    Off& f = o[i];
    f.X = pl.take<float>(vec);
    f.Y = pl.take<float>(vec);
    f.XM = pl.take<float>(vec);
    f.YM = pl.take<float>(vec);
    f.DX = pl.take<float>(q.dtype == AMM_DTYPE_F32 ? vec : 1);
    f.DXh = pl.take<uint16_t>(q.dtype == AMM_DTYPE_BF16 ? vec : 1);
    f.DXl = pl.take<uint16_t>(q.dtype == AMM_DTYPE_BF16 ? vec : 1);
    f.cnt = pl.take<uint16_t>(cols * amm::kBT);
    f.list = pl.take<uint32_t>(cols);
    f.tcount = pl.take<uint32_t>(p.ntiles);
    f.tpre = pl.take<uint32_t>(p.ntiles + 1);
    f.colc = pl.take<uint32_t>(q.n);
    f.Tc = pl.take<float>(q.n * amm::kBT);
    f.BP = pl.take<double>((size_t)p.nvblk * amm::kBT * 3);
    f.BQ = pl.take<double>((size_t)p.nvblk * amm::kBT * 2);
    f.bt = pl.take<BTrial>(amm::kBT);
    f.alpha = pl.take<float>(q.n);
    f.nrm2 = pl.take<double>(q.n);
    f.cnrm2 = pl.take<double>(q.n);
    f.cdf0 = pl.take<uint64_t>(q.n);
    f.cdf1 = pl.take<uint64_t>(q.n);
    f.z = pl.take<float>(dpad8);
    f.v = pl.take<float>(dpad8);
    f.eigvec = pl.take<double>(dpad8);
    b->sb[i].dpad8 = dpad8;
    max_n = std::max<uint64_t>(max_n, q.n);
    uint64_t zs = cdiv(q.n, kZChunk);
    if (amm::colstats_fuses_z(q.dtype, (uint32_t)q.d))
      zs = std::max<uint64_t>(zs, amm::colstats_z_blocks(q.dtype, (uint32_t)q.d, q.n));
    max_zs = std::max<uint64_t>(max_zs, zs * dpad8);
  }
  if (hipMalloc((void**)&b->spool, pl.bytes) != hipSuccess)
    return bail(fail(AMM_E_NOMEM, "device allocation of " + std::to_string(pl.bytes) + " bytes failed"));
  if (hipMemset(b->spool, 0, pl.bytes) != hipSuccess) return bail(fail(AMM_E_HIP, "hipMemset failed"));
  char* s0 = b->spool;
  for (uint32_t i = 0; i < count; ++i) {
    BProb& p = b->hp[i];
    SetupBufs& s = b->sb[i];
    const Off& f = o[i];
    p.X = reinterpret_cast<float*>(s0 + f.X);
    p.Y = reinterpret_cast<float*>(s0 + f.Y);
    p.XM = reinterpret_cast<float*>(s0 + f.XM);
    p.YM = reinterpret_cast<float*>(s0 + f.YM);
    p.DX = reinterpret_cast<float*>(s0 + f.DX);
    p.DXh = reinterpret_cast<uint16_t*>(s0 + f.DXh);
    p.DXl = reinterpret_cast<uint16_t*>(s0 + f.DXl);
    p.cnt = reinterpret_cast<uint16_t*>(s0 + f.cnt);
    p.list = reinterpret_cast<uint32_t*>(s0 + f.list);
    p.tcount = reinterpret_cast<uint32_t*>(s0 + f.tcount);
    p.tpre = reinterpret_cast<uint32_t*>(s0 + f.tpre);
    p.colc = reinterpret_cast<uint32_t*>(s0 + f.colc);
    p.Tc = reinterpret_cast<float*>(s0 + f.Tc);
    p.BP = reinterpret_cast<double*>(s0 + f.BP);
    p.BQ = reinterpret_cast<double*>(s0 + f.BQ);
    p.bt = reinterpret_cast<BTrial*>(s0 + f.bt);
    s.alpha = reinterpret_cast<float*>(s0 + f.alpha);
    s.nrm2 = reinterpret_cast<double*>(s0 + f.nrm2);
    s.cnrm2 = reinterpret_cast<double*>(s0 + f.cnrm2);
    s.cdf0 = reinterpret_cast<uint64_t*>(s0 + f.cdf0);
    s.cdf1 = reinterpret_cast<uint64_t*>(s0 + f.cdf1);
    s.z = reinterpret_cast<float*>(s0 + f.z);
    s.v = reinterpret_cast<float*>(s0 + f.v);
    s.st = reinterpret_cast<DevState*>(s0 + st_all) + i;
    s.eigvec = reinterpret_cast<double*>(s0 + f.eigvec);
    p.st = s.st;
    p.cdf0 = s.cdf0;
    p.cdf1 = s.cdf1;
    p.alpha = s.alpha;
    p.z = s.z;
    p.v = s.v;
    p.eigvec = nullptr;
  }
  // Non-temporal loads of A in the contractions where every problem has
  // d >= 4096 (+5% on the dense passes of 16384 x 65536 and 64 x 4096^2),
  // and wherever the batch's matrices exceed twice the 256 MB Infinity
  // Cache, so that no column survives from one pass to the next anyway
  // (256 matrices, 36 GB BF16: 12.32 against 13.01 ms per dense pass).
  {
    bool all_tall = true;
    uint64_t bytes = 0;
    for (uint32_t i = 0; i < count; ++i) {
      all_tall = all_tall && probs[i].d >= 4096;
      bytes += probs[i].d * probs[i].n * (probs[i].dtype == AMM_DTYPE_BF16 ? 2u : 4u);
    }
    b->nt_loads = all_tall || bytes > kNtMinBytes;
  }
  // The union of a product in one launch (a workgroup per problem) where
  // every problem's count rows are small enough for one workgroup to read
  // (64 x 4096^2 with one trial: three launches of about 5 us each become
  // one); otherwise the three grid-wide passes.
  b->union_fused = true;
  for (uint32_t i = 0; i < count; ++i)
    b->union_fused = b->union_fused && cdiv(probs[i].n, amm::kBTile) <= amm::kUnionMaxTiles &&
                     (uint64_t)trials * probs[i].n * 2u <= kUnionFusedBytes;
  uint64_t tiles = 0;
  b->tiles.resize(count);
  for (uint32_t i = 0; i < count; ++i) {
    b->tiles[i] = (uint32_t)amm::cdf_tiles(b->hp[i].n);
    tiles += b->tiles[i];
  }
  if (hipMalloc((void**)&b->tile_max, tiles * 8) != hipSuccess ||
      hipMalloc((void**)&b->tile_sum, tiles * 8) != hipSuccess ||
      hipMalloc((void**)&b->tile_lp, tiles * 8) != hipSuccess ||
      hipMalloc((void**)&b->tile_a2, tiles * 8) != hipSuccess ||
      hipMalloc((void**)&b->zP2, std::max<uint64_t>(max_zs, 1) * 4) != hipSuccess ||
      hipMalloc((void**)&b->zP2s, std::max<uint64_t>(cdiv(max_n, kZChunk), 1) * 8) != hipSuccess ||
      hipMalloc((void**)&b->dP, count * sizeof(BProb)) != hipSuccess ||
      hipMalloc((void**)&b->dS, count * sizeof(amm::BSetup)) != hipSuccess)
    return bail(fail(AMM_E_NOMEM, "device allocation failed"));
  {
    std::vector<amm::BSetup> hs(count);
    uint64_t t0 = 0;
    for (uint32_t i = 0; i < count; ++i) {
      const BProb& p = b->hp[i];
      const SetupBufs& s = b->sb[i];
      amm::BSetup& q = hs[i];
      q = amm::BSetup{};
      q.A = p.A;
      q.ld = p.ld;
      q.d = p.d;
      q.n = p.n;
      q.dtype = p.dtype;
      q.fusable = amm::colstats_fuses_z(p.dtype, p.d) ? 1u : 0u;
      q.alpha = s.alpha;
      q.nrm2 = s.nrm2;
      q.cnrm2 = s.cnrm2;
      q.cdf[0] = s.cdf0;
      q.cdf[1] = s.cdf1;
      q.st = s.st;
      q.z = s.z;
      q.v = s.v;
      q.dpad = s.dpad8;
      q.tile_max = b->tile_max + t0;
      q.tile_sum = b->tile_sum + t0;
      q.tile_lp = b->tile_lp + t0;
      q.tile_a2 = b->tile_a2 + t0;
      q.ntiles = b->tiles[i];
      t0 += b->tiles[i];
    }
    if (copy_sync(b->stream, b->dS, hs.data(), count * sizeof(amm::BSetup), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(AMM_E_HIP, "hipMemcpy failed"));
  }
  if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return bail(fail(AMM_E_HIP, "hipEventCreate failed"));
  default_tuning(b);
  *out = b;
  return AMM_OK;
}

int amm_batch_destroy(amm_batch* b) {
  if (!b) return AMM_OK;
  std::lock_guard<std::recursive_mutex> lk(amm_host::device_mutex());
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  void* bufs[] = {b->spool, b->tpool, b->rpool, b->tile_max, b->tile_sum, b->tile_lp, b->tile_a2,
                  b->zP2,   b->zP2s,  b->offs,  b->dP,       b->dS,       b->dsig,    b->dvec};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return AMM_OK;
}

int amm_batch_set_exact_eigenvector(amm_batch* b, uint32_t problem, const double* v, uint64_t len) {
  if (!b) return fail(AMM_E_INVALID, "batch is NULL");
  if (problem >= b->np) return fail(AMM_E_INVALID, "problem index out of range");
  BProb& p = b->hp[problem];
  if (!v) {
    p.eigvec = nullptr;
  } else {
    if (len != p.d) return fail(AMM_E_INVALID, "eigenvector length must equal d");
    BHIPCHK(copy_sync(b->stream, b->sb[problem].eigvec, v, len * sizeof(double), hipMemcpyHostToDevice));
    p.eigvec = b->sb[problem].eigvec;
  }
  b->plan_valid = false;
  drop_graph(b);
  return AMM_OK;
}

int amm_batch_power_method(amm_batch* b, const amm_params* p, double* time_ms, double* eig, double* err) {
  if (!b) return fail(AMM_E_INVALID, "batch is NULL");
  int rc;
  if ((rc = run(b, p))) return rc;
  const uint64_t S = p->step;
  if (time_ms) {
    std::vector<int64_t> t(S);
    BHIPCHK(copy_sync(b->stream, t.data(), b->ts, S * sizeof(int64_t), hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < S; ++i) time_ms[i] = (double)t[i] / (double)b->wall_khz;
  }
  if (!eig && !err) return AMM_OK;
  // one copy of the whole output pool, then each problem's first ntr trials
  std::vector<char> h(b->rpool_bytes);
  BHIPCHK(copy_sync(b->stream, h.data(), b->rpool, h.size(), hipMemcpyDeviceToHost));
  for (int which = 0; which < 2; ++which) {
    double* dst = which ? err : eig;
    if (!dst) continue;
    for (uint32_t i = 0; i < b->np; ++i) {
      const char* src = reinterpret_cast<const char*>(which ? b->hp[i].err : b->hp[i].eig);
      std::memcpy(dst + (uint64_t)i * b->ntr * S, h.data() + (src - b->rpool),
                  (size_t)b->ntr * S * sizeof(double));
    }
  }
  return AMM_OK;
}

int amm_batch_last_modes(amm_batch* b, uint32_t* modes, uint64_t len) {
  if (!b || !modes) return fail(AMM_E_INVALID, "NULL argument");
  if (!b->ran) return fail(AMM_E_STATE, "no run yet");
  const uint64_t S = b->last_step;
  if (len < (uint64_t)b->np * b->ntr * S) return fail(AMM_E_INVALID, "modes buffer too small");
  std::vector<char> h(b->rpool_bytes);
  BHIPCHK(copy_sync(b->stream, h.data(), b->rpool, h.size(), hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < b->np; ++i) {
    const char* src = reinterpret_cast<const char*>(b->hp[i].modes);
    std::memcpy(modes + (uint64_t)i * b->ntr * S, h.data() + (src - b->rpool),
                (size_t)b->ntr * S * sizeof(uint32_t));
  }
  return AMM_OK;
}

int amm_batch_get_vector(amm_batch* b, uint32_t problem, uint32_t trial, int32_t which, float* host,
                         uint64_t len) {
  if (!b || !host) return fail(AMM_E_INVALID, "NULL argument");
  if (problem >= b->np || trial >= b->ntr) return fail(AMM_E_INVALID, "problem or trial out of range");
  const BProb& p = b->hp[problem];
  if (len < p.d) return fail(AMM_E_INVALID, "buffer shorter than d");
  const float* src = (which == 1 ? p.Y : p.X) + (uint64_t)trial * p.dstage;
  BHIPCHK(hipStreamSynchronize(b->stream));
  BHIPCHK(copy_sync(b->stream, host, src, p.d * sizeof(float), hipMemcpyDeviceToHost));
  return AMM_OK;
}

int amm_batch_last_run_ms(amm_batch* b, double* ms) {
  if (!b || !ms) return fail(AMM_E_INVALID, "NULL argument");
  *ms = b->last_ms;
  return AMM_OK;
}

int amm_batch_set_tuning(amm_batch* b, uint32_t problem, uint32_t dot_blocks, uint32_t passes,
                         uint32_t chunks) {
  if (!b) return fail(AMM_E_INVALID, "batch is NULL");
  if (problem >= b->np) return fail(AMM_E_INVALID, "problem index out of range");
  const Tuning t{dot_blocks, passes, chunks};
  int rc;
  if ((rc = check_tuning(b->hp[problem], t))) return rc;
  b->tune[problem] = t;
  b->plan_valid = false;
  drop_graph(b);
  return AMM_OK;
}

int amm_batch_get_tuning(amm_batch* b, uint32_t problem, uint32_t* dot_blocks, uint32_t* passes,
                         uint32_t* chunks) {
  if (!b) return fail(AMM_E_INVALID, "batch is NULL");
  if (problem >= b->np) return fail(AMM_E_INVALID, "problem index out of range");
  const Tuning& t = b->tune[problem];
  if (dot_blocks) *dot_blocks = t.dot_blocks;
  if (passes) *passes = t.passes;
  if (chunks) *chunks = t.chunks;
  return AMM_OK;
}

// Union of the trials' samples of iteration `iter` for one problem: runs
// the setup and the draw + compaction kernels only (every trial in the step
// kind of `kernel`: previous -> plain CDF, proposed -> centred CDF) and
// copies the compacted columns (ascending) with their 16 weights.
int amm_batch_sample(amm_batch* b, int32_t kernel, uint32_t iter, uint32_t trial0, uint32_t problem,
                     uint32_t* cols_host, float* w_host, uint64_t* count) {
  if (!b || !count) return fail(AMM_E_INVALID, "NULL argument");
  if (kernel != AMM_KERNEL_PREVIOUS && kernel != AMM_KERNEL_PROPOSED)
    return fail(AMM_E_INVALID, "kernel must be previousAMM or proposedAMM");
  if (problem >= b->np) return fail(AMM_E_INVALID, "problem index out of range");
  BHIPCHK(hipSetDevice(b->device));
  amm_params p{};
  p.kernel = kernel;
  p.step = 1;
  p.vr_period = 0;
  p.trial = trial0;
  int rc;
  if ((rc = validate(b, &p))) return rc;
  if ((rc = prepare(b, &p))) return rc;
  drop_graph(b);
  const amm::BRun r = run_args(b, p);
  BHIPCHK(amm::launch_bbegin(b->dP, b->g_vec, r, b->stream));
  if ((rc = enq_setup(b, kernel))) return rc;
  b->stats_valid = false;
  BHIPCHK(amm::launch_bsample(b->dP, b->g_sample, r, iter, b->stream));
  BHIPCHK(amm::launch_bunion(b->dP, b->g_prep, r, b->union_fused, b->stream));
  // the per-trial weights of every union slot, as the products form them
  BHIPCHK(amm::launch_bweight(b->dP, b->g_weight, r, 0u, 1u, b->stream));
  const BProb& q = b->hp[problem];
  uint32_t U = 0;
  BHIPCHK(copy_sync(b->stream, &U, q.tpre + q.ntiles, sizeof(uint32_t), hipMemcpyDeviceToHost));
  std::vector<uint32_t> cols(U);
  std::vector<float> W((size_t)U * amm::kBT);
  if (U) {
    BHIPCHK(copy_sync(b->stream, cols.data(), q.colc, cols.size() * 4, hipMemcpyDeviceToHost));
    BHIPCHK(copy_sync(b->stream, W.data(), q.Tc, W.size() * 4, hipMemcpyDeviceToHost));
  }
  for (uint32_t m = 0; m < U; ++m) {
    if (cols_host) cols_host[m] = cols[m];
    if (w_host)
      for (uint32_t k = 0; k < b->ntr; ++k) w_host[(size_t)m * b->ntr + k] = W[(size_t)m * amm::kBT + k];
  }
  *count = U;
  return AMM_OK;
}

// sigma_max of every (problem, trial) after a run with parameters p, and
// the singular vectors of trial 0 (see include/amm.h).
int amm_batch_sigma_max(amm_batch* b, const amm_params* p, double* sigma_host, float* const* u_dev,
                        float* const* v_dev) {
  if (!b) return fail(AMM_E_INVALID, "batch is NULL");
  int rc;
  if ((rc = run(b, p))) return rc;
  const amm::BRun r = run_args(b, *p);
  const size_t ns = (size_t)b->np * b->ntr;
  // kept for the batch's lifetime (np and ntr are fixed at creation)
  std::unique_lock<std::recursive_mutex> lk(amm_host::device_mutex());
  if (!b->dsig && hipMalloc((void**)&b->dsig, ns * sizeof(double)) != hipSuccess) {
    b->dsig = nullptr;
    return fail(AMM_E_NOMEM, "device allocation failed");
  }
  if (!b->dvec && hipMalloc((void**)&b->dvec, 2 * (size_t)b->np * sizeof(float*)) != hipSuccess) {
    b->dvec = nullptr;
    return fail(AMM_E_NOMEM, "device allocation failed");
  }
  lk.unlock();
  double* dsig = b->dsig;
  float** dptr = nullptr;
  std::vector<float*> hp;
  if (u_dev || v_dev) {
    hp.assign(2 * (size_t)b->np, nullptr);
    for (uint32_t i = 0; i < b->np; ++i) {
      if (u_dev) hp[i] = u_dev[i];
      if (v_dev) hp[b->np + i] = v_dev[i];
    }
    dptr = b->dvec;
    BHIPCHK(hipMemcpyAsync(dptr, hp.data(), hp.size() * sizeof(float*), hipMemcpyHostToDevice, b->stream));
  }
  // t = A^T x for every trial: one exact half-step over the dense union
  hipError_t e = amm::launch_bset_mode(b->dP, b->np, amm::kModeExactNoMemo, b->stream);
  if (e == hipSuccess) e = amm::launch_bnorm(b->dP, b->g_norm, r, b->stream);
  if (e == hipSuccess) e = amm::launch_bunion(b->dP, b->g_prep, r, b->union_fused, b->stream);
  for (int dt = 0; dt < 2 && e == hipSuccess; ++dt) e = amm::launch_bdot(dt, b->nt_loads, b->dP, b->g_dot[dt], b->np, b->stream);
  if (e == hipSuccess) e = amm::launch_bweight(b->dP, b->g_weight, r, 0u, 0u, b->stream);
  if (e == hipSuccess)
    e = amm::launch_bsigma(b->dP, r, dsig, u_dev && dptr ? dptr : nullptr,
                           v_dev && dptr ? dptr + b->np : nullptr, b->stream);
  if (e == hipSuccess && sigma_host)
    e = hipMemcpyAsync(sigma_host, dsig, ns * sizeof(double), hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);  // also keeps hp alive until its copy ran
  if (e != hipSuccess) return fail(AMM_E_HIP, std::string("batched sigma_max failed: ") + hipGetErrorString(e));
  return AMM_OK;
}

}  // extern "C"

namespace {

// Candidate shapes of the tuners: multipliers of the default workgroup
// targets of both contractions, and the second contraction's passes.
struct Cand {
  double ks_mult, ch_mult;
  uint32_t passes;
};
const Cand kCands[] = {{1, 1, 0},   {0.5, 0.5, 0},   {2, 2, 0}, {1, 1, 1}, {1, 1, 2}, {2, 0.5, 0},
                       {0.5, 2, 0}, {0.25, 0.25, 0}, {4, 4, 0}, {4, 4, 1}, {4, 4, 2}, {8, 8, 0}};

void set_candidate(amm_batch* b, const Cand& cd) {
  double tot = 0.0;
  for (const BProb& q : b->hp) tot += (double)q.ntiles * (double)q.d;
  for (uint32_t i = 0; i < b->np; ++i) {
    const BProb& q = b->hp[i];
    const double share = tot > 0 ? (double)q.ntiles * (double)q.d / tot : 1.0;
    b->tune[i] = default_tuning_one(q, share, cd.ks_mult, cd.ch_mult, cd.passes);
  }
  b->plan_valid = false;
  drop_graph(b);
}

// Mean time of one pass of both contractions over the dense union (every
// column present, all trials exact) with the current shapes.
int time_dense(amm_batch* b, uint32_t iters, double* ms_out) {
  amm_params p{};
  p.kernel = AMM_KERNEL_EXACT;
  p.step = 1;
  int rc;
  if ((rc = prepare(b, &p))) return rc;
  drop_graph(b);
  const amm::BRun r = run_args(b, p);
  BHIPCHK(amm::launch_bbegin(b->dP, b->g_vec, r, b->stream));
  BHIPCHK(amm::launch_bunion(b->dP, b->g_prep, r, b->union_fused, b->stream));
  for (uint32_t w = 0; w < 2; ++w) {  // warm-up, then timed
    BHIPCHK(hipEventRecord(b->ev0, b->stream));
    for (uint32_t k = 0; k < (w ? iters : 1); ++k) {
      for (int dt = 0; dt < 2; ++dt) BHIPCHK(amm::launch_bdot(dt, b->nt_loads, b->dP, b->g_dot[dt], b->np, b->stream));
      BHIPCHK(amm::launch_bweight(b->dP, b->g_weight, r, 0u, 0u, b->stream));
      for (int dt = 0; dt < 2; ++dt)
        for (int pa = 0; pa < 2; ++pa)
          BHIPCHK(amm::launch_baxpy(dt, (uint32_t)pa + 1, b->nt_loads, b->dP, b->g_axpy[dt][pa], b->np, 0u, b->stream));
    }
    BHIPCHK(hipEventRecord(b->ev1, b->stream));
  }
  BHIPCHK(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  BHIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  *ms_out = (double)ms / iters;
  return AMM_OK;
}

}  // namespace

extern "C" {

// Time candidate shapes of the two contractions on the live problems (one
// exact-step iteration: every column present) and keep the fastest per
// problem size class.  best_ms: time of one iteration's contractions with
// the chosen shapes.
int amm_batch_autotune(amm_batch* b, uint32_t iters, double* best_ms) {
  if (!b) return fail(AMM_E_INVALID, "batch is NULL");
  if (iters < 1) iters = 1;
  BHIPCHK(hipSetDevice(b->device));
  int rc;
  double best = 1e30;
  std::vector<Tuning> best_t = b->tune;
  const std::vector<Tuning> orig = b->tune;
  auto restore = [&](int r) {  // a failed candidate leaves the shapes as they were
    b->tune = orig;
    b->plan_valid = false;
    drop_graph(b);
    return r;
  };
  for (const Cand& cd : kCands) {
    set_candidate(b, cd);
    double per = 0.0;
    if ((rc = time_dense(b, iters, &per))) return restore(rc);
    if (per < best) {
      best = per;
      best_t = b->tune;
    }
  }
  b->tune = best_t;
  b->plan_valid = false;
  drop_graph(b);
  if (best_ms) *best_ms = best;
  return AMM_OK;
}

int amm_batch_autotune_run(amm_batch* b, const amm_params* p, uint32_t reps, double* best_ms) {
  if (!b || !p) return fail(AMM_E_INVALID, "NULL argument");
  if (reps < 1) reps = 1;
  BHIPCHK(hipSetDevice(b->device));
  int rc;
  if ((rc = validate(b, p))) return rc;
  double best = 1e30;
  std::vector<Tuning> best_t = b->tune;
  const std::vector<Tuning> orig = b->tune;
  auto restore = [&](int r) {  // a failed run leaves the shapes as they were
    b->tune = orig;
    b->plan_valid = false;
    drop_graph(b);
    return r;
  };
  // best of `reps` runs after the one that captures the graph
  auto eval = [&](double* t) -> int {
    *t = 1e30;
    for (uint32_t k = 0; k <= reps; ++k) {
      int r;
      if ((r = run(b, p))) return r;
      if (k) *t = std::min(*t, b->last_ms);
    }
    return AMM_OK;
  };
  auto consider = [&](double t) {
    if (t < best) {
      best = t;
      best_t = b->tune;
    }
  };
  double t = 0.0;
  if ((rc = eval(&t))) return restore(rc);  // the shapes in place (kept on a tie)
  consider(t);
  // coordinate descent over the multipliers of the default targets: the
  // chunk count of the second contraction first (it moves run times most),
  // then the first contraction's workgroups, then the passes
  Cand cur{1, 1, 0};
  double cur_t = 1e30;
  auto step = [&](const Cand& cd) -> int {
    set_candidate(b, cd);
    double tc = 0.0;
    int r;
    if ((r = eval(&tc))) return r;
    consider(tc);
    if (tc < cur_t) {
      cur_t = tc;
      cur = cd;
    }
    return AMM_OK;
  };
  if ((rc = step(cur))) return restore(rc);
  const Cand base_ch = cur;
  for (double ch : {0.25, 0.5, 2.0, 4.0, 8.0, 16.0})
    if ((rc = step(Cand{base_ch.ks_mult, ch, base_ch.passes}))) return restore(rc);
  const Cand base_ks = cur;
  for (double ks : {0.25, 0.5, 2.0, 4.0})
    if ((rc = step(Cand{ks, base_ks.ch_mult, base_ks.passes}))) return restore(rc);
  const Cand base_pa = cur;
  for (uint32_t pa : {1u, 2u})
    if ((rc = step(Cand{base_pa.ks_mult, base_pa.ch_mult, pa}))) return restore(rc);
  b->tune = best_t;
  b->plan_valid = false;
  drop_graph(b);
  if (best_ms) *best_ms = best;
  return AMM_OK;
}

int amm_batch_time_dense(amm_batch* b, uint32_t iters, double* ms) {
  if (!b || !ms) return fail(AMM_E_INVALID, "NULL argument");
  if (iters < 1) iters = 1;
  BHIPCHK(hipSetDevice(b->device));
  return time_dense(b, iters, ms);
}

}  // extern "C"
