// amm_noalloc.hip -- the exact product y = A (A^T x) with no device
// allocation: for callers that cannot call hipMalloc (code under hipGraph
// capture, tight serving loops, harnesses that forbid allocation).
//
// Every intermediate of the product (the padded copy of x, A^T x split by
// rows, the per-chunk partials of A t, the finish block partials and y
// before it is copied out) lives in `scratch`, one file-scope device array
// of kScratchBytes.  Each GPU the library runs on holds its own copy until
// the process ends; its device address is taken once per device with
// hipGetSymbolAddress and only that address is handed to kernels.  Calls
// that overlap on one GPU share the array and race: the caller orders them
// (one stream, or events).  Nothing in the array needs to start at zero:
// every region is written before it is read within a call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "amm.h"
#include "amm_device.hpp"
#include "amm_launch.hpp"

namespace amm_host {
int fail(int code, const std::string& msg);
}

namespace {

constexpr size_t kScratchBytes = size_t(16) << 20;  // 16 MiB per GPU
constexpr int kMaxDevices = 64;
constexpr uint32_t kTargetBlocks = 1024;  // workgroups each contraction aims for
constexpr uint32_t kMaxChunkEntries = 2048;  // k_gather_axpy's LDS chunk limit

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

int fail(int code, const std::string& msg) { return amm_host::fail(code, msg); }

// dst[i] = src[i] for i < n, 0 for n <= i < npad
__global__ __launch_bounds__(256) void k_copy_pad(const float* __restrict__ src, uint64_t n,
                                                  float* __restrict__ dst, uint64_t npad) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < npad) dst[i] = i < n ? src[i] : 0.0f;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Device address of this device's scratch array (taken on first use, with
// the device current).
int scratch_ptr(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(AMM_E_INVALID, "device ordinal beyond 64");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(AMM_E_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = (unsigned char*)g_scratch[dev];
  return AMM_OK;
}

struct Layout {
  uint32_t R = 1, rps = 0;  // row splits of A^T x and rows per split
  uint32_t cs = 1, S = 1;   // entries per chunk and chunks of A t
  size_t off_x = 0, off_y = 0, off_mode = 0, off_st = 0, off_bp = 0, off_p1 = 0, off_p2 = 0, off_p2s = 0;
  size_t bytes = 0;
};

// Launch plan sized to the scratch budget: the contractions' usual targets
// (about kTargetBlocks workgroups each), with fewer chunks of A t when
// their partials would not fit.
bool plan(uint64_t d, uint64_t n, int V, Layout& L, std::string& why) {
  const uint64_t dpad = (d + 7) / 8 * 8;
  const uint64_t groups = (n + 15) / 16;  // four waves x four columns per workgroup
  const uint32_t step = 64u * V * 2;
  uint64_t R = groups >= kTargetBlocks ? 1 : (kTargetBlocks + groups - 1) / groups;
  R = std::min<uint64_t>(R, std::max<uint64_t>(1, (d + step - 1) / step));
  uint64_t rps = (d + R - 1) / R;
  rps = (rps + step - 1) / step * step;
  R = (d + rps - 1) / rps;
  L.R = (uint32_t)R;
  L.rps = (uint32_t)rps;
  size_t off = 0;
  L.off_x = off;
  off += align256(dpad * 4);
  L.off_y = off;
  off += align256(dpad * 4);
  L.off_mode = off;
  off += 256;
  L.off_st = off;
  off += align256(sizeof(amm::DevState));
  L.off_bp = off;
  off += align256((3 * amm::finish_blocks(dpad) + 64) * 8);
  L.off_p1 = off;
  off += align256((size_t)R * n * 4);
  if (off >= kScratchBytes) {
    why = "d = " + std::to_string(d) + ", n = " + std::to_string(n) + " needs more than the " +
          std::to_string(kScratchBytes >> 20) + " MiB scratch of the allocation-free exact product";
    return false;
  }
  const uint64_t rb = (dpad + 256u * V - 1) / (256u * V);
  const uint64_t S_t = std::max<uint64_t>(1, (kTargetBlocks + rb - 1) / rb);
  uint64_t cs = std::min<uint64_t>(kMaxChunkEntries, std::max<uint64_t>(1, (n + S_t - 1) / S_t));
  uint64_t S = (n + cs - 1) / cs;
  const size_t room = kScratchBytes - off - 512;
  const uint64_t S_fit = room / (dpad * 4 + 8);
  if (S > S_fit) {
    if (S_fit == 0) {
      why = "d = " + std::to_string(d) + " leaves no room for one partial of A t in the " +
            std::to_string(kScratchBytes >> 20) + " MiB scratch";
      return false;
    }
    cs = (n + S_fit - 1) / S_fit;
    if (cs > kMaxChunkEntries) {
      why = "d = " + std::to_string(d) + ", n = " + std::to_string(n) + " needs more than the " +
            std::to_string(kScratchBytes >> 20) + " MiB scratch of the allocation-free exact product";
      return false;
    }
    S = (n + cs - 1) / cs;
  }
  L.cs = (uint32_t)cs;
  L.S = (uint32_t)S;
  L.off_p2 = off;
  off += align256((size_t)S * dpad * 4);
  L.off_p2s = off;
  off += align256((size_t)S * 8);
  L.bytes = off;
  if (off > kScratchBytes) {
    why = "internal: scratch plan exceeds its budget";
    return false;
  }
  return true;
}

#define HIPCHK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(AMM_E_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" {

int amm_exact_product_noalloc(const void* A, int32_t dtype, uint64_t d, uint64_t n, uint64_t ld,
                              const float* x_dev, float* y_dev, void* stream) {
  if (!A) return fail(AMM_E_INVALID, "A is NULL");
  if (!x_dev || !y_dev) return fail(AMM_E_INVALID, "x and y must be device pointers");
  if (dtype != AMM_DTYPE_F32 && dtype != AMM_DTYPE_BF16) return fail(AMM_E_INVALID, "unknown dtype");
  if (d < 1 || n < 1) return fail(AMM_E_INVALID, "d and n must be >= 1");
  if (d >= (1ull << 31) || n >= (1ull << 31)) return fail(AMM_E_INVALID, "d and n must be < 2^31");
  if (ld < d) return fail(AMM_E_INVALID, "ld must be >= d");
  const int V = dtype == AMM_DTYPE_BF16 ? 8 : 4;
  if (ld % V) return fail(AMM_E_INVALID, "ld must be a multiple of 16 bytes of elements");
  if (((uintptr_t)A) & 15u) return fail(AMM_E_INVALID, "A must be 16-byte aligned");
  Layout L;
  std::string why;
  if (!plan(d, n, V, L, why)) return fail(AMM_E_INVALID, why);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  unsigned char* base = nullptr;
  int rc;
  if ((rc = scratch_ptr(dev, &base))) return rc;
  const hipStream_t s = (hipStream_t)stream;
  const uint64_t dpad = (d + 7) / 8 * 8;
  float* xp = (float*)(base + L.off_x);
  float* yp = (float*)(base + L.off_y);
  uint32_t* mode = (uint32_t*)(base + L.off_mode);
  amm::DevState* st = (amm::DevState*)(base + L.off_st);
  double* BP = (double*)(base + L.off_bp);
  float* P1 = (float*)(base + L.off_p1);
  float* P2 = (float*)(base + L.off_p2);
  double* P2s = (double*)(base + L.off_p2s);

  k_copy_pad<<<(unsigned)((dpad + 255) / 256), 256, 0, s>>>(x_dev, d, xp, dpad);
  HIPCHK(hipGetLastError());
  HIPCHK(amm::launch_fill_u32(mode, 1, amm::kModeExactNoMemo, s));
  amm::GatherDotArgs g{};
  g.A = A;
  g.ld = ld;
  g.d = (uint32_t)d;
  g.cap = (uint32_t)n;
  g.x = xp;
  g.R = L.R;
  g.rows_per_split = L.rps;
  g.P1 = P1;
  HIPCHK(amm::launch_gather_dot(dtype, g, s));
  amm::GatherAxpyArgs a{};
  a.A = A;
  a.ld = ld;
  a.d = (uint32_t)d;
  a.dpad = dpad;
  a.cap = (uint32_t)n;
  a.cs = L.cs;
  a.S = L.S;
  a.P1 = P1;
  a.R = L.R;
  a.st = st;
  a.P2 = P2;
  a.P2s = P2s;
  HIPCHK(amm::launch_gather_axpy(dtype, a, s));
  amm::FinishArgs f{};
  f.P2 = P2;
  f.S_exact = L.S;
  f.S_amm = L.S;
  f.dpad = dpad;
  f.d = (uint32_t)d;
  f.x = xp;
  f.y = yp;
  f.BP = BP;
  f.st = st;
  f.mode = mode;
  f.it = 0;
  HIPCHK(amm::launch_finish(f, s));
  k_copy_pad<<<(unsigned)((d + 255) / 256), 256, 0, s>>>(yp, d, y_dev, d);
  HIPCHK(hipGetLastError());
  return AMM_OK;
}

}  // extern "C"
