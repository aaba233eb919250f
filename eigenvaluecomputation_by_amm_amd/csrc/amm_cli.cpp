// amm_cli.cpp -- the AMM command line: same flags, defaults, stdout CSV and
// result/*.binary files as the reference's AMM.cu, driving libamm.so.
//
//   AMM [-m M] [-c C] [-r NREPS] [-step S] [-kernel exact|previousAMM|
//       proposedAMM|proposedAMM_v2] [-matrixType gaussian|lognormal|uniform_P1_P2]
//       [-filepath F] [-seed S] [-sanity] [-display] [-vr N] [-vr_adaptive T]
//       [-verbose V] [-dtype f32|bf16]
//
// host_memory mirrors host_memory.h:23-192, device_memory the parts of
// device_memory.h:116-498 the path needs, powerMethod_by_* the four drivers'
// output stage (powerMethod_by_exact.cuh:71-99 and its three siblings), and
// main AMM.cu:74-95.  Deliberate differences: argument errors print
// "[error] ..." and exit 1 instead of assert(); with no -filepath the result
// file is named after the matrix type; a missing data/exact_eigenvector
// prints a warning and reads as zeros (NormError 1) instead of undefined
// memory; the Gershgorin offset is the true maximum.
#include <hip/hip_runtime.h>
#include <inttypes.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "amm.h"

template <typename T>
inline T DIV_CEIL(const T x, const T y) {
  return (x + y - 1) / y;
}

static std::vector<std::string> split(const std::string& str, char sep) {
  std::vector<std::string> v;
  std::stringstream ss(str);
  std::string buffer;
  while (std::getline(ss, buffer, sep)) v.push_back(buffer);
  return v;
}

static void die(const char* what) {
  std::fprintf(stderr, "[error] %s: %s\n", what, amm_last_error());
  std::exit(1);
}

// ---------------------------------------------------------------------------
class host_memory {  // host_memory.h:23-192
 public:
  uint64_t original_m = 128;
  uint64_t m = 128;
  uint32_t c = 0;
  uint64_t step = 10;
  uint64_t vr_period = 0;
  std::string str_vr_adaptive_threshold = "";
  double vr_adaptive_threshold = 0.0;
  std::string kernel = "exact";
  std::string matrixType = "gaussian_0.0_1.0";
  std::string filepath = "";
  std::string dtype = "f32";
  int verbose = 0;
  uint32_t nreps = 10;
  uint32_t seed;
  float diagonal_offset = 0.0f;
  bool sanity_check = false;
  bool variance_reduction = false;
  bool display = false;
  bool dump_matrix = false;

  host_memory() { seed = (uint32_t)time(NULL); }

  bool parser(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
      auto next = [&](const char* flag) -> const char* {
        if (i + 1 >= argc) {
          std::printf("[error] %s needs a value.\n", flag);
          std::exit(1);
        }
        return argv[++i];
      };
      if (!strcmp(argv[i], "-m")) {
        m = std::atoi(next("-m"));
        m = DIV_CEIL(m, (uint64_t)128) * 128;
        original_m = m;
      } else if (!strcmp(argv[i], "-c")) {
        c = std::atoi(next("-c"));
      } else if (!strcmp(argv[i], "-r")) {
        nreps = std::atoi(next("-r"));
      } else if (!strcmp(argv[i], "-step")) {
        step = std::atoi(next("-step"));
      } else if (!strcmp(argv[i], "-kernel")) {
        kernel = next("-kernel");
      } else if (!strcmp(argv[i], "-matrixType")) {
        matrixType = next("-matrixType");
      } else if (!strcmp(argv[i], "-filepath")) {
        filepath = next("-filepath");
      } else if (!strcmp(argv[i], "-seed")) {
        seed = std::atoi(next("-seed"));
      } else if (!strcmp(argv[i], "-sanity")) {
        sanity_check = true;
      } else if (!strcmp(argv[i], "-display")) {
        display = true;
      } else if (!strcmp(argv[i], "-vr")) {
        vr_period = std::atoi(next("-vr"));
        variance_reduction = (vr_period > 0);
      } else if (!strcmp(argv[i], "-vr_adaptive")) {
        str_vr_adaptive_threshold = next("-vr_adaptive");
        vr_adaptive_threshold = std::atof(str_vr_adaptive_threshold.c_str());
        variance_reduction = (vr_adaptive_threshold > 0.0);
      } else if (!strcmp(argv[i], "-verbose")) {
        verbose = std::atoi(next("-verbose"));
      } else if (!strcmp(argv[i], "-dtype")) {
        dtype = next("-dtype");
      } else if (!strcmp(argv[i], "-dump_matrix")) {
        dump_matrix = true;
      } else {
        std::printf("[warning] unknown parameter: %s\n", argv[i]);
      }
    }

    if (kernel == "exact") {
      c = (uint32_t)m;
      sanity_check = false;
    } else if (kernel != "previousAMM" && kernel != "proposedAMM" && kernel != "proposedAMM_v2") {
      std::printf("[error] '-kernel %s' is not valid.\n", kernel.c_str());
      return false;
    }
    if (dtype != "f32" && dtype != "bf16") {
      std::printf("[error] '-dtype %s' is not valid.\n", dtype.c_str());
      return false;
    }

    if (filepath != "") {
      std::ifstream ifs(filepath, std::ios::in | std::ios::binary);
      if (ifs.fail()) {
        std::cerr << "Failed to open file." << std::endl;
        return false;
      }
      uint32_t val = 0;
      ifs.read((char*)&val, sizeof(uint32_t));
      if (!ifs || val == 0) {
        std::cerr << "Failed to read the matrix size." << std::endl;
        return false;
      }
      original_m = val;
      m = DIV_CEIL(val, (uint32_t)128) * 128;
      if (kernel == "exact") c = (uint32_t)m;
    }

    if (c == 0) c = (uint32_t)m;

    // host_memory.h:166-169 (asserts in the reference)
    if (!(1 <= c && c <= m)) {
      std::printf("[error] c must satisfy 1 <= c <= m (c = %u, m = %" PRIu64 ").\n", c, m);
      return false;
    }
    if (c % 8 != 0) {
      std::printf("[error] c must be a multiple of 8 (c = %u).\n", c);
      return false;
    }
    if (kernel == "proposedAMM_v2" && (vr_period < 1 || vr_period > step)) {
      std::printf("[error] proposedAMM_v2 needs 1 <= vr <= step.\n");
      return false;
    }

    if (verbose >= 2) {
      std::printf("[info] Parameters are as follows.\n");
      std::printf("\tkernel           : %s\n", kernel.c_str());
      std::printf("\tm                : %" PRIu64 "\n", m);
      std::printf("\tc                : %u\n", c);
      std::printf("\tnreps            : %u\n", nreps);
      std::printf("\tseed             : %d\n", seed);
      std::printf("\tdtype            : %s\n", dtype.c_str());
      std::printf("\tdiagonal offset  : %f\n", diagonal_offset);
      std::printf("\tvr period        : %" PRIu64 "\n", vr_period);
      std::printf("[info] Sanity check      : %s\n", sanity_check ? "on" : "off");
      std::printf("[info] Variance reduction: %s\n", variance_reduction ? "on" : "off");
    }
    return true;
  }
};

// ---------------------------------------------------------------------------
struct device_memory {  // device_memory.h:116-498
  void* dA = nullptr;     // m x m, FP32 or BF16
  float* dA32 = nullptr;  // FP32 staging when dtype is bf16
  uint64_t m = 128;
  uint64_t original_m = 128;
  std::string kernel;
  amm_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  std::ofstream ofs_output;
  uint32_t rep = 0;

  explicit device_memory(host_memory& p) {
    m = p.m;
    original_m = p.original_m;
    kernel = p.kernel;
    if (hipStreamCreate(&stream) != hipSuccess) die("hipStreamCreate");
    const size_t elems = (size_t)m * m;
    if (hipMalloc(&dA32, sizeof(float) * elems) != hipSuccess) die("hipMalloc");
    if (p.dtype == "bf16") {
      if (hipMalloc(&dA, sizeof(uint16_t) * elems) != hipSuccess) die("hipMalloc");
    } else {
      dA = dA32;
    }

    std::string base;
    if (p.filepath != "") {
      auto ary = split(p.filepath, '/');
      base = ary.empty() ? p.filepath : ary[ary.size() - 1];
    } else {
      base = p.matrixType;
    }
    std::string path_output("result/eigenvalues_by");
    path_output += "_" + kernel;
    path_output += "_" + base;
    path_output += "_r=" + std::to_string(p.nreps);
    path_output += "_step=" + std::to_string(p.step);
    if (p.kernel != "exact") {
      path_output += "_c=" + std::to_string(p.c);
      if (p.vr_adaptive_threshold == 0.0)
        path_output += "_vr=" + std::to_string(p.vr_period);
      else
        path_output += "_vr=adaptive-" + p.str_vr_adaptive_threshold;
    }
    path_output += ".binary";
    ofs_output.open(path_output, std::ios::out | std::ios::binary);
    if (!ofs_output) std::fprintf(stderr, "[warning] cannot open %s for writing\n", path_output.c_str());
  }

  ~device_memory() {
    ofs_output.close();
    if (ctx) amm_ctx_destroy(ctx);
    if (dA && dA != dA32) (void)hipFree(dA);
    if (dA32) (void)hipFree(dA32);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // matrix.dat: the m x m matrix as stored (padded), float32 in memory order,
  // with the diagonal offset taken off the diagonal again -- the file the
  // reference's output_matrix writes (device_memory.h:436-473) for an FP64
  // eigen-solve (calculate_largest_eigenvalue.py).
  void output_matrix(float diagonal_offset) const {
    std::vector<float> h((size_t)m * m);
    if (hipMemcpy(h.data(), dA32, sizeof(float) * h.size(), hipMemcpyDeviceToHost) != hipSuccess)
      die("hipMemcpy");
    for (uint64_t i = 0; i < m; ++i) h[i * m + i] -= diagonal_offset;
    std::ofstream fs("matrix.dat", std::ios::out | std::ios::binary);
    fs.write((const char*)h.data(), sizeof(float) * h.size());
    if (!fs) std::fprintf(stderr, "[warning] cannot write matrix.dat\n");
  }

  // device_memory.h:316-337: generate or read A, then the column statistics
  void write_data(host_memory& p) {
    const size_t elems = (size_t)m * m;
    if (p.filepath != "") {
      std::vector<float> h(elems, 0.0f);
      std::ifstream ifs(p.filepath, std::ios::in | std::ios::binary);
      ifs.seekg(sizeof(uint32_t));
      std::vector<float> row(original_m);
      for (uint64_t y = 0; y < original_m; ++y) {
        ifs.read((char*)row.data(), sizeof(float) * original_m);
        if (!ifs) {
          std::fprintf(stderr, "[error] %s is truncated (row %" PRIu64 ")\n", p.filepath.c_str(), y);
          std::exit(1);
        }
        std::copy(row.begin(), row.end(), h.begin() + y * m);
      }
      if (hipMemcpy(dA32, h.data(), sizeof(float) * elems, hipMemcpyHostToDevice) != hipSuccess)
        die("hipMemcpy");
      p.diagonal_offset = 0.0f;
    } else {
      auto ary = split(p.matrixType, '_');
      const std::string type = ary.empty() ? "" : ary[0];
      const float p1 = ary.size() <= 1 ? 0.0f : (float)std::atof(ary[1].c_str());
      const float p2 = ary.size() <= 2 ? 0.0f : (float)std::atof(ary[2].c_str());
      int dist;
      if (type == "gaussian") dist = AMM_DIST_GAUSSIAN;
      else if (type == "lognormal") dist = AMM_DIST_LOGNORMAL;
      else if (type == "uniform") dist = AMM_DIST_UNIFORM;
      else {
        std::printf("[error] '-matrixType %s' is not valid.\n", p.matrixType.c_str());
        std::exit(1);
      }
      if (amm_generate_symmetric(dA32, m, m, dist, p1, p2, p.seed, stream, &p.diagonal_offset))
        die("amm_generate_symmetric");
    }
    if (p.dump_matrix) output_matrix(p.diagonal_offset);
    int dt = AMM_DTYPE_F32;
    if (dA != dA32) {
      if (amm_f32_to_bf16(dA32, dA, elems, stream)) die("amm_f32_to_bf16");
      if (hipStreamSynchronize(stream) != hipSuccess) die("hipStreamSynchronize");
      dt = AMM_DTYPE_BF16;
    }
    if (amm_ctx_create(dA, dt, m, m, m, stream, &ctx)) die("amm_ctx_create");
    if (amm_column_stats(ctx, nullptr, nullptr, nullptr, nullptr)) die("amm_column_stats");

    // data/exact_eigenvector (device_memory.h:246-254)
    std::vector<double> v(m, 0.0);
    std::ifstream ifs("data/exact_eigenvector", std::ios::in | std::ios::binary);
    if (ifs) {
      ifs.read((char*)v.data(), sizeof(double) * original_m);
      if (!ifs) std::fprintf(stderr, "[warning] data/exact_eigenvector is shorter than %" PRIu64 "\n",
                             original_m);
    } else if (p.verbose >= 1) {
      std::fprintf(stderr, "[warning] data/exact_eigenvector not found; NormError reads 1\n");
    }
    if (amm_set_exact_eigenvector(ctx, v.data(), m)) die("amm_set_exact_eigenvector");
  }
};

// ---------------------------------------------------------------------------
// One run of a power method and, with is_output, its stdout rows and binary
// records (powerMethod_by_exact.cuh:71-99).
static void power_method(host_memory& p, device_memory& dm, int kernel, bool is_output) {
  uint64_t step = p.step;
  if (!is_output) {
    // warm-up lengths (AMM.cu:14,30,46,62 with powerMethod_by_previous.cuh:131);
    // a warm-up of 0 iterations still draws its x0 (powerMethod_by_previous.cuh:128),
    // so it still consumes one repetition of the random stream
    if (kernel == AMM_KERNEL_PREVIOUS || kernel == AMM_KERNEL_PROPOSED) step = p.vr_period;
    if (step == 0) {
      dm.rep++;
      return;
    }
  }
  amm_params P{};
  P.kernel = kernel;
  P.c = p.c;
  P.step = step;
  P.vr_period = p.vr_period;
  P.vr_adaptive_threshold = p.vr_adaptive_threshold;
  P.seed = 1234ull + p.seed;
  P.trial = dm.rep++;
  P.sanity = p.sanity_check ? 1 : 0;
  // the last of -vr / -vr_adaptive decides whether any variance-reduction
  // step happens (host_memory.h:108-120, powerMethod_by_previous.cuh:135)
  P.flags = AMM_FLAG_KEEP_STATS | (p.variance_reduction ? 0u : AMM_FLAG_NO_VR);
  std::vector<double> t(step), e(step), r(step);
  if (amm_power_method(dm.ctx, &P, t.data(), e.data(), r.data())) die("amm_power_method");
  if (!is_output) return;
  for (uint64_t i = 0; i < step; ++i) {
    double time = t[i];
    double eig = std::max(0.0, std::sqrt(std::max(0.0, e[i])) - (double)p.diagonal_offset);
    double err = 1.0 - r[i] * r[i];
    if (p.display) std::printf("%.12lf,%.12lf,%.12lf\n", time, eig, err);
    dm.ofs_output.write((char*)&time, sizeof(double));
    dm.ofs_output.write((char*)&eig, sizeof(double));
    dm.ofs_output.write((char*)&err, sizeof(double));
  }
}

void powerMethod_by_exact(host_memory& p, device_memory& dm, const bool is_output = true) {
  power_method(p, dm, AMM_KERNEL_EXACT, is_output);
}
void powerMethod_by_previous(host_memory& p, device_memory& dm, const bool is_output = true) {
  power_method(p, dm, AMM_KERNEL_PREVIOUS, is_output);
}
void powerMethod_by_proposed(host_memory& p, device_memory& dm, const bool is_output = true) {
  power_method(p, dm, AMM_KERNEL_PROPOSED, is_output);
}
void powerMethod_by_proposed_v2(host_memory& p, device_memory& dm, const bool is_output = true) {
  power_method(p, dm, AMM_KERNEL_PROPOSED_V2, is_output);
}

// AMM.cu:7-72
static void execute_kernel(host_memory& p, device_memory& dm, const bool warmup = false) {
  void (*fn)(host_memory&, device_memory&, const bool) = nullptr;
  if (p.kernel == "exact") fn = powerMethod_by_exact;
  else if (p.kernel == "previousAMM") fn = powerMethod_by_previous;
  else if (p.kernel == "proposedAMM") fn = powerMethod_by_proposed;
  else if (p.kernel == "proposedAMM_v2") fn = powerMethod_by_proposed_v2;
  if (!fn) return;
  if (warmup) {
    fn(p, dm, false);
  } else {
    std::printf("Time,Eigenvalue,NormError\n");
    for (uint32_t i = 0; i < p.nreps; ++i) fn(p, dm, true);
  }
}

int main(int argc, char* argv[]) {
  host_memory p;
  if (!p.parser(argc, argv)) return 1;
  device_memory dm(p);
  dm.write_data(p);
  execute_kernel(p, dm, true);
  if (hipDeviceSynchronize() != hipSuccess) die("hipDeviceSynchronize");
  execute_kernel(p, dm);
  std::fflush(stdout);
  return 0;
}
