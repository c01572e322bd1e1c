// tuner.cpp -- exhaustive autotuning of the products' configurations.
//
// Eigen's own answer to "which blocking for this shape" is a cache-model
// heuristic (evaluateProductBlockingSizesHeuristic,
// Eigen/src/Core/products/GeneralBlockPanelKernel.h:201-405) plus an offline
// sweep tool (benchmarks/Tuning/bench_blocking_sizes.cpp).  Here every
// candidate configuration is timed on the device with HIP events and the
// fastest one is kept in a table keyed by (kind, dtype, opa, opb, m, n, k):
// kind 0 is the plain FP32 / FP64 kernels (24 configurations), kinds 2..5
// the packed products (split-BF16 with 2 or 3 terms, bf16 and half inputs;
// eh_packed_kind).  The table can be saved to / loaded from a text file
// stamped with the GPU architecture.  It is also the yardstick the learned
// selector is measured against (regret = t(choice)/t(best) - 1).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>
#include <vector>

#include "../../include/eigen_amd.h"
#include "internal.hpp"
#include "kernels.hpp"

namespace eh {

constexpr int kPlainKind = 0;
// Version of the configuration sets the table indexes (the plain kernels'
// 24, the packed kernels' tiles and splits): a table saved under another
// version names other kernels and is refused.
constexpr int kKernelSetVersion = 3;
using ShapeKey = std::tuple<int, int, int, int, int64_t, int64_t, int64_t>;
static std::mutex g_tune_mutex;
static std::map<ShapeKey, int> g_table;

static ShapeKey key_of(int kind, int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k) {
  return ShapeKey{kind, kind == kPlainKind ? dtype : 0, opa != EH_OP_N, opb != EH_OP_N, m, n, k};
}

static int kind_num_configs(int kind) { return kind == kPlainKind ? kNumConfigs : packed_num_configs(kind); }

// One stream for all timing runs (the split-K workspace is cached per stream).
static eh_status tuner_stream(hipStream_t* s) {
  static hipStream_t stream = nullptr;
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  if (!stream) {
    hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      stream = nullptr;
      return hip_fail(e, "tuner stream");
    }
  }
  *s = stream;
  return EH_SUCCESS;
}

// Median time of `reps` runs of configuration `config` (after one warm-up) on
// the caller's buffers, C = A*B with beta = 0; kind as in the table key.
static eh_status time_config(int kind, int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, const void* A,
                             const void* B, void* C, int config, int reps, float* ms) {
  hipStream_t s;
  eh_status sst = tuner_stream(&s);
  if (sst != EH_SUCCESS) return sst;
  hipError_t e;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const double one = 1.0, zero = 0.0;
  const float onef = 1.0f, zerof = 0.0f;
  const void* al = dtype == EH_F32 ? (const void*)&onef : (const void*)&one;
  const void* be = dtype == EH_F32 ? (const void*)&zerof : (const void*)&zero;
  const int64_t lda = opa == EH_OP_N ? std::max<int64_t>(m, 1) : std::max<int64_t>(k, 1);
  const int64_t ldb = opb == EH_OP_N ? std::max<int64_t>(k, 1) : std::max<int64_t>(n, 1);
  std::vector<float> t;
  eh_status st = EH_SUCCESS;
  for (int r = 0; r <= reps && st == EH_SUCCESS; ++r) {
    hipEventRecord(e0, s);
    if (kind == kPlainKind)
      st = gemm_dispatch(dtype, opa, opb, m, n, k, al, A, lda, B, ldb, be, C, std::max<int64_t>(m, 1), EH_PREC_PLAIN,
                         config, s);
    else
      st = packed_product(kind, opa, opb, m, n, k, A, B, C, config, s);
    hipEventRecord(e1, s);
    if (st != EH_SUCCESS) break;
    e = hipEventSynchronize(e1);
    if (e != hipSuccess) {
      st = hip_fail(e, "tuner sync");
      break;
    }
    float v = 0.0f;
    hipEventElapsedTime(&v, e0, e1);
    if (r > 0) t.push_back(v);  // first run is the warm-up
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (st != EH_SUCCESS) return st;
  std::sort(t.begin(), t.end());
  *ms = t.empty() ? 0.0f : t[t.size() / 2];
  return EH_SUCCESS;
}

// Tune one shape on scratch buffers (used by the table mode on a miss):
// uniform FP32 / FP64 operands, or 16-bit operands of value one.
static eh_status tune_shape(int kind, int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, int* best) {
  const bool half = kind == EH_PACKED_BF16 || kind == EH_PACKED_F16;
  const size_t esz = half ? 2 : (dtype == EH_F64 ? 8 : 4), csz = dtype == EH_F64 ? 8 : 4;
  void *A = nullptr, *B = nullptr, *C = nullptr;
  hipError_t e = hipMalloc(&A, std::max<int64_t>(m * k, 1) * esz);
  if (e == hipSuccess) e = hipMalloc(&B, std::max<int64_t>(k * n, 1) * esz);
  if (e == hipSuccess) e = hipMalloc(&C, std::max<int64_t>(m * n, 1) * csz);
  eh_status st = e == hipSuccess ? EH_SUCCESS : fail(EH_ERR_ALLOC, "tuner scratch");
  if (half) {
    const unsigned short one = kind == EH_PACKED_BF16 ? 0x3F80 : 0x3C00;
    if (st == EH_SUCCESS) e = hipMemsetD16(static_cast<hipDeviceptr_t>(A), one, (size_t)std::max<int64_t>(m * k, 1));
    if (st == EH_SUCCESS && e == hipSuccess)
      e = hipMemsetD16(static_cast<hipDeviceptr_t>(B), one, (size_t)std::max<int64_t>(k * n, 1));
    if (st == EH_SUCCESS && e != hipSuccess) st = hip_fail(e, "tuner scratch fill");
  } else {
    if (st == EH_SUCCESS) st = eh_fill((eh_dtype)dtype, A, m * k, 1, 0, -1.0, 1.0, nullptr);
    if (st == EH_SUCCESS) st = eh_fill((eh_dtype)dtype, B, k * n, 2, 0, -1.0, 1.0, nullptr);
  }
  if (st == EH_SUCCESS) {  // the fills run on the null stream, the timing on the tuner's own
    e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) st = hip_fail(e, "tuner scratch sync");
  }
  float best_ms = 0.0f;
  *best = -1;
  const int nc = kind_num_configs(kind);
  for (int c = 0; c < nc && st == EH_SUCCESS; ++c) {
    float ms = 0.0f;
    st = time_config(kind, dtype, opa, opb, m, n, k, A, B, C, c, 2, &ms);
    if (st == EH_SUCCESS && (*best < 0 || ms < best_ms)) {
      best_ms = ms;
      *best = c;
    }
  }
  hipFree(A);
  hipFree(B);
  hipFree(C);
  return st;
}

static int table_plan_kind(int kind, int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k) {
  {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    auto it = g_table.find(key_of(kind, dtype, opa, opb, m, n, k));
    if (it != g_table.end()) return it->second;
  }
  int best = -1;
  if (fixed_scratch_active()) return -1;  // tuning allocates: the caller's planner fallback decides
  if (tune_shape(kind, dtype, opa, opb, m, n, k, &best) != EH_SUCCESS || best < 0) return -1;
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  g_table[key_of(kind, dtype, opa, opb, m, n, k)] = best;
  return best;
}

int table_plan(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k) {
  return table_plan_kind(kPlainKind, dtype, opa, opb, m, n, k);
}

int table_plan_packed(int kind, int opa, int opb, int64_t m, int64_t n, int64_t k) {
  return table_plan_kind(kind, EH_F32, opa, opb, m, n, k);
}

// Architecture name of the current device ("gfx950"), for the table stamp.
static std::string device_arch() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
  std::string a = prop.gcnArchName;
  const size_t colon = a.find(':');  // drop feature suffixes (":sramecc+:xnack-")
  return colon == std::string::npos ? a : a.substr(0, colon);
}

}  // namespace eh

using namespace eh;

extern "C" {

eh_status eh_time_config(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, const void* A, const void* B,
                         void* C, int config, int reps, float* ms) {
  if (!ms) return fail(EH_ERR_INVALID, "ms is NULL");
  if (config < 0 || config >= kNumConfigs) return fail(EH_ERR_INVALID, "config out of range");
  if (reps < 1) return fail(EH_ERR_INVALID, "reps must be >= 1");
  return time_config(kPlainKind, dtype, opa, opb, m, n, k, A, B, C, config, reps, ms);
}

eh_status eh_time_config_packed(eh_packed_kind kind, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k,
                                const void* A, const void* B, void* C, int config, int reps, float* ms) {
  if (!packed_kind_valid(kind)) return fail(EH_ERR_INVALID, "bad packed kind");
  if (!ms) return fail(EH_ERR_INVALID, "ms is NULL");
  if (config < 0 || config >= packed_num_configs(kind)) return fail(EH_ERR_INVALID, "config out of range");
  if (reps < 1) return fail(EH_ERR_INVALID, "reps must be >= 1");
  return time_config(kind, EH_F32, opa, opb, m, n, k, A, B, C, config, reps, ms);
}

int eh_num_packed_configs(eh_packed_kind kind) { return packed_kind_valid(kind) ? packed_num_configs(kind) : 0; }

eh_status eh_packed_config_info(eh_packed_kind kind, int config, int* tile_m, int* tile_n, int* split_k) {
  if (!packed_kind_valid(kind)) return fail(EH_ERR_INVALID, "bad packed kind");
  if (config < 0 || config >= packed_num_configs(kind)) return fail(EH_ERR_INVALID, "config out of range");
  int tile, s;
  config_decode(config, &tile, &s);
  const bool one = kind == EH_PACKED_BF16 || kind == EH_PACKED_F16;
  if (tile_m) *tile_m = tile == 0 ? 128 : 256;
  if (tile_n) *tile_n = tile == 0 ? 128 : (one || tile == 3 ? 256 : 128);
  if (split_k) *split_k = s;
  return EH_SUCCESS;
}

eh_status eh_plan_packed(eh_packed_kind kind, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, int* config) {
  if (!config) return fail(EH_ERR_INVALID, "config is NULL");
  if (!packed_kind_valid(kind)) return fail(EH_ERR_INVALID, "bad packed kind");
  if (m < 0 || n < 0 || k < 0) return fail(EH_ERR_INVALID, "m, n, k must be >= 0");
  *config = plan_packed(kind, opa, opb, m, n, k);
  return EH_SUCCESS;
}

eh_status eh_tuner_record(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, int config) {
  if (config < 0 || config >= kNumConfigs) return fail(EH_ERR_INVALID, "config out of range");
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  g_table[key_of(kPlainKind, dtype, opa, opb, m, n, k)] = config;
  return EH_SUCCESS;
}

eh_status eh_tuner_record_packed(eh_packed_kind kind, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k,
                                 int config) {
  if (!packed_kind_valid(kind)) return fail(EH_ERR_INVALID, "bad packed kind");
  if (config < 0 || config >= packed_num_configs(kind)) return fail(EH_ERR_INVALID, "config out of range");
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  g_table[key_of(kind, EH_F32, opa, opb, m, n, k)] = config;
  return EH_SUCCESS;
}

int64_t eh_tuner_count(void) {
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  return (int64_t)g_table.size();
}

eh_status eh_tuner_clear(void) {
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  g_table.clear();
  return EH_SUCCESS;
}

eh_status eh_tuner_save(const char* path) {
  if (!path) return fail(EH_ERR_INVALID, "path is NULL");
  const std::string arch = device_arch();
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(EH_ERR_INVALID, std::string("cannot write ") + path);
  std::fprintf(f, "# eigen_amd tuning table v2 arch=%s kernels=%d\n# kind dtype opa opb m n k config\n", arch.c_str(),
               kKernelSetVersion);
  for (const auto& kv : g_table)
    std::fprintf(f, "%d %d %d %d %lld %lld %lld %d\n", std::get<0>(kv.first), std::get<1>(kv.first),
                 std::get<2>(kv.first), std::get<3>(kv.first), (long long)std::get<4>(kv.first),
                 (long long)std::get<5>(kv.first), (long long)std::get<6>(kv.first), kv.second);
  std::fclose(f);
  return EH_SUCCESS;
}

eh_status eh_tuner_load(const char* path) {
  if (!path) return fail(EH_ERR_INVALID, "path is NULL");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(EH_ERR_INVALID, std::string("cannot read ") + path);
  std::map<ShapeKey, int> loaded;
  char line[256];
  int lineno = 0;
  auto bad = [&](const std::string& why) {
    std::fclose(f);
    return fail(EH_ERR_INVALID, std::string(path) + " line " + std::to_string(lineno) + ": " + why);
  };
  while (std::fgets(line, sizeof line, f)) {
    ++lineno;
    if (line[0] == '#') {
      if (const char* a = std::strstr(line, "arch=")) {
        std::string arch(a + 5);
        arch = arch.substr(0, arch.find_first_of(" \r\n"));
        if (arch != device_arch()) return bad("table made for " + arch + ", this device is " + device_arch());
      }
      if (const char* kv = std::strstr(line, "kernels=")) {
        const int v = std::atoi(kv + 8);
        if (v != kKernelSetVersion)
          return bad("table made for kernel set " + std::to_string(v) + ", this library has " +
                     std::to_string(kKernelSetVersion));
      }
      continue;
    }
    if (line[0] == '\n') continue;
    long long v[8];
    int cnt = 0;
    std::istringstream in(line);
    while (cnt < 8 && (in >> v[cnt])) ++cnt;
    std::string rest;
    if (in >> rest) return bad("malformed");
    int kind, d, a, b, c;
    long long m, n, k;
    if (cnt == 7) {  // v1: plain kernels
      kind = kPlainKind;
      d = (int)v[0], a = (int)v[1], b = (int)v[2], m = v[3], n = v[4], k = v[5], c = (int)v[6];
    } else if (cnt == 8) {
      kind = (int)v[0], d = (int)v[1], a = (int)v[2], b = (int)v[3], m = v[4], n = v[5], k = v[6], c = (int)v[7];
    } else {
      return bad("malformed");
    }
    if ((kind != kPlainKind && !packed_kind_valid(kind)) || c < 0 || c >= kind_num_configs(kind) ||
        (d != EH_F32 && d != EH_F64) || m < 0 || n < 0 || k < 0)
      return bad("malformed");
    loaded[key_of(kind, d, a, b, m, n, k)] = c;
  }
  std::fclose(f);
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  for (const auto& kv : loaded) g_table[kv.first] = kv.second;
  return EH_SUCCESS;
}

}  // extern "C"
