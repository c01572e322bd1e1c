// internal.hpp -- declarations shared by the host-side translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <mutex>
#include <string>

#include "../../include/eigen_amd.h"
#include "cplx.hpp"

namespace eh {

constexpr int kNumTiles = 6;
constexpr int kNumSplits = 4;
constexpr int kNumConfigs = kNumTiles * kNumSplits;  // 24 candidate configurations

eh_status fail(eh_status s, const std::string& msg);
eh_status hip_fail(hipError_t e, const char* what);

void config_decode(int config, int* tile, int* splits);
int config_tile_m(int config);
int config_tile_n(int config);
int heuristic_config(int dtype, int64_t m, int64_t n, int64_t k);
eh_tune_mode tune_mode();

// Planner (planner.cpp): configuration for a shape under the current mode.
eh_status plan_config(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, int* config);

// Packed products (capi.cpp): the split-BF16 FP32 products and the bf16 /
// half input products, kind = eh_packed_kind.  Their configurations are
// tile * kNumSplits + split index over the packed kernel's tiles.
bool packed_kind_valid(int kind);
int packed_num_configs(int kind);
int packed_rule_config(int kind, int64_t m, int64_t n, int64_t k);
int plan_packed(int kind, int opa, int opb, int64_t m, int64_t n, int64_t k);
eh_status packed_product(int kind, int opa, int opb, int64_t m, int64_t n, int64_t k, const void* A, const void* B,
                         void* C, int config, hipStream_t s);

// Per-stream device workspace (capi.cpp).  A lease holds the stream's
// workspace lock and a reference to its buffer: hold it until every launch
// of the call that uses the buffer is queued, so that another thread's call
// on the same stream can neither interleave its launches with ours (both
// would use the one buffer) nor free the buffer under us.
struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
  Workspace() = default;
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
  ~Workspace();
};
struct WorkspaceLease {
  std::unique_lock<std::recursive_mutex> lock;
  std::shared_ptr<Workspace> buf;
};
// *ptr = `bytes` of device workspace for work queued on `stream`.  Normally
// the per-stream workspace (grown with hipMalloc on demand).  While a
// FixedScratch is installed on the calling thread, *ptr is that
// caller-provided buffer instead (no allocation, no lock), and a request
// larger than it fails with EH_ERR_UNSUPPORTED.
eh_status workspace_get(hipStream_t stream, size_t bytes, WorkspaceLease* lease, void** ptr);

// Installs a fixed device buffer as this thread's workspace for its lifetime
// (the allocation-free entry points); while installed, the planner does not
// tune unseen shapes either (tuning allocates).
struct FixedScratch {
  FixedScratch(void* ptr, size_t bytes);
  ~FixedScratch();
  FixedScratch(const FixedScratch&) = delete;
  FixedScratch& operator=(const FixedScratch&) = delete;
  void* prev_ptr;
  size_t prev_bytes;
};
bool fixed_scratch_active();

// Event timing of kernel launches for measurement (eh_kernel_timing): while
// enabled, a KernelTimer records one HIP event on `s` when constructed and
// one when destroyed, around the launch it scopes (not while `s` is being
// captured into a graph).
struct KernelTimer {
  KernelTimer(int cls, hipStream_t s);
  ~KernelTimer();
  KernelTimer(const KernelTimer&) = delete;
  KernelTimer& operator=(const KernelTimer&) = delete;
  int cls;
  hipStream_t s;
  hipEvent_t start = nullptr;
};

// A stream-ordered temporary (hipMallocAsync) released by hipFreeAsync on the
// same stream when the owner goes out of scope, i.e. after the work queued
// on that stream that uses it.
struct StreamTemp {
  void* ptr = nullptr;
  hipStream_t s = nullptr;
  eh_status get(size_t bytes, hipStream_t stream, const char* what) {
    s = stream;
    hipError_t e = hipMallocAsync(&ptr, bytes < 256 ? 256 : bytes, stream);
    if (e != hipSuccess) {
      ptr = nullptr;
      return fail(EH_ERR_ALLOC, std::string(what) + ": " + hipGetErrorString(e));
    }
    return EH_SUCCESS;
  }
  StreamTemp() = default;
  StreamTemp(const StreamTemp&) = delete;
  StreamTemp& operator=(const StreamTemp&) = delete;
  ~StreamTemp() {
    if (ptr) (void)hipFreeAsync(ptr, s);
  }
};
eh_status check_gemm_args(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                          const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, const void* C,
                          int64_t ldc);
// Plain-precision device product C = alpha*op(A)*op(B) + beta*C (operands
// already checked) on the planner's configuration; tri != 0 updates only the
// lower (1) or upper (2) triangle of C.  k == 0 or alpha == 0 only scales C.
template <typename T>
eh_status device_product(int opa, int opb, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda,
                         const T* B, int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t stream, int tri = 0);
// Complex product C = alpha*op(A)*op(B) + beta*C on interleaved storage (op
// C conjugates), as one real product over 2x2 real blocks (complex_gemm.cpp),
// plain FP32 / FP64 MFMA; temporaries from the stream-ordered allocator.
template <typename R>
eh_status complex_product(int opa, int opb, int64_t m, int64_t n, int64_t k, cplx<R> alpha, const cplx<R>* A,
                          int64_t lda, const cplx<R>* B, int64_t ldb, cplx<R> beta, cplx<R>* C, int64_t ldc,
                          hipStream_t s);
eh_status gemm_dispatch(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                        const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                        int64_t ldc, int prec, int config, void* stream);

}  // namespace eh
