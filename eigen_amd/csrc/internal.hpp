// internal.hpp -- declarations shared by the host-side translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/eigen_amd.h"

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

eh_status workspace_get(hipStream_t stream, size_t bytes, void** out);

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
eh_status gemm_dispatch(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                        const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                        int64_t ldc, int prec, int config, void* stream);

}  // namespace eh
