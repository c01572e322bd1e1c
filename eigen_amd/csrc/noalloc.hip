// noalloc.hip -- eh_gemm_noalloc: eh_gemm with no device allocation.
//
// The product kernels keep split-K partial sums, packed bf16 operand planes
// and matrix-vector partial sums in a device workspace that eh_gemm takes
// from the per-stream pool (grown with hipMalloc on demand).  Callers that
// cannot allocate -- code under graph capture, serving loops, harnesses that
// forbid it -- use this entry point instead: the same dispatch and kernels,
// with the workspace replaced by one fixed device array of kScratchBytes per
// GPU, compiled into the library and resident for the life of the process.
// Requests that need more scratch than that fail with EH_ERR_UNSUPPORTED
// before anything is launched.  The array is shared by every call on the
// GPU: calls that overlap in time on one GPU race on it (the caller
// serialises them, e.g. by using one stream).  Nothing in it needs to start
// at zero: every kernel writes each scratch element it later reads.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/eigen_amd.h"
#include "internal.hpp"

namespace eh {

constexpr size_t kScratchBytes = (size_t)EH_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device, looked up once per device with
// that device current (each device holds its own copy of the array).  In
// host code the name `scratch` is only a host-side handle for the symbol;
// kernels and copies get these addresses.
constexpr int kMaxDevices = 64;
static void* g_scratch_addr[kMaxDevices];
static std::mutex g_scratch_mutex;

static eh_status scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(EH_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return EH_SUCCESS;
}

}  // namespace eh

using namespace eh;

extern "C" eh_status eh_gemm_noalloc(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k,
                                     const void* alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                                     const void* beta, void* C, int64_t ldc, eh_precision prec, void* stream) {
  void* ptr = nullptr;
  eh_status st = scratch_address(&ptr);
  if (st != EH_SUCCESS) return st;
  FixedScratch fixed(ptr, kScratchBytes);
  return gemm_dispatch(dtype, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, prec, -1, stream);
}
