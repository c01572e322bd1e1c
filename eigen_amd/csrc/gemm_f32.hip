// gemm_f32.hip -- FP32 instantiations of the plain MFMA product and its launchers.
#include "gemm_mfma.hpp"

namespace eh {

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T, int BM, int BN, int WGM, int WGN, int RS, bool TA, bool TB>
static hipError_t launch_one(const GemmParams<T>& p, int splits, hipStream_t s) {
  dim3 grid((unsigned)grid_tiles<BM, BN>(p), splits);
  hipLaunchKernelGGL((gemm_mfma_kernel<T, BM, BN, WGM, WGN, TA, TB, RS>), grid, dim3(64 * WGM * WGN), 0, s, p);
  return hipGetLastError();
}

// RS: register stages of the global loads (2 where the registers fit
// without lowering occupancy)
template <typename T, int BM, int BN, int WGM, int WGN, int RS>
static hipError_t launch_tile(const GemmParams<T>& p, bool ta, bool tb, int splits, hipStream_t s) {
  if (!ta && !tb) return launch_one<T, BM, BN, WGM, WGN, RS, false, false>(p, splits, s);
  if (ta && !tb) return launch_one<T, BM, BN, WGM, WGN, RS, true, false>(p, splits, s);
  if (!ta && tb) return launch_one<T, BM, BN, WGM, WGN, RS, false, true>(p, splits, s);
  return launch_one<T, BM, BN, WGM, WGN, RS, true, true>(p, splits, s);
}

template <typename T>
hipError_t launch_gemm_mfma(const GemmParams<T>& p, int tile, bool ta, bool tb, int splits, hipStream_t s) {
  switch (tile) {
    case 0: return launch_tile<T, 64, 64, 2, 2, 2>(p, ta, tb, splits, s);
    case 1: return launch_tile<T, 64, 128, 2, 2, 2>(p, ta, tb, splits, s);
    case 2: return launch_tile<T, 128, 64, 2, 2, 2>(p, ta, tb, splits, s);
    case 3: return launch_tile<T, 128, 128, 2, 2, 2>(p, ta, tb, splits, s);
    // 8 waves (two per SIMD), wave tiles of 64x64
    case 4: return launch_tile<T, 128, 256, 2, 4, 2>(p, ta, tb, splits, s);
    case 5: return launch_tile<T, 256, 128, 4, 2, 2>(p, ta, tb, splits, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_splitk_reduce(const GemmParams<T>& p, int splits, hipStream_t s) {
  const int64_t total = (int64_t)p.M * p.N;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(splitk_reduce_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, splits);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale_c(T* C, int M, int N, int64_t ldc, T beta, hipStream_t s, int tri) {
  const int64_t total = (int64_t)M * N;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(scale_c_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, C, M, N, ldc, beta, tri);
  return hipGetLastError();
}

template hipError_t launch_gemm_mfma<float>(const GemmParams<float>&, int, bool, bool, int, hipStream_t);
template hipError_t launch_splitk_reduce<float>(const GemmParams<float>&, int, hipStream_t);
template hipError_t launch_scale_c<float>(float*, int, int, int64_t, float, hipStream_t, int);

template <typename T>
hipError_t launch_strided_combine(const T* W, int M, int N, T* C, int64_t incc, int64_t ldc, T alpha, T beta,
                                  hipStream_t s) {
  const int64_t total = (int64_t)M * N;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(strided_combine_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, W, M, N, C, incc, ldc, alpha,
                     beta);
  return hipGetLastError();
}
template hipError_t launch_strided_combine<float>(const float*, int, int, float*, int64_t, int64_t, float, float, hipStream_t);

}  // namespace eh
