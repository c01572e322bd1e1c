// Helpers shared by the Eigen drop-in test programs: device buffers and exact
// small-integer test matrices (every partial sum of their products is exact in
// FP32 and FP64, so device and host results must agree bit for bit).
#pragma once
#include <hip/hip_runtime.h>

#include <Eigen/Core>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK_HIP(x)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      std::fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      std::exit(2);                                                                      \
    }                                                                                    \
  } while (0)

template <typename T>
struct DeviceBuffer {
  T* ptr = nullptr;
  size_t n = 0;
  explicit DeviceBuffer(size_t count) : n(count) { CHECK_HIP(hipMalloc(&ptr, sizeof(T) * (count ? count : 1))); }
  ~DeviceBuffer() { (void)hipFree(ptr); }
  void upload(const T* h) { CHECK_HIP(hipMemcpy(ptr, h, sizeof(T) * n, hipMemcpyHostToDevice)); }
  void download(T* h) const { CHECK_HIP(hipMemcpy(h, ptr, sizeof(T) * n, hipMemcpyDeviceToHost)); }
};

// Deterministic integers in [-8, 8].
template <typename Mat>
void fill_int(Mat& m, unsigned seed) {
  unsigned s = seed * 2654435761u + 12345u;
  for (Eigen::Index j = 0; j < m.cols(); ++j)
    for (Eigen::Index i = 0; i < m.rows(); ++i) {
      s = s * 1664525u + 1013904223u;
      m(i, j) = typename Mat::Scalar(int((s >> 16) % 17) - 8);
    }
}

// One line per case, flushed, so that the cases run so far are on record
// even if the program dies.
inline int report(const char* name, bool ok) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", name);
  std::fflush(stdout);
  return ok ? 0 : 1;
}
