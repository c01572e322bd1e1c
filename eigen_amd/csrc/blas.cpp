// blas.cpp -- sgemm_ / dgemm_ and the other real level-3 routines (?gemmtr_,
// ?syrk_, ?syr2k_, ?symm_, ?trmm_, ?trsm_) with the semantics of Eigen's BLAS
// (blas/level3_impl.h:12-758) on top of the device product.
//
// The reference validates in a fixed order and reports through xerbla_
// (:61-78), returns early for m == 0 or n == 0 (:80), pre-scales C by beta
// (:82-87), returns for k == 0 (:89) and then runs the GEBP product, which
// itself returns immediately for alpha == 0 (GeneralMatrixMatrix.h:130).
// Here beta is fused into the kernel epilogue with the same observable
// result: beta == 0 overwrites C without reading it, alpha == 0 or k == 0
// leaves only beta*C and never touches A or B.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/eigen_amd.h"
#include "../../include/eigen_amd_blas.h"

namespace {

enum { OP_N = 0, OP_T = 1, OP_C = 2, OP_BAD = 0xff };

int op_code(char x) {
  if (x == 'N' || x == 'n') return OP_N;
  if (x == 'T' || x == 't') return OP_T;
  if (x == 'C' || x == 'c') return OP_C;
  return OP_BAD;
}

hipStream_t g_stream = nullptr;
std::mutex g_stage_mutex;

// Grow-only device staging buffers for host operands.
struct Stage {
  void* ptr = nullptr;
  size_t bytes = 0;
  bool ensure(size_t want) {
    if (bytes >= want) return true;
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (hipMalloc(&ptr, want) != hipSuccess) {
      ptr = nullptr;
      return false;
    }
    bytes = want;
    return true;
  }
};
Stage g_sa, g_sb, g_sc;

bool is_device_ptr(const void* p) {
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory: clear the error
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

void report(const char* what) { std::fprintf(stderr, "eigen_amd BLAS: %s: %s\n", what, eh_last_error()); }

// dtype of an element type: float and complex<float> run as EH_F32,
// double and complex<double> as EH_F64
template <typename T>
struct DtOf {
  static constexpr eh_dtype value = sizeof(T) == 4 ? EH_F32 : EH_F64;
};
template <typename R>
struct DtOf<std::complex<R>> {
  static constexpr eh_dtype value = sizeof(R) == 4 ? EH_F32 : EH_F64;
};
template <typename T>
constexpr eh_dtype dtype_of() {
  return DtOf<T>::value;
}

// C-ABI signatures the level-3 wrappers dispatch to (real or complex entry)
using SymmFn = eh_status (*)(eh_dtype, eh_side, eh_uplo, int64_t, int64_t, const void*, const void*, int64_t,
                             const void*, int64_t, const void*, void*, int64_t, void*);
using TriFn = eh_status (*)(eh_dtype, eh_side, eh_uplo, eh_op, eh_diag, int64_t, int64_t, const void*, const void*,
                            int64_t, void*, int64_t, void*);
using Rank2Fn = eh_status (*)(eh_dtype, eh_uplo, eh_op, int64_t, int64_t, const void*, const void*, int64_t,
                              const void*, int64_t, const void*, void*, int64_t, void*);
using RankFn = eh_status (*)(eh_dtype, eh_uplo, eh_op, int64_t, int64_t, const void*, const void*, int64_t,
                             const void*, void*, int64_t, void*);
using GemmtrFn = eh_status (*)(eh_dtype, eh_uplo, eh_op, eh_op, int64_t, int64_t, const void*, const void*, int64_t,
                               const void*, int64_t, const void*, void*, int64_t, void*);

template <typename T>
void gemm_impl(const char* name, const char* opa, const char* opb, const int* m, const int* n, const int* k,
               const T* palpha, const T* a, const int* lda, const T* b, const int* ldb, const T* pbeta, T* c,
               const int* ldc) {
  int info = 0;
  const int oa = op_code(*opa), ob = op_code(*opb);
  if (oa == OP_BAD)
    info = 1;
  else if (ob == OP_BAD)
    info = 2;
  else if (*m < 0)
    info = 3;
  else if (*n < 0)
    info = 4;
  else if (*k < 0)
    info = 5;
  else if (*lda < std::max(1, oa == OP_N ? *m : *k))
    info = 8;
  else if (*ldb < std::max(1, ob == OP_N ? *k : *n))
    info = 10;
  else if (*ldc < std::max(1, *m))
    info = 13;
  if (info) {
    xerbla_(name, &info, 6);
    return;
  }
  if (*m == 0 || *n == 0) return;

  const eh_dtype dt = sizeof(T) == 4 ? EH_F32 : EH_F64;
  const T alpha = *palpha, beta = *pbeta;
  const bool need_ab = (*k > 0) && alpha != T(0);
  const bool need_c_in = beta != T(0);
  const int arows = oa == OP_N ? *m : *k, acols = oa == OP_N ? *k : *m;
  const int brows = ob == OP_N ? *k : *n, bcols = ob == OP_N ? *n : *k;

  const bool dev_a = !need_ab || is_device_ptr(a);
  const bool dev_b = !need_ab || is_device_ptr(b);
  const bool dev_c = is_device_ptr(c);
  hipStream_t s = g_stream;

  if (dev_a && dev_b && dev_c) {
    eh_status st = eh_gemm(dt, (eh_op)oa, (eh_op)ob, *m, *n, *k, &alpha, a, *lda, b, *ldb, &beta, c, *ldc,
                           EH_PREC_AUTO, s);
    if (st != EH_SUCCESS) return report(name);
    if (hipStreamSynchronize(s) != hipSuccess) std::fprintf(stderr, "eigen_amd BLAS: stream sync failed\n");
    return;
  }

  // Host operands: stage them into packed device buffers (ld = rows).
  std::lock_guard<std::mutex> lk(g_stage_mutex);
  const T* da = a;
  const T* db = b;
  T* dc = c;
  int64_t dlda = *lda, dldb = *ldb, dldc = *ldc;
  auto copy_in = [&](Stage& st, const T* host, int rows, int cols, int ld) -> T* {
    if (!st.ensure((size_t)std::max(rows, 1) * std::max(cols, 1) * sizeof(T))) return nullptr;
    if (rows > 0 && cols > 0 &&
        hipMemcpy2DAsync(st.ptr, rows * sizeof(T), host, (size_t)ld * sizeof(T), rows * sizeof(T), cols,
                         hipMemcpyHostToDevice, s) != hipSuccess)
      return nullptr;
    return static_cast<T*>(st.ptr);
  };
  if (need_ab && !dev_a) {
    da = copy_in(g_sa, a, arows, acols, *lda);
    dlda = std::max(arows, 1);
    if (!da) return (void)std::fprintf(stderr, "eigen_amd BLAS: staging A failed\n");
  }
  if (need_ab && !dev_b) {
    db = copy_in(g_sb, b, brows, bcols, *ldb);
    dldb = std::max(brows, 1);
    if (!db) return (void)std::fprintf(stderr, "eigen_amd BLAS: staging B failed\n");
  }
  if (!dev_c) {
    if (!g_sc.ensure((size_t)(*m) * (*n) * sizeof(T))) return (void)std::fprintf(stderr, "eigen_amd BLAS: staging C failed\n");
    dc = static_cast<T*>(g_sc.ptr);
    dldc = *m;
    if (need_c_in && hipMemcpy2DAsync(dc, (size_t)(*m) * sizeof(T), c, (size_t)(*ldc) * sizeof(T),
                                      (size_t)(*m) * sizeof(T), *n, hipMemcpyHostToDevice, s) != hipSuccess)
      return (void)std::fprintf(stderr, "eigen_amd BLAS: staging C failed\n");
  }
  eh_status st = eh_gemm(dt, (eh_op)oa, (eh_op)ob, *m, *n, *k, &alpha, da, dlda, db, dldb, &beta, dc, dldc,
                         EH_PREC_AUTO, s);
  if (st != EH_SUCCESS) return report(name);
  if (!dev_c && hipMemcpy2DAsync(c, (size_t)(*ldc) * sizeof(T), dc, (size_t)(*m) * sizeof(T),
                                 (size_t)(*m) * sizeof(T), *n, hipMemcpyDeviceToHost, s) != hipSuccess)
    std::fprintf(stderr, "eigen_amd BLAS: copying C back failed\n");
  if (hipStreamSynchronize(s) != hipSuccess) std::fprintf(stderr, "eigen_amd BLAS: stream sync failed\n");
}

// ?syrk_ with the semantics of blas/level3_impl.h:594-689: validation order
// and info codes 1 (uplo), 2 (trans), 3 (n), 4 (k), 7 (lda), 10 (ldc); only
// the uplo triangle of C is read or written.
// Also complex ?syrk_ (trans C invalid) and ?herk_ (blas/level3_impl.h:835-912:
// alpha and beta real, trans T invalid): element type T, scalar type S.
template <typename T, typename S = T>
void syrk_impl(const char* name, const char* uplo, const char* op, const int* n, const int* k, const S* palpha,
               const T* a, const int* lda, const S* pbeta, T* c, const int* ldc, RankFn fn = eh_syrk,
               int bad_op = OP_BAD) {
  int info = 0;
  const int ul = (*uplo == 'U' || *uplo == 'u') ? EH_UPPER : (*uplo == 'L' || *uplo == 'l') ? EH_LOWER : -1;
  const int o = op_code(*op);
  if (ul < 0)
    info = 1;
  else if (o == OP_BAD || o == bad_op)
    info = 2;
  else if (*n < 0)
    info = 3;
  else if (*k < 0)
    info = 4;
  else if (*lda < std::max(1, o == OP_N ? *n : *k))
    info = 7;
  else if (*ldc < std::max(1, *n))
    info = 10;
  if (info) {
    xerbla_(name, &info, 6);
    return;
  }
  const S alpha = *palpha, beta = *pbeta;
  const bool need_a = *k > 0 && alpha != S(0);
  if (*n == 0 || (!need_a && beta == S(1))) return;

  const eh_dtype dt = dtype_of<T>();
  const int arows = o == OP_N ? *n : *k, acols = o == OP_N ? *k : *n;
  const bool dev_a = !need_a || is_device_ptr(a);
  const bool dev_c = is_device_ptr(c);
  hipStream_t s = g_stream;
  std::unique_lock<std::mutex> lk(g_stage_mutex, std::defer_lock);
  const T* da = a;
  T* dc = c;
  int64_t dlda = *lda, dldc = *ldc;
  if (!dev_a || !dev_c) lk.lock();
  if (!dev_a) {
    if (!g_sa.ensure((size_t)std::max(arows, 1) * std::max(acols, 1) * sizeof(T)) ||
        hipMemcpy2DAsync(g_sa.ptr, arows * sizeof(T), a, (size_t)(*lda) * sizeof(T), arows * sizeof(T), acols,
                         hipMemcpyHostToDevice, s) != hipSuccess)
      return (void)std::fprintf(stderr, "eigen_amd BLAS: staging A failed\n");
    da = static_cast<const T*>(g_sa.ptr);
    dlda = std::max(arows, 1);
  }
  if (!dev_c) {  // the whole of C travels, so the other triangle comes back unchanged
    if (!g_sc.ensure((size_t)(*n) * (*n) * sizeof(T)) ||
        hipMemcpy2DAsync(g_sc.ptr, (size_t)(*n) * sizeof(T), c, (size_t)(*ldc) * sizeof(T), (size_t)(*n) * sizeof(T),
                         *n, hipMemcpyHostToDevice, s) != hipSuccess)
      return (void)std::fprintf(stderr, "eigen_amd BLAS: staging C failed\n");
    dc = static_cast<T*>(g_sc.ptr);
    dldc = *n;
  }
  eh_status st = fn(dt, (eh_uplo)ul, (eh_op)o, *n, *k, &alpha, da, dlda, &beta, dc, dldc, s);
  if (st != EH_SUCCESS) return report(name);
  if (!dev_c && hipMemcpy2DAsync(c, (size_t)(*ldc) * sizeof(T), dc, (size_t)(*n) * sizeof(T),
                                 (size_t)(*n) * sizeof(T), *n, hipMemcpyDeviceToHost, s) != hipSuccess)
    std::fprintf(stderr, "eigen_amd BLAS: copying C back failed\n");
  if (hipStreamSynchronize(s) != hipSuccess) std::fprintf(stderr, "eigen_amd BLAS: stream sync failed\n");
}

// ---- the other level-3 routines ---------------------------------------------
int side_code(char x) { return (x == 'L' || x == 'l') ? EH_LEFT : (x == 'R' || x == 'r') ? EH_RIGHT : -1; }
int uplo_code(char x) { return (x == 'U' || x == 'u') ? EH_UPPER : (x == 'L' || x == 'l') ? EH_LOWER : -1; }
int diag_code(char x) { return (x == 'N' || x == 'n') ? EH_NON_UNIT : (x == 'U' || x == 'u') ? EH_UNIT : -1; }

// Device view of one rows x cols matrix argument: the pointer itself when it
// is device memory, else a packed copy (ld = rows) in `st`, filled from the
// host when `in`; copy_back returns a staged copy to the host.
// Arguments the routine does not read (alpha == 0) keep the caller's
// pointer and leading dimension.
template <typename T>
struct DevMat {
  DevMat(const T* p_, int ld_) : p(const_cast<T*>(p_)), ld(ld_) {}
  T* p;
  int64_t ld;
  bool staged = false;
  int rows = 0, cols = 0;
  int64_t host_ld = 0;
};

template <typename T>
bool dev_view(Stage& st, const T* host, int rows, int cols, int ld, bool in, hipStream_t s, DevMat<T>* d) {
  d->rows = rows;
  d->cols = cols;
  d->host_ld = ld;
  if (rows <= 0 || cols <= 0 || is_device_ptr(host)) {
    d->p = const_cast<T*>(host);
    d->ld = ld;
    return true;
  }
  if (!st.ensure((size_t)rows * cols * sizeof(T))) return false;
  d->p = static_cast<T*>(st.ptr);
  d->ld = rows;
  d->staged = true;
  return !in || hipMemcpy2DAsync(st.ptr, rows * sizeof(T), host, (size_t)ld * sizeof(T), rows * sizeof(T), cols,
                                 hipMemcpyHostToDevice, s) == hipSuccess;
}

template <typename T>
bool copy_back(const DevMat<T>& d, T* host, hipStream_t s) {
  return !d.staged || hipMemcpy2DAsync(host, (size_t)d.host_ld * sizeof(T), d.p, d.ld * sizeof(T),
                                       d.rows * sizeof(T), d.cols, hipMemcpyDeviceToHost, s) == hipSuccess;
}

void finish(hipStream_t s) {
  if (hipStreamSynchronize(s) != hipSuccess) std::fprintf(stderr, "eigen_amd BLAS: stream sync failed\n");
}

// ?gemmtr_ with the checks of blas/level3_impl.h:181-197: info 1 uplo,
// 2 transa, 3 transb, 4 n, 5 k, 8 lda, 10 ldb, 13 ldc.
template <typename T>
void gemmtr_impl(const char* name, const char* uplo, const char* opa, const char* opb, const int* n, const int* k,
                 const T* palpha, const T* a, const int* lda, const T* b, const int* ldb, const T* pbeta, T* c,
                 const int* ldc, GemmtrFn fn = eh_gemmtr) {
  int info = 0;
  const int ul = uplo_code(*uplo), oa = op_code(*opa), ob = op_code(*opb);
  if (ul < 0)
    info = 1;
  else if (oa == OP_BAD)
    info = 2;
  else if (ob == OP_BAD)
    info = 3;
  else if (*n < 0)
    info = 4;
  else if (*k < 0)
    info = 5;
  else if (*lda < std::max(1, oa == OP_N ? *n : *k))
    info = 8;
  else if (*ldb < std::max(1, ob == OP_N ? *k : *n))
    info = 10;
  else if (*ldc < std::max(1, *n))
    info = 13;
  if (info) return xerbla_(name, &info, 7);
  if (*n == 0) return;
  const T alpha = *palpha, beta = *pbeta;
  const bool need_ab = *k > 0 && alpha != T(0);
  if (!need_ab && beta == T(1)) return;
  hipStream_t s = g_stream;
  std::lock_guard<std::mutex> lk(g_stage_mutex);
  DevMat<T> A(a, *lda), B(b, *ldb), C(c, *ldc);
  if ((need_ab && (!dev_view(g_sa, a, oa == OP_N ? *n : *k, oa == OP_N ? *k : *n, *lda, true, s, &A) ||
                   !dev_view(g_sb, b, ob == OP_N ? *k : *n, ob == OP_N ? *n : *k, *ldb, true, s, &B))) ||
      !dev_view(g_sc, c, *n, *n, *ldc, true, s, &C))  // all of C travels: the other triangle returns unchanged
    return (void)std::fprintf(stderr, "eigen_amd BLAS: staging failed\n");
  if (fn(dtype_of<T>(), (eh_uplo)ul, (eh_op)oa, (eh_op)ob, *n, *k, &alpha, A.p, A.ld, B.p, B.ld, &beta, C.p, C.ld,
         s) != EH_SUCCESS)
    return report(name);
  if (!copy_back(C, c, s)) std::fprintf(stderr, "eigen_amd BLAS: copying C back failed\n");
  finish(s);
}

// ?syr2k_ with the checks of blas/level3_impl.h:702-716: info 1 uplo,
// 2 trans, 3 n, 4 k, 7 lda, 9 ldb, 12 ldc; also ?her2k_ (:915-942, beta real,
// trans T invalid) and complex ?syr2k_ (trans C invalid): `bad_op`.
template <typename T, typename SB = T>
void syr2k_impl(const char* name, const char* uplo, const char* op, const int* n, const int* k, const T* palpha,
                const T* a, const int* lda, const T* b, const int* ldb, const SB* pbeta, T* c, const int* ldc,
                Rank2Fn fn = eh_syr2k, int bad_op = OP_BAD) {
  int info = 0;
  const int ul = uplo_code(*uplo), o = op_code(*op);
  const int rows = o == OP_N ? *n : *k, cols = o == OP_N ? *k : *n;
  if (ul < 0)
    info = 1;
  else if (o == OP_BAD || o == bad_op)
    info = 2;
  else if (*n < 0)
    info = 3;
  else if (*k < 0)
    info = 4;
  else if (*lda < std::max(1, rows))
    info = 7;
  else if (*ldb < std::max(1, rows))
    info = 9;
  else if (*ldc < std::max(1, *n))
    info = 12;
  if (info) return xerbla_(name, &info, 6);
  if (*n == 0) return;
  const T alpha = *palpha;
  const SB beta = *pbeta;
  const bool need_ab = *k > 0 && alpha != T(0);
  if (!need_ab && beta == SB(1)) return;
  hipStream_t s = g_stream;
  std::lock_guard<std::mutex> lk(g_stage_mutex);
  DevMat<T> A(a, *lda), B(b, *ldb), C(c, *ldc);
  if ((need_ab && (!dev_view(g_sa, a, rows, cols, *lda, true, s, &A) ||
                   !dev_view(g_sb, b, rows, cols, *ldb, true, s, &B))) ||
      !dev_view(g_sc, c, *n, *n, *ldc, true, s, &C))
    return (void)std::fprintf(stderr, "eigen_amd BLAS: staging failed\n");
  if (fn(dtype_of<T>(), (eh_uplo)ul, (eh_op)o, *n, *k, &alpha, A.p, A.ld, B.p, B.ld, &beta, C.p, C.ld, s) !=
      EH_SUCCESS)
    return report(name);
  if (!copy_back(C, c, s)) std::fprintf(stderr, "eigen_amd BLAS: copying C back failed\n");
  finish(s);
}

// ?symm_ with the checks of blas/level3_impl.h:514-532: info 1 side,
// 2 uplo, 3 m, 4 n, 7 lda, 9 ldb, 12 ldc.
template <typename T>
void symm_impl(const char* name, const char* side, const char* uplo, const int* m, const int* n, const T* palpha,
               const T* a, const int* lda, const T* b, const int* ldb, const T* pbeta, T* c, const int* ldc,
               SymmFn fn = eh_symm) {
  int info = 0;
  const int sd = side_code(*side), ul = uplo_code(*uplo);
  const int ka = sd == EH_LEFT ? *m : *n;
  if (sd < 0)
    info = 1;
  else if (ul < 0)
    info = 2;
  else if (*m < 0)
    info = 3;
  else if (*n < 0)
    info = 4;
  else if (*lda < std::max(1, ka))
    info = 7;
  else if (*ldb < std::max(1, *m))
    info = 9;
  else if (*ldc < std::max(1, *m))
    info = 12;
  if (info) return xerbla_(name, &info, 6);
  if (*m == 0 || *n == 0) return;
  const T alpha = *palpha, beta = *pbeta;
  if (alpha == T(0) && beta == T(1)) return;
  hipStream_t s = g_stream;
  std::lock_guard<std::mutex> lk(g_stage_mutex);
  DevMat<T> A(a, *lda), B(b, *ldb), C(c, *ldc);
  const bool need_ab = alpha != T(0);
  if ((need_ab && (!dev_view(g_sa, a, ka, ka, *lda, true, s, &A) ||
                   !dev_view(g_sb, b, *m, *n, *ldb, true, s, &B))) ||
      !dev_view(g_sc, c, *m, *n, *ldc, beta != T(0), s, &C))
    return (void)std::fprintf(stderr, "eigen_amd BLAS: staging failed\n");
  if (fn(dtype_of<T>(), (eh_side)sd, (eh_uplo)ul, *m, *n, &alpha, A.p, A.ld, B.p, B.ld, &beta, C.p, C.ld, s) !=
      EH_SUCCESS)
    return report(name);
  if (!copy_back(C, c, s)) std::fprintf(stderr, "eigen_amd BLAS: copying C back failed\n");
  finish(s);
}

// ?trmm_ / ?trsm_ with the checks of blas/level3_impl.h:339-357 and
// :478-496: info 1 side, 2 uplo, 3 transa, 4 diag, 5 m, 6 n, 9 lda, 11 ldb.
template <typename T, bool SOLVE>
void tri_impl(const char* name, const char* side, const char* uplo, const char* opa, const char* diag, const int* m,
              const int* n, const T* palpha, const T* a, const int* lda, T* b, const int* ldb,
              TriFn fn = SOLVE ? eh_trsm : eh_trmm) {
  int info = 0;
  const int sd = side_code(*side), ul = uplo_code(*uplo), o = op_code(*opa), dg = diag_code(*diag);
  const int ka = sd == EH_LEFT ? *m : *n;
  if (sd < 0)
    info = 1;
  else if (ul < 0)
    info = 2;
  else if (o == OP_BAD)
    info = 3;
  else if (dg < 0)
    info = 4;
  else if (*m < 0)
    info = 5;
  else if (*n < 0)
    info = 6;
  else if (*lda < std::max(1, ka))
    info = 9;
  else if (*ldb < std::max(1, *m))
    info = 11;
  if (info) return xerbla_(name, &info, 6);
  if (*m == 0 || *n == 0) return;
  const T alpha = *palpha;
  hipStream_t s = g_stream;
  std::lock_guard<std::mutex> lk(g_stage_mutex);
  DevMat<T> A(a, *lda), B(b, *ldb);
  if ((alpha != T(0) && !dev_view(g_sa, a, ka, ka, *lda, true, s, &A)) ||
      !dev_view(g_sb, b, *m, *n, *ldb, alpha != T(0), s, &B))
    return (void)std::fprintf(stderr, "eigen_amd BLAS: staging failed\n");
  if (fn(dtype_of<T>(), (eh_side)sd, (eh_uplo)ul, (eh_op)o, (eh_diag)dg, *m, *n, &alpha, A.p, A.ld, B.p, B.ld,
         s) != EH_SUCCESS)
    return report(name);
  if (!copy_back(B, b, s)) std::fprintf(stderr, "eigen_amd BLAS: copying B back failed\n");
  finish(s);
}

// ?gemv_ with the checks of blas/level2_impl.h:45-61: info 1 trans, 2 m,
// 3 n, 6 lda, 8 incx, 11 incy; quick return for m == 0, n == 0, or
// alpha == 0 with beta == 1.  Host vectors are gathered into contiguous
// staging (a negative increment runs from the last element) and scattered
// back.
template <typename T>
void gemv_impl(const char* name, const char* op, const int* m, const int* n, const T* palpha, const T* a,
               const int* lda, const T* x, const int* incx, const T* pbeta, T* y, const int* incy) {
  int info = 0;
  const int o = op_code(*op);
  if (o == OP_BAD)
    info = 1;
  else if (*m < 0)
    info = 2;
  else if (*n < 0)
    info = 3;
  else if (*lda < std::max(1, *m))
    info = 6;
  else if (*incx == 0)
    info = 8;
  else if (*incy == 0)
    info = 11;
  if (info) return xerbla_(name, &info, 6);
  const T alpha = *palpha, beta = *pbeta;
  if (*m == 0 || *n == 0 || (alpha == T(0) && beta == T(1))) return;
  const int lx = o == OP_N ? *n : *m, ly = o == OP_N ? *m : *n;
  hipStream_t s = g_stream;
  std::lock_guard<std::mutex> lk(g_stage_mutex);
  DevMat<T> A(a, *lda);
  if (alpha != T(0) && !dev_view(g_sa, a, *m, *n, *lda, true, s, &A))
    return (void)std::fprintf(stderr, "eigen_amd BLAS: staging failed\n");
  // element i of a strided vector of length len
  auto at = [](int64_t i, int64_t len, int inc) { return (inc > 0 ? i : i - (len - 1)) * inc; };
  const T* dx = x;
  int64_t dincx = *incx;
  std::vector<T> hx, hy;
  if (alpha != T(0) && !is_device_ptr(x)) {
    hx.resize(lx);
    for (int64_t i = 0; i < lx; ++i) hx[i] = x[at(i, lx, *incx)];
    if (!g_sb.ensure((size_t)lx * sizeof(T)) ||
        hipMemcpyAsync(g_sb.ptr, hx.data(), (size_t)lx * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess)
      return (void)std::fprintf(stderr, "eigen_amd BLAS: staging x failed\n");
    dx = static_cast<const T*>(g_sb.ptr);
    dincx = 1;
  }
  T* dy = y;
  int64_t dincy = *incy;
  const bool host_y = !is_device_ptr(y);
  if (host_y) {
    hy.resize(ly);
    if (beta != T(0))
      for (int64_t i = 0; i < ly; ++i) hy[i] = y[at(i, ly, *incy)];
    if (!g_sc.ensure((size_t)ly * sizeof(T)) ||
        (beta != T(0) &&
         hipMemcpyAsync(g_sc.ptr, hy.data(), (size_t)ly * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess))
      return (void)std::fprintf(stderr, "eigen_amd BLAS: staging y failed\n");
    dy = static_cast<T*>(g_sc.ptr);
    dincy = 1;
  }
  if (eh_gemv(dtype_of<T>(), (eh_op)o, *m, *n, &alpha, A.p, A.ld, dx, dincx, &beta, dy, dincy, s) != EH_SUCCESS)
    return report(name);
  if (host_y) {
    if (hipMemcpyAsync(hy.data(), dy, (size_t)ly * sizeof(T), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return (void)std::fprintf(stderr, "eigen_amd BLAS: copying y back failed\n");
    for (int64_t i = 0; i < ly; ++i) y[at(i, ly, *incy)] = hy[i];
  }
  finish(s);
}

// cgemv_ / zgemv_: ?gemv_'s checks on complex elements (Z is a pair of reals).
template <typename T>
void complex_gemv_impl(const char* name, const char* op, const int* m, const int* n, const T* palpha,
                       const std::complex<T>* a, const int* lda, const std::complex<T>* x, const int* incx,
                       const T* pbeta, std::complex<T>* y, const int* incy) {
  using Z = std::complex<T>;
  int info = 0;
  const int o = op_code(*op);
  if (o == OP_BAD)
    info = 1;
  else if (*m < 0)
    info = 2;
  else if (*n < 0)
    info = 3;
  else if (*lda < std::max(1, *m))
    info = 6;
  else if (*incx == 0)
    info = 8;
  else if (*incy == 0)
    info = 11;
  if (info) return xerbla_(name, &info, 6);
  const bool alpha_zero = palpha[0] == T(0) && palpha[1] == T(0);
  const bool beta_one = pbeta[0] == T(1) && pbeta[1] == T(0), beta_zero = pbeta[0] == T(0) && pbeta[1] == T(0);
  if (*m == 0 || *n == 0 || (alpha_zero && beta_one)) return;
  const int lx = o == OP_N ? *n : *m, ly = o == OP_N ? *m : *n;
  hipStream_t s = g_stream;
  std::lock_guard<std::mutex> lk(g_stage_mutex);
  DevMat<Z> A(a, *lda);
  if (!alpha_zero && !dev_view(g_sa, a, *m, *n, *lda, true, s, &A))
    return (void)std::fprintf(stderr, "eigen_amd BLAS: staging failed\n");
  auto at = [](int64_t i, int64_t len, int inc) { return (inc > 0 ? i : i - (len - 1)) * inc; };
  const Z* dx = x;
  int64_t dincx = *incx;
  std::vector<Z> hx, hy;
  if (!alpha_zero && !is_device_ptr(x)) {
    hx.resize(lx);
    for (int64_t i = 0; i < lx; ++i) hx[i] = x[at(i, lx, *incx)];
    if (!g_sb.ensure((size_t)lx * sizeof(Z)) ||
        hipMemcpyAsync(g_sb.ptr, hx.data(), (size_t)lx * sizeof(Z), hipMemcpyHostToDevice, s) != hipSuccess)
      return (void)std::fprintf(stderr, "eigen_amd BLAS: staging x failed\n");
    dx = static_cast<const Z*>(g_sb.ptr);
    dincx = 1;
  }
  Z* dy = y;
  int64_t dincy = *incy;
  const bool host_y = !is_device_ptr(y);
  if (host_y) {
    hy.resize(ly);
    if (!beta_zero)
      for (int64_t i = 0; i < ly; ++i) hy[i] = y[at(i, ly, *incy)];
    if (!g_sc.ensure((size_t)ly * sizeof(Z)) ||
        (!beta_zero &&
         hipMemcpyAsync(g_sc.ptr, hy.data(), (size_t)ly * sizeof(Z), hipMemcpyHostToDevice, s) != hipSuccess))
      return (void)std::fprintf(stderr, "eigen_amd BLAS: staging y failed\n");
    dy = static_cast<Z*>(g_sc.ptr);
    dincy = 1;
  }
  if (eh_gemv_complex(dtype_of<T>(), (eh_op)o, *m, *n, palpha, A.p, A.ld, dx, dincx, pbeta, dy, dincy, s) !=
      EH_SUCCESS)
    return report(name);
  if (host_y) {
    if (hipMemcpyAsync(hy.data(), dy, (size_t)ly * sizeof(Z), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return (void)std::fprintf(stderr, "eigen_amd BLAS: copying y back failed\n");
    for (int64_t i = 0; i < ly; ++i) y[at(i, ly, *incy)] = hy[i];
  }
  finish(s);
}

// cgemm_ / zgemm_: ?gemm_'s checks, on complex elements (Z is a pair of reals).
template <typename T>
void complex_gemm_impl(const char* name, const char* opa, const char* opb, const int* m, const int* n, const int* k,
                       const T* palpha, const std::complex<T>* a, const int* lda, const std::complex<T>* b,
                       const int* ldb, const T* pbeta, std::complex<T>* c, const int* ldc) {
  using Z = std::complex<T>;
  int info = 0;
  const int oa = op_code(*opa), ob = op_code(*opb);
  if (oa == OP_BAD)
    info = 1;
  else if (ob == OP_BAD)
    info = 2;
  else if (*m < 0)
    info = 3;
  else if (*n < 0)
    info = 4;
  else if (*k < 0)
    info = 5;
  else if (*lda < std::max(1, oa == OP_N ? *m : *k))
    info = 8;
  else if (*ldb < std::max(1, ob == OP_N ? *k : *n))
    info = 10;
  else if (*ldc < std::max(1, *m))
    info = 13;
  if (info) return xerbla_(name, &info, 6);
  if (*m == 0 || *n == 0) return;
  const bool need_ab = *k > 0 && (palpha[0] != T(0) || palpha[1] != T(0));
  const bool beta_one = pbeta[0] == T(1) && pbeta[1] == T(0);
  if (!need_ab && beta_one) return;
  hipStream_t s = g_stream;
  std::lock_guard<std::mutex> lk(g_stage_mutex);
  DevMat<Z> A(a, *lda), B(b, *ldb), C(c, *ldc);
  const bool read_c = pbeta[0] != T(0) || pbeta[1] != T(0);
  if ((need_ab && (!dev_view(g_sa, a, oa == OP_N ? *m : *k, oa == OP_N ? *k : *m, *lda, true, s, &A) ||
                   !dev_view(g_sb, b, ob == OP_N ? *k : *n, ob == OP_N ? *n : *k, *ldb, true, s, &B))) ||
      !dev_view(g_sc, c, *m, *n, *ldc, read_c, s, &C))
    return (void)std::fprintf(stderr, "eigen_amd BLAS: staging failed\n");
  if (eh_gemm_complex(dtype_of<T>(), (eh_op)oa, (eh_op)ob, *m, *n, *k, palpha, A.p, A.ld, B.p, B.ld, pbeta, C.p,
                      C.ld, EH_PREC_AUTO, s) != EH_SUCCESS)
    return report(name);
  if (!copy_back(C, c, s)) std::fprintf(stderr, "eigen_amd BLAS: copying C back failed\n");
  finish(s);
}

// sbgemm_: bfloat16 A and B, float alpha, beta and C, with ?gemm_'s
// validation order and info codes (the reference declares this OpenBLAS
// entry point, GeneralMatrixMatrix_BLAS.h:127-131, and has no
// implementation of its own).
void sbgemm_impl(const char* opa, const char* opb, const int* m, const int* n, const int* k, const float* palpha,
                 const unsigned short* a, const int* lda, const unsigned short* b, const int* ldb, const float* pbeta,
                 float* c, const int* ldc) {
  static const char name[] = "SBGEMM";
  int info = 0;
  const int oa = op_code(*opa), ob = op_code(*opb);
  if (oa == OP_BAD)
    info = 1;
  else if (ob == OP_BAD)
    info = 2;
  else if (*m < 0)
    info = 3;
  else if (*n < 0)
    info = 4;
  else if (*k < 0)
    info = 5;
  else if (*lda < std::max(1, oa == OP_N ? *m : *k))
    info = 8;
  else if (*ldb < std::max(1, ob == OP_N ? *k : *n))
    info = 10;
  else if (*ldc < std::max(1, *m))
    info = 13;
  if (info) return xerbla_(name, &info, 6);
  if (*m == 0 || *n == 0) return;
  const float alpha = *palpha, beta = *pbeta;
  const bool need_ab = *k > 0 && alpha != 0.0f;
  if (!need_ab && beta == 1.0f) return;
  hipStream_t s = g_stream;
  std::lock_guard<std::mutex> lk(g_stage_mutex);
  DevMat<unsigned short> A(a, *lda), B(b, *ldb);
  DevMat<float> C(c, *ldc);
  if ((need_ab && (!dev_view(g_sa, a, oa == OP_N ? *m : *k, oa == OP_N ? *k : *m, *lda, true, s, &A) ||
                   !dev_view(g_sb, b, ob == OP_N ? *k : *n, ob == OP_N ? *n : *k, *ldb, true, s, &B))) ||
      !dev_view(g_sc, c, *m, *n, *ldc, beta != 0.0f, s, &C))
    return (void)std::fprintf(stderr, "eigen_amd BLAS: staging failed\n");
  if (eh_gemm_bf16((eh_op)oa, (eh_op)ob, *m, *n, *k, &alpha, A.p, A.ld, B.p, B.ld, &beta, C.p, C.ld, s) != EH_SUCCESS)
    return report(name);
  if (!copy_back(C, c, s)) std::fprintf(stderr, "eigen_amd BLAS: copying C back failed\n");
  finish(s);
}

}  // namespace

extern "C" {

__attribute__((weak)) void xerbla_(const char* name, int* info, size_t len) {
  std::printf("eigen_amd BLAS ERROR #%i: %.*s\n", *info, (int)len, name);
}

void sgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const float* alpha,
            const float* a, const int* lda, const float* b, const int* ldb, const float* beta, float* c,
            const int* ldc) {
  gemm_impl<float>("SGEMM ", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void dgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const double* alpha,
            const double* a, const int* lda, const double* b, const int* ldb, const double* beta, double* c,
            const int* ldc) {
  gemm_impl<double>("DGEMM ", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void ssyrk_(const char* uplo, const char* trans, const int* n, const int* k, const float* alpha, const float* a,
            const int* lda, const float* beta, float* c, const int* ldc) {
  syrk_impl<float>("SSYRK ", uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
}

void dsyrk_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha, const double* a,
            const int* lda, const double* beta, double* c, const int* ldc) {
  syrk_impl<double>("DSYRK ", uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
}

void sgemmtr_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
              const float* alpha, const float* a, const int* lda, const float* b, const int* ldb, const float* beta,
              float* c, const int* ldc) {
  gemmtr_impl<float>("SGEMMTR", uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void dgemmtr_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
              const double* alpha, const double* a, const int* lda, const double* b, const int* ldb,
              const double* beta, double* c, const int* ldc) {
  gemmtr_impl<double>("DGEMMTR", uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void ssyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const float* alpha, const float* a,
             const int* lda, const float* b, const int* ldb, const float* beta, float* c, const int* ldc) {
  syr2k_impl<float>("SSYR2K", uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void dsyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha, const double* a,
             const int* lda, const double* b, const int* ldb, const double* beta, double* c, const int* ldc) {
  syr2k_impl<double>("DSYR2K", uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void ssymm_(const char* side, const char* uplo, const int* m, const int* n, const float* alpha, const float* a,
            const int* lda, const float* b, const int* ldb, const float* beta, float* c, const int* ldc) {
  symm_impl<float>("SSYMM ", side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);
}

void dsymm_(const char* side, const char* uplo, const int* m, const int* n, const double* alpha, const double* a,
            const int* lda, const double* b, const int* ldb, const double* beta, double* c, const int* ldc) {
  symm_impl<double>("DSYMM ", side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);
}

void strmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const float* alpha, const float* a, const int* lda, float* b, const int* ldb) {
  tri_impl<float, false>("STRMM ", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}

void dtrmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const double* alpha, const double* a, const int* lda, double* b, const int* ldb) {
  tri_impl<double, false>("DTRMM ", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}

void strsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const float* alpha, const float* a, const int* lda, float* b, const int* ldb) {
  tri_impl<float, true>("STRSM ", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}

void dtrsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const double* alpha, const double* a, const int* lda, double* b, const int* ldb) {
  tri_impl<double, true>("DTRSM ", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}

void sgemv_(const char* trans, const int* m, const int* n, const float* alpha, const float* a, const int* lda,
            const float* x, const int* incx, const float* beta, float* y, const int* incy) {
  gemv_impl<float>("SGEMV ", trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
}

void dgemv_(const char* trans, const int* m, const int* n, const double* alpha, const double* a, const int* lda,
            const double* x, const int* incx, const double* beta, double* y, const int* incy) {
  gemv_impl<double>("DGEMV ", trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
}

void cgemv_(const char* trans, const int* m, const int* n, const void* alpha, const void* a, const int* lda,
            const void* x, const int* incx, const void* beta, void* y, const int* incy) {
  using Z = std::complex<float>;
  complex_gemv_impl<float>("CGEMV ", trans, m, n, static_cast<const float*>(alpha), static_cast<const Z*>(a), lda,
                           static_cast<const Z*>(x), incx, static_cast<const float*>(beta), static_cast<Z*>(y), incy);
}

void zgemv_(const char* trans, const int* m, const int* n, const void* alpha, const void* a, const int* lda,
            const void* x, const int* incx, const void* beta, void* y, const int* incy) {
  using Z = std::complex<double>;
  complex_gemv_impl<double>("ZGEMV ", trans, m, n, static_cast<const double*>(alpha), static_cast<const Z*>(a), lda,
                            static_cast<const Z*>(x), incx, static_cast<const double*>(beta), static_cast<Z*>(y),
                            incy);
}

void cgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const void* alpha,
            const void* a, const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc) {
  using Z = std::complex<float>;
  complex_gemm_impl<float>("CGEMM ", transa, transb, m, n, k, static_cast<const float*>(alpha),
                           static_cast<const Z*>(a), lda, static_cast<const Z*>(b), ldb,
                           static_cast<const float*>(beta), static_cast<Z*>(c), ldc);
}

void zgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const void* alpha,
            const void* a, const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc) {
  using Z = std::complex<double>;
  complex_gemm_impl<double>("ZGEMM ", transa, transb, m, n, k, static_cast<const double*>(alpha),
                            static_cast<const Z*>(a), lda, static_cast<const Z*>(b), ldb,
                            static_cast<const double*>(beta), static_cast<Z*>(c), ldc);
}

void sbgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const float* alpha,
             const void* a, const int* lda, const void* b, const int* ldb, const float* beta, float* c,
             const int* ldc) {
  sbgemm_impl(transa, transb, m, n, k, alpha, static_cast<const unsigned short*>(a), lda,
              static_cast<const unsigned short*>(b), ldb, beta, c, ldc);
}

// ---- complex level-3 (blas/level3_impl.h for complex scalars) --------------
#define EH_CX(R) std::complex<R>
#define EH_CP(R, p) static_cast<const EH_CX(R)*>(p)
#define EH_MP(R, p) static_cast<EH_CX(R)*>(p)

void ctrsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const void* alpha, const void* a, const int* lda, void* b, const int* ldb) {
  tri_impl<EH_CX(float), true>("CTRSM ", side, uplo, transa, diag, m, n, EH_CP(float, alpha), EH_CP(float, a), lda,
                               EH_MP(float, b), ldb, eh_trsm_complex);
}
void ztrsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const void* alpha, const void* a, const int* lda, void* b, const int* ldb) {
  tri_impl<EH_CX(double), true>("ZTRSM ", side, uplo, transa, diag, m, n, EH_CP(double, alpha), EH_CP(double, a), lda,
                                EH_MP(double, b), ldb, eh_trsm_complex);
}
void ctrmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const void* alpha, const void* a, const int* lda, void* b, const int* ldb) {
  tri_impl<EH_CX(float), false>("CTRMM ", side, uplo, transa, diag, m, n, EH_CP(float, alpha), EH_CP(float, a), lda,
                                EH_MP(float, b), ldb, eh_trmm_complex);
}
void ztrmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const void* alpha, const void* a, const int* lda, void* b, const int* ldb) {
  tri_impl<EH_CX(double), false>("ZTRMM ", side, uplo, transa, diag, m, n, EH_CP(double, alpha), EH_CP(double, a),
                                 lda, EH_MP(double, b), ldb, eh_trmm_complex);
}
void csymm_(const char* side, const char* uplo, const int* m, const int* n, const void* alpha, const void* a,
            const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc) {
  symm_impl<EH_CX(float)>("CSYMM ", side, uplo, m, n, EH_CP(float, alpha), EH_CP(float, a), lda, EH_CP(float, b), ldb,
                          EH_CP(float, beta), EH_MP(float, c), ldc, eh_symm_complex);
}
void zsymm_(const char* side, const char* uplo, const int* m, const int* n, const void* alpha, const void* a,
            const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc) {
  symm_impl<EH_CX(double)>("ZSYMM ", side, uplo, m, n, EH_CP(double, alpha), EH_CP(double, a), lda,
                           EH_CP(double, b), ldb, EH_CP(double, beta), EH_MP(double, c), ldc, eh_symm_complex);
}
void chemm_(const char* side, const char* uplo, const int* m, const int* n, const void* alpha, const void* a,
            const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc) {
  symm_impl<EH_CX(float)>("CHEMM ", side, uplo, m, n, EH_CP(float, alpha), EH_CP(float, a), lda, EH_CP(float, b), ldb,
                          EH_CP(float, beta), EH_MP(float, c), ldc, eh_hemm);
}
void zhemm_(const char* side, const char* uplo, const int* m, const int* n, const void* alpha, const void* a,
            const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc) {
  symm_impl<EH_CX(double)>("ZHEMM ", side, uplo, m, n, EH_CP(double, alpha), EH_CP(double, a), lda,
                           EH_CP(double, b), ldb, EH_CP(double, beta), EH_MP(double, c), ldc, eh_hemm);
}
void csyrk_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
            const int* lda, const void* beta, void* c, const int* ldc) {
  syrk_impl<EH_CX(float)>("CSYRK ", uplo, trans, n, k, EH_CP(float, alpha), EH_CP(float, a), lda, EH_CP(float, beta),
                          EH_MP(float, c), ldc, eh_syrk_complex, OP_C);
}
void zsyrk_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
            const int* lda, const void* beta, void* c, const int* ldc) {
  syrk_impl<EH_CX(double)>("ZSYRK ", uplo, trans, n, k, EH_CP(double, alpha), EH_CP(double, a), lda,
                           EH_CP(double, beta), EH_MP(double, c), ldc, eh_syrk_complex, OP_C);
}
void cherk_(const char* uplo, const char* trans, const int* n, const int* k, const float* alpha, const void* a,
            const int* lda, const float* beta, void* c, const int* ldc) {
  syrk_impl<EH_CX(float), float>("CHERK ", uplo, trans, n, k, alpha, EH_CP(float, a), lda, beta, EH_MP(float, c), ldc,
                                 eh_herk, OP_T);
}
void zherk_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha, const void* a,
            const int* lda, const double* beta, void* c, const int* ldc) {
  syrk_impl<EH_CX(double), double>("ZHERK ", uplo, trans, n, k, alpha, EH_CP(double, a), lda, beta, EH_MP(double, c),
                                   ldc, eh_herk, OP_T);
}
void csyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
             const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc) {
  syr2k_impl<EH_CX(float)>("CSYR2K", uplo, trans, n, k, EH_CP(float, alpha), EH_CP(float, a), lda, EH_CP(float, b),
                           ldb, EH_CP(float, beta), EH_MP(float, c), ldc, eh_syr2k_complex, OP_C);
}
void zsyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
             const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc) {
  syr2k_impl<EH_CX(double)>("ZSYR2K", uplo, trans, n, k, EH_CP(double, alpha), EH_CP(double, a), lda,
                            EH_CP(double, b), ldb, EH_CP(double, beta), EH_MP(double, c), ldc, eh_syr2k_complex, OP_C);
}
void cher2k_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
             const int* lda, const void* b, const int* ldb, const float* beta, void* c, const int* ldc) {
  syr2k_impl<EH_CX(float), float>("CHER2K", uplo, trans, n, k, EH_CP(float, alpha), EH_CP(float, a), lda,
                                  EH_CP(float, b), ldb, beta, EH_MP(float, c), ldc, eh_her2k, OP_T);
}
void zher2k_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
             const int* lda, const void* b, const int* ldb, const double* beta, void* c, const int* ldc) {
  syr2k_impl<EH_CX(double), double>("ZHER2K", uplo, trans, n, k, EH_CP(double, alpha), EH_CP(double, a), lda,
                                    EH_CP(double, b), ldb, beta, EH_MP(double, c), ldc, eh_her2k, OP_T);
}
void cgemmtr_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
              const void* alpha, const void* a, const int* lda, const void* b, const int* ldb, const void* beta,
              void* c, const int* ldc) {
  gemmtr_impl<EH_CX(float)>("CGEMMTR", uplo, transa, transb, n, k, EH_CP(float, alpha), EH_CP(float, a), lda,
                            EH_CP(float, b), ldb, EH_CP(float, beta), EH_MP(float, c), ldc, eh_gemmtr_complex);
}
void zgemmtr_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
              const void* alpha, const void* a, const int* lda, const void* b, const int* ldb, const void* beta,
              void* c, const int* ldc) {
  gemmtr_impl<EH_CX(double)>("ZGEMMTR", uplo, transa, transb, n, k, EH_CP(double, alpha), EH_CP(double, a), lda,
                             EH_CP(double, b), ldb, EH_CP(double, beta), EH_MP(double, c), ldc, eh_gemmtr_complex);
}
#undef EH_CX
#undef EH_CP
#undef EH_MP

void eh_blas_set_stream(void* stream) { g_stream = static_cast<hipStream_t>(stream); }
void* eh_blas_get_stream(void) { return g_stream; }

}  // extern "C"
