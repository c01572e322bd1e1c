// capi.cpp -- the C-ABI: argument checks, planning, workspace, dispatch.
//
// Host-side counterpart of Eigen's GEMM driver
// (Eigen/src/Core/products/GeneralMatrixMatrix.h:117-246 and :422-515): it
// takes the product request, applies the BLAS quick-return rules
// (blas/level3_impl.h:80-95, GeneralMatrixMatrix.h:130), chooses the tile /
// split-K configuration (the role of computeProductBlockingSizes,
// GeneralBlockPanelKernel.h:201-452) and launches the device kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/eigen_amd.h"
#include "internal.hpp"
#include "kernels.hpp"

namespace eh {

thread_local std::string g_last_error;

eh_status fail(eh_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

eh_status hip_fail(hipError_t e, const char* what) {
  return fail(EH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---- configuration table --------------------------------------------------
// config = tile * 4 + split index; tiles as in launch_gemm_mfma, split-K
// factors {1, 2, 4, 8}.
static const int kTileM[kNumTiles] = {64, 64, 128, 128, 128, 256};
static const int kTileN[kNumTiles] = {64, 128, 64, 128, 256, 128};
static const int kSplits[kNumSplits] = {1, 2, 4, 8};

void config_decode(int config, int* tile, int* splits) {
  *tile = config / kNumSplits;
  *splits = kSplits[config % kNumSplits];
}

int config_tile_m(int config) { return kTileM[config / kNumSplits]; }
int config_tile_n(int config) { return kTileN[config / kNumSplits]; }

static int split_index(int s) {
  for (int i = 0; i < kNumSplits; ++i)
    if (kSplits[i] == s) return i;
  return 0;
}

// The closed-form rule (the method without the learned selector): the
// largest preferred tile that still gives at least one wave of workgroups
// over the 256 CUs, then split K while the grid stays under one wave and
// each split keeps at least 512 of depth.
int heuristic_config(int dtype, int64_t m, int64_t n, int64_t k) {
  (void)dtype;
  // 128x128 first: on gfx950 the 256-wide tiles run at occupancy 1 and
  // measured slower than 128x128 for FP32 and FP64 alike.
  static const int order[] = {3, 2, 1, 0};
  int tile = 0;
  for (int t : order) {
    int64_t tiles = ((m + kTileM[t] - 1) / kTileM[t]) * ((n + kTileN[t] - 1) / kTileN[t]);
    if (tiles >= 256) {
      tile = t;
      break;
    }
  }
  int64_t tiles = ((m + kTileM[tile] - 1) / kTileM[tile]) * ((n + kTileN[tile] - 1) / kTileN[tile]);
  int s = 1;
  while (s < 8 && tiles * s < 256 && k / (2 * s) >= 512) s *= 2;
  return tile * kNumSplits + split_index(s);
}

// ---- packed products: configurations ----------------------------------------
// The split-BF16 FP32 products and the bf16 / half input products run the
// packed kernel (launch_gemm_split_packed), whose configurations are
// tile * kNumSplits + split index like the plain ones, over its own tiles:
// split terms 0 = 128x128, 1 = 256x128 (4 waves), 2 = 256x128 (8 waves),
// 3 = 256x256 (8 waves, 16-deep stages); one term 0 = 128x128, 1 = 256x256.
static bool packed_one_term(int kind) { return kind == EH_PACKED_BF16 || kind == EH_PACKED_F16; }

int packed_num_configs(int kind) { return (packed_one_term(kind) ? 2 : 4) * kNumSplits; }

bool packed_kind_valid(int kind) {
  return kind == EH_PACKED_SPLIT2 || kind == EH_PACKED_SPLIT3 || kind == EH_PACKED_BF16 || kind == EH_PACKED_F16;
}

// The fixed rule: the largest tile that gives one wave of workgroups over
// the 256 CUs (256x256, then for the split terms 256x128 with 8 waves),
// 128x128 below; then split K while the grid stays under one wave and each
// split keeps at least 512 of depth.
int packed_rule_config(int kind, int64_t m, int64_t n, int64_t k) {
  const bool one = packed_one_term(kind);
  const int64_t t256 = ((m + 255) / 256) * ((n + 255) / 256);
  const int64_t t256x128 = ((m + 255) / 256) * ((n + 127) / 128);
  int tile = 0;
  int64_t tiles = ((m + 127) / 128) * ((n + 127) / 128);
  if (t256 >= 256) {
    tile = one ? 1 : 3;
    tiles = t256;
  } else if (!one && t256x128 >= 256) {
    tile = 2;
    tiles = t256x128;
  }
  int s = 1;
  while (s < 8 && tiles * s < 256 && k / (2 * s) >= 512) s *= 2;
  return tile * kNumSplits + split_index(s);
}

int table_plan_packed(int kind, int opa, int opb, int64_t m, int64_t n, int64_t k);

// Configuration of a packed product under the current tuning mode: the
// timing table (tuned on first use) in EH_TUNE_TABLE, the rule otherwise
// (the learned selector is trained on the plain kernels only).
int plan_packed(int kind, int opa, int opb, int64_t m, int64_t n, int64_t k) {
  if (tune_mode() == EH_TUNE_TABLE) {
    const int c = table_plan_packed(kind, opa, opb, m, n, k);
    if (c >= 0 && c < packed_num_configs(kind)) return c;
  }
  return packed_rule_config(kind, m, n, k);
}

bool mlp_has_default();

static std::mutex g_mutex;
static int g_mode = -1;  // -1: not yet chosen (EIGEN_AMD_TUNE, else MLP when compiled in)
static int g_sgemm_prec = -1;  // -1: not yet read from the environment

static eh_precision default_sgemm_precision() {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (g_sgemm_prec < 0) {
    g_sgemm_prec = EH_PREC_PLAIN;
    if (const char* e = std::getenv("EIGEN_AMD_SGEMM_PRECISION")) {
      if (!std::strcmp(e, "split3")) g_sgemm_prec = EH_PREC_SPLIT3;
      if (!std::strcmp(e, "split2")) g_sgemm_prec = EH_PREC_SPLIT2;
    }
  }
  return (eh_precision)g_sgemm_prec;
}

// ---- workspace --------------------------------------------------------------
// One growable device buffer per (device, stream) for split-K partials, packed
// operands and work lists; hipStreamPerThread names a different stream in
// every thread, so it is keyed per thread as well.  A call holds the stream's
// lock (recursive: a call may re-enter through a sub-product) from taking the
// buffer until its launches are queued, so calls from several threads on one
// stream queue their launch sequences one after another.  Growing replaces
// the buffer; the old one is freed when its last reference goes, with
// hipFree, which waits for the device, so work queued with it completes
// before the memory is released.
struct WorkspaceSlot {
  std::recursive_mutex mu;
  std::shared_ptr<Workspace> buf;
};
static std::map<std::tuple<int, hipStream_t, std::thread::id>, std::unique_ptr<WorkspaceSlot>> g_ws;

Workspace::~Workspace() {
  if (ptr) (void)hipFree(ptr);
}

static thread_local void* t_fixed_ptr = nullptr;
static thread_local size_t t_fixed_bytes = 0;

FixedScratch::FixedScratch(void* ptr, size_t bytes) : prev_ptr(t_fixed_ptr), prev_bytes(t_fixed_bytes) {
  t_fixed_ptr = ptr;
  t_fixed_bytes = bytes;
}
FixedScratch::~FixedScratch() {
  t_fixed_ptr = prev_ptr;
  t_fixed_bytes = prev_bytes;
}
bool fixed_scratch_active() { return t_fixed_ptr != nullptr; }

eh_status workspace_get(hipStream_t stream, size_t bytes, WorkspaceLease* out, void** ptr) {
  if (t_fixed_ptr) {
    if (bytes > t_fixed_bytes)
      return fail(EH_ERR_UNSUPPORTED, "product needs " + std::to_string(bytes) +
                                          " bytes of scratch; the fixed scratch holds " + std::to_string(t_fixed_bytes));
    *ptr = t_fixed_ptr;
    return EH_SUCCESS;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  const std::thread::id tid = stream == hipStreamPerThread ? std::this_thread::get_id() : std::thread::id();
  WorkspaceSlot* slot;
  {
    std::lock_guard<std::mutex> lk(g_mutex);
    std::unique_ptr<WorkspaceSlot>& sp = g_ws[std::make_tuple(dev, stream, tid)];
    if (!sp) sp.reset(new WorkspaceSlot);
    slot = sp.get();  // slots are never removed
  }
  WorkspaceLease lease;
  lease.lock = std::unique_lock<std::recursive_mutex>(slot->mu);
  if (!slot->buf || slot->buf->bytes < bytes) {
    const size_t have = slot->buf ? slot->buf->bytes : 0;
    auto fresh = std::make_shared<Workspace>();
    const size_t want = std::max(bytes, std::max(have, (size_t)1 << 20));
    slot->buf.reset();  // the old buffer goes once no lease refers to it
    e = hipMalloc(&fresh->ptr, want);
    if (e != hipSuccess) {
      fresh->ptr = nullptr;
      return fail(EH_ERR_ALLOC, std::string("hipMalloc(workspace): ") + hipGetErrorString(e));
    }
    fresh->bytes = want;
    slot->buf = std::move(fresh);
  }
  lease.buf = slot->buf;
  *ptr = lease.buf->ptr;
  *out = std::move(lease);
  return EH_SUCCESS;
}

// ---- kernel event timing ------------------------------------------------------
static std::atomic<int> g_timing_on{0};
static std::mutex g_timing_mutex;
static std::vector<std::pair<hipEvent_t, hipEvent_t>> g_timed[EH_KC_COUNT];

KernelTimer::KernelTimer(int c, hipStream_t st) : cls(c), s(st) {
  if (!g_timing_on.load(std::memory_order_relaxed) || fixed_scratch_active()) return;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return;
  if (hipEventCreate(&start) != hipSuccess) {
    start = nullptr;
    return;
  }
  if (hipEventRecord(start, s) != hipSuccess) {
    (void)hipEventDestroy(start);
    start = nullptr;
  }
}

KernelTimer::~KernelTimer() {
  if (!start) return;
  hipEvent_t stop = nullptr;
  if (hipEventCreate(&stop) != hipSuccess || hipEventRecord(stop, s) != hipSuccess) {
    if (stop) (void)hipEventDestroy(stop);
    (void)hipEventDestroy(start);
    return;
  }
  std::lock_guard<std::mutex> lk(g_timing_mutex);
  g_timed[cls].emplace_back(start, stop);
}

static void timing_clear_locked() {
  for (auto& v : g_timed) {
    for (auto& ev : v) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    v.clear();
  }
}

// ---- argument checks ------------------------------------------------------
static bool valid_op(int op) { return op == EH_OP_N || op == EH_OP_T || op == EH_OP_C; }

eh_status check_gemm_args(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                          const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, const void* C,
                          int64_t ldc) {
  if (dtype != EH_F32 && dtype != EH_F64) return fail(EH_ERR_INVALID, "dtype must be EH_F32 or EH_F64");
  if (!valid_op(opa)) return fail(EH_ERR_INVALID, "opa must be N, T or C");
  if (!valid_op(opb)) return fail(EH_ERR_INVALID, "opb must be N, T or C");
  if (m < 0 || n < 0 || k < 0) return fail(EH_ERR_INVALID, "m, n, k must be >= 0");
  if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) return fail(EH_ERR_UNSUPPORTED, "m, n, k must fit int32");
  if (lda < std::max<int64_t>(1, opa == EH_OP_N ? m : k)) return fail(EH_ERR_INVALID, "lda too small");
  if (ldb < std::max<int64_t>(1, opb == EH_OP_N ? k : n)) return fail(EH_ERR_INVALID, "ldb too small");
  if (ldc < std::max<int64_t>(1, m)) return fail(EH_ERR_INVALID, "ldc too small");
  if (!alpha || !beta) return fail(EH_ERR_INVALID, "alpha and beta must point to host scalars");
  if (m > 0 && n > 0) {
    if (!C) return fail(EH_ERR_INVALID, "C is NULL");
    if (k > 0 && (!A || !B)) return fail(EH_ERR_INVALID, "A or B is NULL");
  }
  return EH_SUCCESS;
}

// ---- dispatch ---------------------------------------------------------------
// Byte extents of the operands as the kernels address them, with the margin
// the k tail and edge tiles may touch.  Kernels index operands with 32-bit
// offsets from the base pointer, so a product whose operands exceed that is
// cut into sub-products first.
constexpr uint64_t kMaxExtent = 0xFFFFFFFFull;

template <typename T>
static bool operand_fits(bool kcont, int64_t rows, int64_t k, int64_t ld) {
  const uint64_t sz = sizeof(T);
  if (kcont) return ((uint64_t)(rows - 1) * ld + k) * sz < kMaxExtent;      // rows of k values
  return ((uint64_t)(k + 64) * ld + rows) * sz < kMaxExtent;                // k columns of rows
}

// Split-BF16 FP32 product: pack both operands into bf16 planes (one pass
// over each), then run the packed product kernel (and the split-K reduction
// when K is cut).
static eh_status launch_split_product(const GemmParams<float>& g, bool ta, bool tb, int terms, int config,
                                      hipStream_t stream) {
  const int m = g.M, n = g.N, k = g.K;
  const int KS = split_kslices(k);
  const int kind = terms == 3 ? EH_PACKED_SPLIT3 : EH_PACKED_SPLIT2;
  if (config < 0) config = plan_packed(kind, ta, tb, m, n, k);
  if (config >= packed_num_configs(kind)) return fail(EH_ERR_INVALID, "packed config out of range");
  int tile = 0, splits = 1;
  config_decode(config, &tile, &splits);
  const int per = (KS + splits - 1) / splits;
  splits = (KS + per - 1) / per;
  const size_t bytesA = split_packed_bytes(m, k, terms), bytesB = split_packed_bytes(n, k, terms);
  const size_t bytesW = splits > 1 ? (size_t)splits * m * n * sizeof(float) : 0;
  void* ws = nullptr;
  WorkspaceLease lease;
  eh_status st = workspace_get(stream, bytesA + bytesB + bytesW, &lease, &ws);
  if (st != EH_SUCCESS) return st;
  unsigned char* base = static_cast<unsigned char*>(ws);
  hipError_t e;
  {
    KernelTimer t(EH_KC_PACK, stream);
    e = launch_pack_split(g.A, g.lda, ta, m, k, terms, base, stream);
  }
  if (e == hipSuccess) {
    KernelTimer t(EH_KC_PACK, stream);
    e = launch_pack_split(g.B, g.ldb, !tb, n, k, terms, base + bytesA, stream);
  }
  if (e != hipSuccess) return hip_fail(e, "pack launch");
  SplitParams sp{};
  sp.Ap = base;
  sp.Bp = base + bytesA;
  sp.C = g.C;
  sp.W = splits > 1 ? reinterpret_cast<float*>(base + bytesA + bytesB) : nullptr;
  sp.ldc = g.ldc;
  sp.M = m;
  sp.N = n;
  sp.KS = KS;
  sp.Mp = split_rows_padded(m);
  sp.Np = split_rows_padded(n);
  sp.ks_per_split = per;
  sp.alpha = g.alpha;
  sp.beta = g.beta;
  {
    KernelTimer t(EH_KC_PACKED_PRODUCT, stream);
    e = launch_gemm_split_packed(sp, terms, tile, splits, stream);
  }
  if (e != hipSuccess) return hip_fail(e, "split product launch");
  if (splits > 1) {
    GemmParams<float> r = g;
    r.W = sp.W;
    e = launch_splitk_reduce<float>(r, splits, stream);
    if (e != hipSuccess) return hip_fail(e, "split-K reduce launch");
  }
  return EH_SUCCESS;
}

// bf16 (or, with f16, IEEE half) inputs, FP32 accumulation and result:
// the operands are repacked into the one-term plane layout (or read in
// place, below), then the packed BF16 (FP16) MFMA product runs with a single
// product per k slice (and the split-K reduction when K is cut).  Operand
// extents must stay below 4 GB (32-bit buffer offsets).
static eh_status launch_bf16_product(bool ta, bool tb, int m, int n, int k, float alpha, const unsigned short* A,
                                     int64_t lda, const unsigned short* B, int64_t ldb, float beta, float* C,
                                     int64_t ldc, hipStream_t stream, bool f16 = false, int config = -1) {
  const uint64_t extA = ta ? (uint64_t)(m - 1) * lda + k : (uint64_t)(k - 1) * lda + m;
  const uint64_t extB = !tb ? (uint64_t)(n - 1) * ldb + k : (uint64_t)(k - 1) * ldb + n;
  if ((extA + 64) * 2 >= kMaxExtent || (extB + 64) * 2 >= kMaxExtent)
    return fail(EH_ERR_UNSUPPORTED, "16-bit operands must span less than 4 GB");
  const int KS = split_kslices(k);
  const int kind = f16 ? EH_PACKED_F16 : EH_PACKED_BF16;
  if (config < 0) config = plan_packed(kind, ta, tb, m, n, k);
  if (config >= packed_num_configs(kind)) return fail(EH_ERR_INVALID, "packed config out of range");
  int tile = 0, splits = 1;
  config_decode(config, &tile, &splits);
  const int per = (KS + splits - 1) / splits;
  splits = (KS + per - 1) / per;
  // An operand whose op(X) rows are k-contiguous (A^T, or B), with K a
  // multiple of 32 and 16-byte aligned rows, is read in place: its rows
  // already hold the packed layout's 64-byte slices.
  auto in_place = [k](bool kcont, const unsigned short* X, int64_t ld) {
    return kcont && k % 32 == 0 && ld % 8 == 0 && reinterpret_cast<uintptr_t>(X) % 16 == 0;
  };
  const bool dA = in_place(ta, A, lda), dB = in_place(!tb, B, ldb);
  const size_t bytesA = dA ? 0 : split_packed_bytes(m, k, 1), bytesB = dB ? 0 : split_packed_bytes(n, k, 1);
  const size_t bytesW = splits > 1 ? (size_t)splits * m * n * sizeof(float) : 0;
  void* ws = nullptr;
  WorkspaceLease lease;
  eh_status st = workspace_get(stream, bytesA + bytesB + bytesW, &lease, &ws);
  if (st != EH_SUCCESS) return st;
  unsigned char* base = static_cast<unsigned char*>(ws);
  hipError_t e = hipSuccess;
  if (!dA) {
    KernelTimer t(EH_KC_PACK, stream);
    e = launch_pack_bf16(A, lda, ta, m, k, base, stream);
  }
  if (e == hipSuccess && !dB) {
    KernelTimer t(EH_KC_PACK, stream);
    e = launch_pack_bf16(B, ldb, !tb, n, k, base + bytesA, stream);
  }
  if (e != hipSuccess) return hip_fail(e, "bf16 pack launch");
  SplitParams sp{};
  sp.Ap = dA ? static_cast<const void*>(A) : base;
  sp.Bp = dB ? static_cast<const void*>(B) : base + bytesA;
  sp.ldAd = dA ? lda : 0;
  sp.ldBd = dB ? ldb : 0;
  sp.C = C;
  sp.W = splits > 1 ? reinterpret_cast<float*>(base + bytesA + bytesB) : nullptr;
  sp.ldc = ldc;
  sp.M = m;
  sp.N = n;
  sp.KS = KS;
  sp.Mp = split_rows_padded(m);
  sp.Np = split_rows_padded(n);
  sp.ks_per_split = per;
  sp.alpha = alpha;
  sp.beta = beta;
  {
    KernelTimer t(EH_KC_PACKED_PRODUCT, stream);
    e = launch_gemm_split_packed(sp, 1, tile, splits, stream, f16);
  }
  if (e != hipSuccess) return hip_fail(e, "16-bit product launch");
  if (splits > 1) {
    GemmParams<float> r{};
    r.C = C;
    r.W = sp.W;
    r.ldc = ldc;
    r.M = m;
    r.N = n;
    r.alpha = alpha;
    r.beta = beta;
    e = launch_splitk_reduce<float>(r, splits, stream);
    if (e != hipSuccess) return hip_fail(e, "split-K reduce launch");
  }
  return EH_SUCCESS;
}

template <typename T>
static eh_status launch_product(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* A,
                                int64_t lda, const T* B, int64_t ldb, T beta, T* C, int64_t ldc, bool split_prec,
                                int terms, int config, hipStream_t stream, int tri = 0) {
  const bool fa = operand_fits<T>(ta, m, k, lda), fb = operand_fits<T>(!tb, n, k, ldb);
  if (!fa || !fb) {
    // Halve the dimension that drives the oversized extent and recurse.
    const bool cut_k = (!fa && !ta) || (!fb && tb);
    const bool cut_m = !fa && ta;
    if (cut_k && k > 1) {
      const int64_t h = k / 2;
      eh_status st = launch_product<T>(ta, tb, m, n, h, alpha, A, lda, B, ldb, beta, C, ldc, split_prec, terms,
                                       config, stream, tri);
      if (st != EH_SUCCESS) return st;
      const T* A2 = ta ? A + h : A + h * lda;
      const T* B2 = tb ? B + h * ldb : B + h;
      return launch_product<T>(ta, tb, m, n, k - h, alpha, A2, lda, B2, ldb, T(1), C, ldc, split_prec, terms,
                               config, stream, tri);
    }
    if (tri != 0) return fail(EH_ERR_UNSUPPORTED, "triangular update with operands past 32-bit addressing");
    if (cut_m && m > 1) {
      const int64_t h = m / 2;
      eh_status st = launch_product<T>(ta, tb, h, n, k, alpha, A, lda, B, ldb, beta, C, ldc, split_prec, terms,
                                       config, stream);
      if (st != EH_SUCCESS) return st;
      return launch_product<T>(ta, tb, m - h, n, k, alpha, ta ? A + h * lda : A + h, lda, B, ldb, beta, C + h, ldc,
                               split_prec, terms, config, stream);
    }
    if (!fb && !tb && n > 1) {
      const int64_t h = n / 2;
      eh_status st = launch_product<T>(ta, tb, m, h, k, alpha, A, lda, B, ldb, beta, C, ldc, split_prec, terms,
                                       config, stream);
      if (st != EH_SUCCESS) return st;
      return launch_product<T>(ta, tb, m, n - h, k, alpha, A, lda, B + h * ldb, ldb, beta, C + h * ldc, ldc,
                               split_prec, terms, config, stream);
    }
    return fail(EH_ERR_UNSUPPORTED, "operand leading dimension too large for 32-bit addressing");
  }

  GemmParams<T> p{};
  p.A = A;
  p.B = B;
  p.C = C;
  p.lda = lda;
  p.ldb = ldb;
  p.ldc = ldc;
  p.M = (int)m;
  p.N = (int)n;
  p.K = (int)k;
  p.alpha = alpha;
  p.beta = beta;
  p.tri = tri;
  p.bytesA = (uint32_t)((ta ? (uint64_t)(m - 1) * lda + k : (uint64_t)(k - 1) * lda + m) * sizeof(T));
  p.bytesB = (uint32_t)((!tb ? (uint64_t)(n - 1) * ldb + k : (uint64_t)(k - 1) * ldb + n) * sizeof(T));

  int tile = 0, splits = 1, kslice = (int)(128 / sizeof(T));
  hipError_t e;
  if constexpr (std::is_same<T, float>::value) {
    if (split_prec) return launch_split_product(p, ta, tb, terms, config, stream);
  }
  {
    if (config < 0 || config >= kNumConfigs) return fail(EH_ERR_INVALID, "config out of range");
    config_decode(config, &tile, &splits);
    // a triangular update of a square result runs square tiles, which launch
    // only the tiles that meet the triangle (balanced over the XCDs)
    if (tri != 0 && m == n && kTileM[tile] != kTileN[tile]) tile = m >= 2048 ? 3 : 0;
  }

  // depth per split: a whole number of k slices
  int64_t per = (k + splits - 1) / splits;
  per = (per + kslice - 1) / kslice * kslice;
  p.k_per_split = (int)per;
  splits = (int)((k + per - 1) / per);  // drop empty trailing splits
  WorkspaceLease lease;
  if (splits > 1) {
    void* ws = nullptr;
    eh_status st = workspace_get(stream, (size_t)splits * m * n * sizeof(T), &lease, &ws);
    if (st != EH_SUCCESS) return st;
    p.W = static_cast<T*>(ws);
  }
  {
    KernelTimer t(EH_KC_PRODUCT, stream);
    e = launch_gemm_mfma<T>(p, tile, ta, tb, splits, stream);
  }
  if (e != hipSuccess) return hip_fail(e, "gemm launch");
  if (splits > 1) {
    e = launch_splitk_reduce<T>(p, splits, stream);
    if (e != hipSuccess) return hip_fail(e, "split-K reduce launch");
  }
  return EH_SUCCESS;
}

// A product with one output column (n == 1) or row (m == 1) is a
// matrix-vector product: C^T = op(B)^T op(A)^T turns the row case into the
// column case.  Returns false (nothing launched) when the matrix spans 4 GB
// or more; the tiled path then handles it.
template <typename T>
static bool vector_product(int opa, int opb, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda,
                           const T* B, int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t stream, eh_status* st) {
  bool trans;
  int rows;
  const T *mat, *x;
  int64_t ldm, incx, incy;
  if (n == 1) {  // y = op(A) * op(B)(:, 0)
    trans = opa != EH_OP_N;
    rows = (int)m;
    mat = A;
    ldm = lda;
    x = B;
    incx = opb == EH_OP_N ? 1 : ldb;
    incy = 1;
  } else {  // m == 1: y^T = op(B)^T * op(A)(0, :)^T, y stored along C's row
    trans = opb == EH_OP_N;
    rows = (int)n;
    mat = B;
    ldm = ldb;
    x = A;
    incx = opa == EH_OP_N ? lda : 1;
    incy = ldc;
  }
  const uint64_t extent =
      ((trans ? (uint64_t)(rows - 1) * ldm + k : (uint64_t)(k - 1) * ldm + rows) + 64) * sizeof(T);
  if (extent >= kMaxExtent) return false;
  const GemvPlan g = gemv_plan(trans, rows, (int)k, (int)sizeof(T));
  void* ws = nullptr;
  WorkspaceLease lease;
  const size_t wsb = gemv_workspace_bytes(g, rows, (int)sizeof(T));
  if (wsb > 0) {
    *st = workspace_get(stream, wsb, &lease, &ws);
    if (*st != EH_SUCCESS) return true;
  }
  KernelTimer t(EH_KC_GEMV, stream);
  hipError_t e = launch_gemv<T>(g, trans, rows, (int)k, alpha, mat, ldm, (uint32_t)(extent - 64 * sizeof(T)), x, incx,
                                beta, C, incy, ws, stream);
  *st = e == hipSuccess ? EH_SUCCESS : hip_fail(e, "gemv launch");
  return true;
}

template <typename T>
static eh_status run_gemm(int opa, int opb, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda,
                          const T* B, int64_t ldb, T beta, T* C, int64_t ldc, eh_precision prec, int config,
                          hipStream_t stream) {
  if (m == 0 || n == 0) return EH_SUCCESS;
  hipError_t e;
  if (k == 0 || alpha == T(0)) {  // product term vanishes: A and B are not read
    if (beta == T(1)) return EH_SUCCESS;
    e = launch_scale_c<T>(C, (int)m, (int)n, ldc, beta, stream);
    return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "scale_c launch");
  }
  const bool ta = opa != EH_OP_N, tb = opb != EH_OP_N;
  bool split_prec = false;
  int terms = 0;
  if (std::is_same<T, float>::value) {
    if (prec == EH_PREC_AUTO) prec = default_sgemm_precision();
    if (prec == EH_PREC_SPLIT2 || prec == EH_PREC_SPLIT3) {
      split_prec = true;
      terms = prec == EH_PREC_SPLIT3 ? 3 : 2;
    }
  } else if (prec != EH_PREC_AUTO && prec != EH_PREC_PLAIN) {
    return fail(EH_ERR_UNSUPPORTED, "FP64 products run on FP64 MFMA only (precision AUTO or PLAIN)");
  }
  if (split_prec && config >= packed_num_configs(terms == 3 ? EH_PACKED_SPLIT3 : EH_PACKED_SPLIT2))
    return fail(EH_ERR_INVALID, "config out of range for the split-BF16 product");
  if (!split_prec && config < 0 && (n == 1 || m == 1)) {
    // HBM-bound matrix-vector product (a split precision keeps its own
    // arithmetic and runs the split product)
    eh_status st = EH_SUCCESS;
    if (vector_product<T>(opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream, &st)) return st;
  }
  if (!split_prec && config < 0) {
    eh_status st = plan_config(std::is_same<T, float>::value ? EH_F32 : EH_F64, opa, opb, m, n, k, &config);
    if (st != EH_SUCCESS) return st;
  }
  return launch_product<T>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, split_prec, terms, config, stream);
}

template <typename T>
eh_status device_product(int opa, int opb, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda,
                         const T* B, int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t stream, int tri) {
  if (tri == 0) return run_gemm<T>(opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, EH_PREC_PLAIN, -1, stream);
  if (m == 0 || n == 0) return EH_SUCCESS;
  if (k == 0 || alpha == T(0)) {
    if (beta == T(1)) return EH_SUCCESS;
    hipError_t e = launch_scale_c<T>(C, (int)m, (int)n, ldc, beta, stream, tri);
    return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "scale_c launch");
  }
  int config = -1;
  eh_status st = plan_config(std::is_same<T, float>::value ? EH_F32 : EH_F64, opa, opb, m, n, k, &config);
  if (st != EH_SUCCESS) return st;
  return launch_product<T>(opa != EH_OP_N, opb != EH_OP_N, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, false, 0,
                           config, stream, tri);
}
template eh_status device_product<float>(int, int, int64_t, int64_t, int64_t, float, const float*, int64_t,
                                         const float*, int64_t, float, float*, int64_t, hipStream_t, int);
template eh_status device_product<double>(int, int, int64_t, int64_t, int64_t, double, const double*, int64_t,
                                          const double*, int64_t, double, double*, int64_t, hipStream_t, int);

// y = alpha*op(A)*x + beta*y with strided vectors (BLAS ?gemv_).  A
// matrix spanning 4 GB or more is cut along the dimension that multiplies
// lda: the input dimension for op = N (later parts accumulate with beta = 1),
// the output rows for op = T.
template <typename T>
static eh_status gemv_run(bool trans, int64_t rows, int64_t k, T alpha, const T* A, int64_t lda, const T* x,
                          int64_t incx, T beta, T* y, int64_t incy, hipStream_t s) {
  if (rows == 0) return EH_SUCCESS;
  if (k == 0 || alpha == T(0)) {  // y = beta*y (A and x not read)
    if (beta == T(1)) return EH_SUCCESS;
    hipError_t e = launch_scale_c<T>(y, 1, (int)rows, incy, beta, s);
    return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "scale launch");
  }
  const uint64_t extent = ((trans ? (uint64_t)(rows - 1) * lda + k : (uint64_t)(k - 1) * lda + rows) + 64) * sizeof(T);
  if (extent >= kMaxExtent) {
    if (!trans && k > 1) {
      const int64_t h = k / 2;
      eh_status st = gemv_run<T>(false, rows, h, alpha, A, lda, x, incx, beta, y, incy, s);
      if (st != EH_SUCCESS) return st;
      return gemv_run<T>(false, rows, k - h, alpha, A + h * lda, lda, x + h * incx, incx, T(1), y, incy, s);
    }
    if (trans && rows > 1) {
      const int64_t h = rows / 2;
      eh_status st = gemv_run<T>(true, h, k, alpha, A, lda, x, incx, beta, y, incy, s);
      if (st != EH_SUCCESS) return st;
      return gemv_run<T>(true, rows - h, k, alpha, A + h * lda, lda, x, incx, beta, y + h * incy, incy, s);
    }
    return fail(EH_ERR_UNSUPPORTED, "matrix-vector operand too large for 32-bit addressing");
  }
  const GemvPlan g = gemv_plan(trans, (int)rows, (int)k, (int)sizeof(T));
  void* ws = nullptr;
  WorkspaceLease lease;
  const size_t wsb = gemv_workspace_bytes(g, (int)rows, (int)sizeof(T));
  if (wsb > 0) {
    eh_status st = workspace_get(s, wsb, &lease, &ws);
    if (st != EH_SUCCESS) return st;
  }
  KernelTimer t(EH_KC_GEMV, s);
  hipError_t e = launch_gemv<T>(g, trans, (int)rows, (int)k, alpha, A, lda, (uint32_t)(extent - 64 * sizeof(T)), x, incx,
                                beta, y, incy, ws, s);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "gemv launch");
}

// C = op(A)*op(B) (alpha 1, beta 0, leading dimensions the operands' rows)
// as the packed product of `kind` with configuration `config`: the tuner's
// timed call.  A and B are FP32 for the split kinds, 16-bit otherwise.
eh_status packed_product(int kind, int opa, int opb, int64_t m, int64_t n, int64_t k, const void* A, const void* B,
                         void* C, int config, hipStream_t s) {
  const int64_t lda = opa == EH_OP_N ? std::max<int64_t>(m, 1) : std::max<int64_t>(k, 1);
  const int64_t ldb = opb == EH_OP_N ? std::max<int64_t>(k, 1) : std::max<int64_t>(n, 1);
  const float one = 1.0f, zero = 0.0f;
  if (!packed_one_term(kind))
    return gemm_dispatch(EH_F32, opa, opb, m, n, k, &one, A, lda, B, ldb, &zero, C, std::max<int64_t>(m, 1),
                         kind == EH_PACKED_SPLIT3 ? EH_PREC_SPLIT3 : EH_PREC_SPLIT2, config, s);
  eh_status st = check_gemm_args(EH_F32, opa, opb, m, n, k, &one, A, lda, B, ldb, &zero, C, std::max<int64_t>(m, 1));
  if (st != EH_SUCCESS) return st;
  if (m == 0 || n == 0 || k == 0) return EH_SUCCESS;
  return launch_bf16_product(opa != EH_OP_N, opb != EH_OP_N, (int)m, (int)n, (int)k, 1.0f,
                             static_cast<const unsigned short*>(A), lda, static_cast<const unsigned short*>(B), ldb,
                             0.0f, static_cast<float*>(C), std::max<int64_t>(m, 1), s, kind == EH_PACKED_F16, config);
}

eh_status gemm_dispatch(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                        const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                        int64_t ldc, int prec, int config, void* stream) {
  eh_status st = check_gemm_args(dtype, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  if (st != EH_SUCCESS) return st;
  if (config >= kNumConfigs) return fail(EH_ERR_INVALID, "config out of range");
  if (prec < EH_PREC_AUTO || prec > EH_PREC_SPLIT3) return fail(EH_ERR_INVALID, "bad precision");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EH_F32)
    return run_gemm<float>(opa, opb, m, n, k, *static_cast<const float*>(alpha), static_cast<const float*>(A), lda,
                           static_cast<const float*>(B), ldb, *static_cast<const float*>(beta),
                           static_cast<float*>(C), ldc, (eh_precision)prec, config, s);
  return run_gemm<double>(opa, opb, m, n, k, *static_cast<const double*>(alpha), static_cast<const double*>(A), lda,
                          static_cast<const double*>(B), ldb, *static_cast<const double*>(beta),
                          static_cast<double*>(C), ldc, (eh_precision)prec, config, s);
}

eh_tune_mode tune_mode() {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (g_mode < 0) {
    g_mode = mlp_has_default() ? EH_TUNE_MLP : EH_TUNE_HEURISTIC;
    if (const char* e = std::getenv("EIGEN_AMD_TUNE")) {
      if (!std::strcmp(e, "heuristic")) g_mode = EH_TUNE_HEURISTIC;
      if (!std::strcmp(e, "mlp")) g_mode = EH_TUNE_MLP;
      if (!std::strcmp(e, "table")) g_mode = EH_TUNE_TABLE;
    }
  }
  return (eh_tune_mode)g_mode;
}

}  // namespace eh

using namespace eh;

extern "C" {

eh_status eh_gemm(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                  const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc,
                  eh_precision prec, void* stream) {
  return gemm_dispatch(dtype, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, prec, -1, stream);
}

eh_status eh_gemm_config(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                         const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                         int64_t ldc, eh_precision prec, int config, void* stream) {
  return gemm_dispatch(dtype, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, prec, config, stream);
}

eh_status eh_gemm_strided(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                          const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                          int64_t incc, int64_t ldc, eh_precision prec, void* stream) {
  if (incc < 1) return fail(EH_ERR_INVALID, "incc must be >= 1");
  if (incc == 1) return eh_gemm(dtype, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, prec, stream);
  eh_status st =
      check_gemm_args(dtype, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, std::max<int64_t>(1, m));
  if (st != EH_SUCCESS) return st;
  if (ldc < 1 || (m > 0 && ldc < (m - 1) * incc + 1)) return fail(EH_ERR_INVALID, "ldc too small for incc");
  if (m == 0 || n == 0) return EH_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool f64 = dtype == EH_F64;
  const size_t es = f64 ? sizeof(double) : sizeof(float);
  const bool vanishes = k == 0 || (f64 ? *static_cast<const double*>(alpha) == 0.0
                                      : *static_cast<const float*>(alpha) == 0.0f);
  void* W = nullptr;
  if (!vanishes) {
    // unscaled product into a stream-ordered temporary
    hipError_t e = hipMallocAsync(&W, (size_t)m * n * es, s);
    if (e != hipSuccess) return fail(EH_ERR_ALLOC, std::string("temporary for the strided result: ") +
                                                       hipGetErrorString(e));
    const double one64 = 1.0, zero64 = 0.0;
    const float one32 = 1.0f, zero32 = 0.0f;
    st = gemm_dispatch(dtype, opa, opb, m, n, k, f64 ? (const void*)&one64 : (const void*)&one32, A, lda, B, ldb,
                       f64 ? (const void*)&zero64 : (const void*)&zero32, W, m, prec, -1, stream);
    if (st != EH_SUCCESS) {
      (void)hipFreeAsync(W, s);
      return st;
    }
  }
  hipError_t e = f64 ? launch_strided_combine<double>(static_cast<const double*>(W), (int)m, (int)n,
                                                      static_cast<double*>(C), incc, ldc,
                                                      *static_cast<const double*>(alpha),
                                                      *static_cast<const double*>(beta), s)
                     : launch_strided_combine<float>(static_cast<const float*>(W), (int)m, (int)n,
                                                     static_cast<float*>(C), incc, ldc,
                                                     *static_cast<const float*>(alpha),
                                                     *static_cast<const float*>(beta), s);
  if (W) (void)hipFreeAsync(W, s);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "strided combine launch");
}

eh_status eh_gemv(eh_dtype dtype, eh_op trans, int64_t m, int64_t n, const void* alpha, const void* A, int64_t lda,
                  const void* x, int64_t incx, const void* beta, void* y, int64_t incy, void* stream) {
  if (dtype != EH_F32 && dtype != EH_F64) return fail(EH_ERR_INVALID, "dtype must be EH_F32 or EH_F64");
  if (!valid_op(trans)) return fail(EH_ERR_INVALID, "trans must be N, T or C");
  if (m < 0 || n < 0) return fail(EH_ERR_INVALID, "m, n must be >= 0");
  if (m > INT32_MAX || n > INT32_MAX) return fail(EH_ERR_UNSUPPORTED, "m, n must fit int32");
  if (lda < std::max<int64_t>(1, m)) return fail(EH_ERR_INVALID, "lda too small");
  if (incx == 0 || incy == 0) return fail(EH_ERR_INVALID, "incx and incy must be nonzero");
  if (!alpha || !beta) return fail(EH_ERR_INVALID, "alpha and beta must point to host scalars");
  if (m == 0 || n == 0) return EH_SUCCESS;
  const bool tr = trans != EH_OP_N;
  const int64_t rows = tr ? n : m, k = tr ? m : n;  // y has `rows` elements, x has k
  if (!y || !A || !x) return fail(EH_ERR_INVALID, "A, x or y is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto run = [&](auto zero) -> eh_status {
    using T = decltype(zero);
    // a negative increment walks the vector from its last element (BLAS)
    const T* xp = static_cast<const T*>(x) + (incx < 0 ? (1 - k) * incx : 0);
    T* yp = static_cast<T*>(y) + (incy < 0 ? (1 - rows) * incy : 0);
    return gemv_run<T>(tr, rows, k, *static_cast<const T*>(alpha), static_cast<const T*>(A), lda, xp, incx,
                       *static_cast<const T*>(beta), yp, incy, s);
  };
  return dtype == EH_F64 ? run(0.0) : run(0.0f);
}

eh_status eh_gemm_bf16(eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const float* alpha, const void* A,
                       int64_t lda, const void* B, int64_t ldb, const float* beta, float* C, int64_t ldc,
                       void* stream) {
  eh_status st = check_gemm_args(EH_F32, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  if (st != EH_SUCCESS) return st;
  if (m == 0 || n == 0) return EH_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (k == 0 || *alpha == 0.0f) {
    if (*beta == 1.0f) return EH_SUCCESS;
    hipError_t e = launch_scale_c<float>(C, (int)m, (int)n, ldc, *beta, s);
    return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "scale_c launch");
  }
  return launch_bf16_product(opa != EH_OP_N, opb != EH_OP_N, (int)m, (int)n, (int)k, *alpha,
                             static_cast<const unsigned short*>(A), lda, static_cast<const unsigned short*>(B), ldb,
                             *beta, C, ldc, s);
}

eh_status eh_gemm_f16(eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const float* alpha, const void* A,
                      int64_t lda, const void* B, int64_t ldb, const float* beta, float* C, int64_t ldc,
                      void* stream) {
  eh_status st = check_gemm_args(EH_F32, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  if (st != EH_SUCCESS) return st;
  if (m == 0 || n == 0) return EH_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (k == 0 || *alpha == 0.0f) {
    if (*beta == 1.0f) return EH_SUCCESS;
    hipError_t e = launch_scale_c<float>(C, (int)m, (int)n, ldc, *beta, s);
    return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "scale_c launch");
  }
  return launch_bf16_product(opa != EH_OP_N, opb != EH_OP_N, (int)m, (int)n, (int)k, *alpha,
                             static_cast<const unsigned short*>(A), lda, static_cast<const unsigned short*>(B), ldb,
                             *beta, C, ldc, s, true);
}

eh_status eh_syrk(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha, const void* A,
                  int64_t lda, const void* beta, void* C, int64_t ldc, void* stream) {
  if (dtype != EH_F32 && dtype != EH_F64) return fail(EH_ERR_INVALID, "dtype must be EH_F32 or EH_F64");
  if (uplo != EH_UPPER && uplo != EH_LOWER) return fail(EH_ERR_INVALID, "uplo must be EH_UPPER or EH_LOWER");
  if (!valid_op(trans)) return fail(EH_ERR_INVALID, "trans must be N, T or C");
  if (n < 0 || k < 0) return fail(EH_ERR_INVALID, "n, k must be >= 0");
  if (n > INT32_MAX || k > INT32_MAX) return fail(EH_ERR_UNSUPPORTED, "n, k must fit int32");
  if (lda < std::max<int64_t>(1, trans == EH_OP_N ? n : k)) return fail(EH_ERR_INVALID, "lda too small");
  if (ldc < std::max<int64_t>(1, n)) return fail(EH_ERR_INVALID, "ldc too small");
  if (!alpha || !beta) return fail(EH_ERR_INVALID, "alpha and beta must point to host scalars");
  if (n == 0) return EH_SUCCESS;
  if (!C || (k > 0 && !A)) return fail(EH_ERR_INVALID, "A or C is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int tri = uplo == EH_LOWER ? 1 : 2;
  const bool tr = trans != EH_OP_N;  // op(A) = A^T: C = alpha*A^T*A + beta*C
  auto run = [&](auto zero) -> eh_status {
    using T = decltype(zero);
    const T al = *static_cast<const T*>(alpha), be = *static_cast<const T*>(beta);
    const T* Ap = static_cast<const T*>(A);
    // k == 0 or alpha == 0 only scales the triangle (A is not read)
    return device_product<T>(tr ? EH_OP_T : EH_OP_N, tr ? EH_OP_N : EH_OP_T, n, n, k, al, Ap, lda, Ap, lda, be,
                             static_cast<T*>(C), ldc, s, tri);
  };
  return dtype == EH_F64 ? run(0.0) : run(0.0f);
}

int eh_num_configs(void) { return kNumConfigs; }

eh_status eh_config_info(int config, int* tile_m, int* tile_n, int* split_k) {
  if (config < 0 || config >= kNumConfigs) return fail(EH_ERR_INVALID, "config out of range");
  int tile, s;
  config_decode(config, &tile, &s);
  if (tile_m) *tile_m = kTileM[tile];
  if (tile_n) *tile_n = kTileN[tile];
  if (split_k) *split_k = s;
  return EH_SUCCESS;
}

eh_status eh_plan(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, int* config) {
  if (!config) return fail(EH_ERR_INVALID, "config is NULL");
  if (dtype != EH_F32 && dtype != EH_F64) return fail(EH_ERR_INVALID, "bad dtype");
  if (m < 0 || n < 0 || k < 0) return fail(EH_ERR_INVALID, "m, n, k must be >= 0");
  return plan_config(dtype, opa, opb, m, n, k, config);
}

eh_status eh_set_tune_mode(eh_tune_mode mode) {
  if (mode != EH_TUNE_HEURISTIC && mode != EH_TUNE_MLP && mode != EH_TUNE_TABLE)
    return fail(EH_ERR_INVALID, "bad tune mode");
  std::lock_guard<std::mutex> lk(g_mutex);
  g_mode = mode;
  return EH_SUCCESS;
}

eh_tune_mode eh_get_tune_mode(void) { return tune_mode(); }

eh_status eh_set_sgemm_precision(eh_precision prec) {
  if (prec < EH_PREC_AUTO || prec > EH_PREC_SPLIT3) return fail(EH_ERR_INVALID, "bad precision");
  std::lock_guard<std::mutex> lk(g_mutex);
  g_sgemm_prec = prec == EH_PREC_AUTO ? (int)EH_PREC_PLAIN : (int)prec;
  return EH_SUCCESS;
}

eh_precision eh_get_sgemm_precision(void) { return default_sgemm_precision(); }

eh_status eh_fill(eh_dtype dtype, void* x, int64_t n, uint64_t seed, int dist, double p0, double p1, void* stream) {
  if (dtype != EH_F32 && dtype != EH_F64) return fail(EH_ERR_INVALID, "bad dtype");
  if (n < 0 || (n > 0 && !x)) return fail(EH_ERR_INVALID, "bad buffer");
  if (dist < 0 || dist > 3) return fail(EH_ERR_INVALID, "bad distribution");
  if (n == 0) return EH_SUCCESS;
  hipError_t e = launch_fill(dtype, x, n, seed, dist, p0, p1, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "fill launch");
}

eh_status eh_split_bf16(const float* x, int64_t n, int terms, uint16_t* t0, uint16_t* t1, uint16_t* t2,
                        void* stream) {
  if (terms < 1 || terms > 3) return fail(EH_ERR_INVALID, "terms must be 1, 2 or 3");
  if (n < 0) return fail(EH_ERR_INVALID, "n < 0");
  if (n > 0 && (!x || !t0 || (terms > 1 && !t1) || (terms > 2 && !t2))) return fail(EH_ERR_INVALID, "NULL buffer");
  if (n == 0) return EH_SUCCESS;
  hipError_t e = launch_split_bf16(x, n, terms, t0, t1, t2, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "split launch");
}

eh_status eh_synchronize(void* stream) {
  hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "hipStreamSynchronize");
}

const char* eh_status_string(eh_status s) {
  switch (s) {
    case EH_SUCCESS: return "success";
    case EH_ERR_INVALID: return "invalid argument";
    case EH_ERR_HIP: return "HIP runtime error";
    case EH_ERR_UNSUPPORTED: return "unsupported";
    case EH_ERR_ALLOC: return "allocation failed";
  }
  return "unknown status";
}

const char* eh_last_error(void) { return g_last_error.c_str(); }

eh_status eh_kernel_timing(int enable) {
  std::lock_guard<std::mutex> lk(g_timing_mutex);
  timing_clear_locked();
  g_timing_on.store(enable ? 1 : 0);
  return EH_SUCCESS;
}

eh_status eh_kernel_timing_read(eh_kernel_class cls, double* total_ms, int64_t* launches) {
  if ((int)cls < 0 || (int)cls >= EH_KC_COUNT) return fail(EH_ERR_INVALID, "bad kernel class");
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  {
    std::lock_guard<std::mutex> lk(g_timing_mutex);
    ev.swap(g_timed[cls]);
  }
  double sum = 0.0;
  eh_status st = EH_SUCCESS;
  for (auto& p : ev) {
    float ms = 0.0f;
    hipError_t e = hipEventSynchronize(p.second);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.first, p.second);
    if (e != hipSuccess && st == EH_SUCCESS) st = hip_fail(e, "kernel timing event");
    sum += ms;
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  if (total_ms) *total_ms = sum;
  if (launches) *launches = (int64_t)ev.size();
  return st;
}

const char* eh_version(void) { return "eigen_amd 0.1 gfx950"; }

}  // extern "C"
