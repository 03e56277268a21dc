// diag.hip -- kernel-level diagnostics (diag.h), built into
// libuncertainty_amd_diag.so only.  Stateless: every call allocates its own
// device buffers (freed on every return path) on one stream of device 0.
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "diag.h"

namespace unc {
}  // namespace unc

using namespace unc;

namespace {

constexpr int GJ_BLOCK = 64;

hipStream_t diag_stream(int *rc) {
  static std::once_flag once;
  static hipStream_t stream = nullptr;
  static int init_rc = UNC_OK;
  std::call_once(once, [] {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || hipSetDevice(0) != hipSuccess ||
        hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      init_rc = UNC_ERR_DEVICE;
    }
  });
  *rc = init_rc;
  return stream;
}

// device buffer released on scope exit
struct DevBuf {
  char *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (hipMalloc((void **)&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return UNC_ERR_OOM;
    }
    return UNC_OK;
  }
};

struct Events {
  hipEvent_t a = nullptr, b = nullptr;
  ~Events() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

#define DC(expr)                                                                          \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      std::fprintf(stderr, "uncertainty_amd diag: %s failed (%s:%d): %s\n", #expr, __FILE__, \
                   __LINE__, hipGetErrorString(e_));                                      \
      return UNC_ERR_DEVICE;                                                              \
    }                                                                                     \
  } while (0)

#define DSTREAM(s)                  \
  int rc_ = UNC_OK;                 \
  hipStream_t s = diag_stream(&rc_); \
  if (rc_) return rc_;              \
  DC(hipSetDevice(0))

__global__ void k_fill_pattern(size_t n, double *A, double scale) {
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < n;
       x += (size_t)gridDim.x * blockDim.x)
    A[x] = scale * (double)((x * 2654435761ull) % 1000003ull) / 1000003.0 - 0.5 * scale;
}

}  // namespace

extern "C" {

int unc_diag_gemm_f64(int M, int N, int K, double alpha, const double *A, int lda, const double *B,
                      int ldb, double beta, double *C, int ldc, double *absmax_out) {
  if (M <= 0 || N <= 0 || K <= 0 || lda < M || ldb < K || ldc < M) return UNC_ERR_BAD_INPUT;
  DSTREAM(s);
  const size_t na = (size_t)lda * K, nb = (size_t)ldb * N, nc = (size_t)ldc * N;
  DevBuf buf;
  if (buf.alloc(sizeof(double) * (na + nb + nc) + 256)) return UNC_ERR_OOM;
  double *dA = (double *)buf.p, *dB = dA + na, *dC = dB + nb;
  unsigned long long *dm = (unsigned long long *)(dC + nc);
  DC(hipMemcpyAsync(dA, A, sizeof(double) * na, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(dB, B, sizeof(double) * nb, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(dC, C, sizeof(double) * nc, hipMemcpyHostToDevice, s));
  DC(hipMemsetAsync(dm, 0, sizeof(unsigned long long), s));
  DC(gemm_f64(M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, dm, nullptr, s));
  unsigned long long mb = 0;
  DC(hipMemcpyAsync(C, dC, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
  DC(hipMemcpyAsync(&mb, dm, sizeof(mb), hipMemcpyDeviceToHost, s));
  DC(hipStreamSynchronize(s));
  if (absmax_out) std::memcpy(absmax_out, &mb, sizeof(double));
  return UNC_OK;
}

int unc_diag_gemm_f64_split(int M, int N, int K, double alpha, const double *A, int lda,
                            const double *B, int ldb, double *C, int ldc, double *absmax_out) {
  if (M <= 0 || N <= 0 || K <= 0 || lda < M || ldb < K || ldc < M) return UNC_ERR_BAD_INPUT;
  if (!gemm_split_applies(M, N, K)) return UNC_ERR_UNSUPPORTED;
  DSTREAM(s);
  const size_t na = (size_t)lda * K, nb = (size_t)ldb * N, nc = (size_t)ldc * N,
               nw = (size_t)M * N;
  DevBuf buf;
  if (buf.alloc(sizeof(double) * (na + nb + nc + nw) + 256)) return UNC_ERR_OOM;
  double *dA = (double *)buf.p, *dB = dA + na, *dC = dB + nb, *dW = dC + nc;
  unsigned long long *dm = (unsigned long long *)(dW + nw);
  DC(hipMemcpyAsync(dA, A, sizeof(double) * na, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(dB, B, sizeof(double) * nb, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(dC, C, sizeof(double) * nc, hipMemcpyHostToDevice, s));
  DC(hipMemsetAsync(dm, 0, sizeof(unsigned long long), s));
  DC(gemm_f64_split(M, N, K, alpha, dA, lda, dB, ldb, dC, ldc, dW, dm, nullptr, s));
  unsigned long long mb = 0;
  DC(hipMemcpyAsync(C, dC, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
  DC(hipMemcpyAsync(&mb, dm, sizeof(mb), hipMemcpyDeviceToHost, s));
  DC(hipStreamSynchronize(s));
  if (absmax_out) std::memcpy(absmax_out, &mb, sizeof(double));
  return UNC_OK;
}

int unc_diag_gemm_f64_sym_batch(int n, int nb, const double *A, const double *B, double *C,
                                double *absmax_out, int reps, double *avg_ms) {
  if (n <= 0 || nb < 1 || nb > 4 || !A || !B || !C || reps < 0) return UNC_ERR_BAD_INPUT;
  DSTREAM(s);
  const size_t nn = (size_t)n * n;
  DevBuf buf;
  if (buf.alloc(sizeof(double) * (1 + 2 * (size_t)nb) * nn + 256)) return UNC_ERR_OOM;
  double *dA = (double *)buf.p, *dB = dA + nn, *dC = dB + nb * nn;
  unsigned long long *dm = (unsigned long long *)(dC + nb * nn);
  DC(hipMemcpyAsync(dA, A, sizeof(double) * nn, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(dB, B, sizeof(double) * nb * nn, hipMemcpyHostToDevice, s));
  DC(hipMemsetAsync(dC, 0, sizeof(double) * nb * nn, s));
  DC(hipMemsetAsync(dm, 0, sizeof(unsigned long long) * 4, s));
  const double *bq[4];
  double *cq[4];
  for (int q = 0; q < 4; ++q) {
    bq[q] = dB + (size_t)(q < nb ? q : 0) * nn;
    cq[q] = dC + (size_t)(q < nb ? q : 0) * nn;
  }
  DC(gemm_f64_sym_batch(n, nb, dA, bq, cq, dm, nullptr, s));
  unsigned long long mb[4] = {0, 0, 0, 0};
  DC(hipMemcpyAsync(C, dC, sizeof(double) * nb * nn, hipMemcpyDeviceToHost, s));
  DC(hipMemcpyAsync(mb, dm, sizeof(mb), hipMemcpyDeviceToHost, s));
  DC(hipStreamSynchronize(s));
  if (absmax_out) std::memcpy(absmax_out, mb, sizeof(double) * nb);
  if (avg_ms && reps > 0) {
    Events ev;
    DC(hipEventCreate(&ev.a));
    DC(hipEventCreate(&ev.b));
    DC(hipEventRecord(ev.a, s));
    for (int r = 0; r < reps; ++r) DC(gemm_f64_sym_batch(n, nb, dA, bq, cq, dm, nullptr, s));
    DC(hipEventRecord(ev.b, s));
    DC(hipEventSynchronize(ev.b));
    float ms = 0.0f;
    DC(hipEventElapsedTime(&ms, ev.a, ev.b));
    *avg_ms = ms / reps;
  }
  return UNC_OK;
}

int unc_diag_gemm_f64_sym(int n, double alpha, const double *A, const double *B, double *C,
                          double *absmax_out, int parts) {
  if (n <= 0 || !A || !B || !C || parts < 0 || parts > 4) return UNC_ERR_BAD_INPUT;
  if (!gemm_split_applies(n, n, n)) return UNC_ERR_UNSUPPORTED;
  DSTREAM(s);
  const size_t nn = (size_t)n * n, nw = gemm_sym_split_work_doubles(n);
  DevBuf buf;
  if (buf.alloc(sizeof(double) * (3 * nn + nw) + 256)) return UNC_ERR_OOM;
  double *dA = (double *)buf.p, *dB = dA + nn, *dC = dB + nn, *dW = dC + nn;
  unsigned long long *dm = (unsigned long long *)(dW + nw);
  DC(hipMemcpyAsync(dA, A, sizeof(double) * nn, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(dB, B, sizeof(double) * nn, hipMemcpyHostToDevice, s));
  DC(hipMemsetAsync(dC, 0, sizeof(double) * nn, s));
  DC(hipMemsetAsync(dm, 0, sizeof(unsigned long long), s));
  DC(gemm_f64_sym(n, alpha, dA, dB, dC, dW, nw, dm, nullptr, s, parts));
  unsigned long long mb = 0;
  DC(hipMemcpyAsync(C, dC, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
  DC(hipMemcpyAsync(&mb, dm, sizeof(mb), hipMemcpyDeviceToHost, s));
  DC(hipStreamSynchronize(s));
  if (absmax_out) std::memcpy(absmax_out, &mb, sizeof(double));
  return UNC_OK;
}

int unc_diag_inverse(int n, const double *A, double *X, int refine) {
  if (n <= 0 || !A || !X) return UNC_ERR_BAD_INPUT;
  DSTREAM(s);
  const size_t nn = (size_t)n * n;
  DevBuf buf;
  if (buf.alloc(sizeof(double) * (5 * nn + (size_t)GJ_BLOCK * (3 * GJ_BLOCK + 2 * n)) + 256))
    return UNC_ERR_OOM;
  double *dA = (double *)buf.p, *dW = dA + nn, *dR = dW + nn, *dD = dR + nn, *dX = dD + nn;
  double *P = dX + nn, *T = P + 3 * GJ_BLOCK * GJ_BLOCK, *Rp = T + (size_t)GJ_BLOCK * n;
  unsigned int *err = (unsigned int *)(Rp + (size_t)GJ_BLOCK * n);
  DC(hipMemsetAsync(err, 0, sizeof(unsigned int), s));
  DC(hipMemcpyAsync(dA, A, sizeof(double) * nn, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(dW, dA, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
  int rc = dense_inverse_gj(n, dW, P, T, Rp, err, s);
  if (rc) return rc;
  double *res = dW;
  if (refine) {
    DC(residual_ex(n, dA, dW, dR, dD, s));  // dD is free until the product
    DC(gemm_f64(n, n, n, 1.0, dW, n, dR, n, 0.0, dD, n, nullptr, nullptr, s));
    DC(sym_update(n, dW, dD, dX, s));
    res = dX;
  }
  unsigned int herr = 0;
  DC(hipMemcpyAsync(X, res, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
  DC(hipMemcpyAsync(&herr, err, sizeof(herr), hipMemcpyDeviceToHost, s));
  DC(hipStreamSynchronize(s));
  return herr ? error_code(herr) : UNC_OK;
}

int unc_diag_residual(int n, const double *A, const double *X, double *R, int tile_cols) {
  if (n <= 0 || !A || !X || !R) return UNC_ERR_BAD_INPUT;
  if (tile_cols != 0 && tile_cols != 16 && tile_cols != 32) return UNC_ERR_BAD_INPUT;
  DSTREAM(s);
  const size_t nn = (size_t)n * n;
  DevBuf buf;
  if (buf.alloc(sizeof(double) * (3 * nn + 2 * (size_t)n))) return UNC_ERR_OOM;
  double *b = (double *)buf.p;
  DC(hipMemcpyAsync(b, A, sizeof(double) * nn, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(b + nn, X, sizeof(double) * nn, hipMemcpyHostToDevice, s));
  DC(residual_ex(n, b, b + nn, b + 2 * nn, b + 3 * nn, s, tile_cols));
  DC(hipMemcpyAsync(R, b + 2 * nn, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
  DC(hipStreamSynchronize(s));
  return UNC_OK;
}

int unc_diag_residual_crt(int n, const double *A, const double *X, double *R, int blocks,
                          int reps, double *avg_ms) {
  if (n <= 0 || !A || !X || !R || reps < 0 || blocks < 1 || blocks > 64) return UNC_ERR_BAD_INPUT;
  DSTREAM(s);
  const size_t nn = (size_t)n * n, wb = residual_crt_work_bytes(n, blocks);
  DevBuf buf;
  if (buf.alloc(sizeof(double) * 3 * nn + wb + 256)) return UNC_ERR_OOM;
  double *b = (double *)buf.p;
  void *work = (void *)(((uintptr_t)(b + 3 * nn) + 255) & ~(uintptr_t)255);
  DC(hipMemcpyAsync(b, A, sizeof(double) * nn, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(b + nn, X, sizeof(double) * nn, hipMemcpyHostToDevice, s));
  DC(residual_crt(n, b, b + nn, b + 2 * nn, work, wb, s));
  DC(hipMemcpyAsync(R, b + 2 * nn, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
  DC(hipStreamSynchronize(s));
  if (avg_ms && reps > 0) {
    Events ev;
    DC(hipEventCreate(&ev.a));
    DC(hipEventCreate(&ev.b));
    DC(hipEventRecord(ev.a, s));
    for (int r = 0; r < reps; ++r) DC(residual_crt(n, b, b + nn, b + 2 * nn, work, wb, s));
    DC(hipEventRecord(ev.b, s));
    DC(hipEventSynchronize(ev.b));
    float ms = 0.0f;
    DC(hipEventElapsedTime(&ms, ev.a, ev.b));
    *avg_ms = ms / reps;
  }
  return UNC_OK;
}

int unc_diag_gemm_ex(int M, int N, int K, double alpha, const double *A, int lda,
                     const double *B, int ldb, int trans_b, double beta, double *C, int ldc,
                     int tri, int nmod, int reps, double *avg_ms) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C || lda < M || ldc < M ||
      ldb < (trans_b ? N : K) || reps < 0 || (nmod != 0 && crt_bits(nmod) == 0))
    return UNC_ERR_BAD_INPUT;
  DSTREAM(s);
  const size_t na = (size_t)lda * K, nb = (size_t)ldb * (trans_b ? K : N), nc = (size_t)ldc * N;
  const size_t wb = nmod ? crt_gemm_work_bytes(M, N, K, nmod) : 0;
  DevBuf buf;
  if (buf.alloc(sizeof(double) * (na + nb + 2 * nc) + wb + 512)) return UNC_ERR_OOM;
  double *dA = (double *)buf.p, *dB = dA + na, *dC = dB + nb, *dC0 = dC + nc;
  void *work = (void *)(((uintptr_t)(dC0 + nc) + 255) & ~(uintptr_t)255);
  DC(hipMemcpyAsync(dA, A, sizeof(double) * na, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(dB, B, sizeof(double) * nb, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(dC, C, sizeof(double) * nc, hipMemcpyHostToDevice, s));
  DC(hipMemcpyAsync(dC0, dC, sizeof(double) * nc, hipMemcpyDeviceToDevice, s));
  auto call = [&]() -> hipError_t {
    return nmod ? gemm_crt(M, N, K, alpha, dA, lda, dB, ldb, trans_b != 0, beta, dC, ldc, tri,
                           nmod, work, wb, s)
                : gemm_f64_ex(M, N, K, alpha, dA, lda, dB, ldb, trans_b != 0, beta, dC, ldc, tri,
                              s);
  };
  DC(call());
  DC(hipMemcpyAsync(C, dC, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
  DC(hipStreamSynchronize(s));
  if (avg_ms && reps > 0) {
    Events ev;
    DC(hipEventCreate(&ev.a));
    DC(hipEventCreate(&ev.b));
    DC(hipEventRecord(ev.a, s));
    for (int r = 0; r < reps; ++r) DC(call());
    DC(hipEventRecord(ev.b, s));
    DC(hipEventSynchronize(ev.b));
    float ms = 0.f;
    DC(hipEventElapsedTime(&ms, ev.a, ev.b));
    *avg_ms = ms / reps;
  }
  return UNC_OK;
}

int unc_diag_chol_inverse(int n, const double *A, double *L, double *X) {
  if (n <= 0 || !A || !L || !X) return UNC_ERR_BAD_INPUT;
  DSTREAM(s);
  const size_t nn = (size_t)n * n;
  DevBuf buf;
  if (buf.alloc(sizeof(double) * 3 * nn + 256)) return UNC_ERR_OOM;
  double *dA = (double *)buf.p, *dX = dA + nn, *dW = dX + nn;
  unsigned int *err = (unsigned int *)(dW + nn);
  DC(hipMemsetAsync(err, 0, sizeof(unsigned int), s));
  DC(hipMemsetAsync(dX, 0, sizeof(double) * nn, s));
  DC(hipMemcpyAsync(dA, A, sizeof(double) * nn, hipMemcpyHostToDevice, s));
  int rc = cholesky_inverse(n, dA, dX, dW, err, s);
  if (rc) return rc;
  unsigned int herr = 0;
  DC(hipMemcpyAsync(L, dA, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
  DC(hipMemcpyAsync(X, dX, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
  DC(hipMemcpyAsync(&herr, err, sizeof(herr), hipMemcpyDeviceToHost, s));
  DC(hipStreamSynchronize(s));
  return herr ? error_code(herr) : UNC_OK;
}

int unc_diag_gemm_bench(int n, int reps, int form, int parts, double *avg_ms) {
  if (n <= 0 || reps <= 0 || !avg_ms || form < 0 || form > 2 || parts < 0 || parts > 4)
    return UNC_ERR_BAD_INPUT;
  DSTREAM(s);
  const size_t nn = (size_t)n * n;
  DevBuf buf;
  if (buf.alloc(sizeof(double) * 6 * nn)) return UNC_ERR_OOM;
  double *b = (double *)buf.p;
  Events ev;
  DC(hipEventCreate(&ev.a));
  DC(hipEventCreate(&ev.b));
  hipLaunchKernelGGL(k_fill_pattern, dim3(4096), dim3(256), 0, s, 2 * nn, b, 1.0);
  auto one = [&]() {
    if (form == 2)
      return gemm_f64_sym(n, 1.0, b, b + nn, b + 2 * nn, b + 3 * nn, 3 * nn, nullptr, nullptr, s,
                          parts);
    if (form == 1)
      return gemm_f64_split(n, n, n, 1.0, b, n, b + nn, n, b + 2 * nn, n, b + 3 * nn, nullptr,
                            nullptr, s);
    return gemm_f64(n, n, n, 1.0, b, n, b + nn, n, 0.0, b + 2 * nn, n, nullptr, nullptr, s);
  };
  DC(one());
  DC(hipEventRecord(ev.a, s));
  for (int r = 0; r < reps; ++r) DC(one());
  DC(hipEventRecord(ev.b, s));
  DC(hipEventSynchronize(ev.b));
  float ms = 0.f;
  DC(hipEventElapsedTime(&ms, ev.a, ev.b));
  *avg_ms = ms / reps;
  return UNC_OK;
}

int unc_diag_residual_bench(int n, int reps, int tile_cols, double *avg_ms) {
  if (n <= 0 || reps <= 0 || !avg_ms) return UNC_ERR_BAD_INPUT;
  if (tile_cols != 0 && tile_cols != 16 && tile_cols != 32) return UNC_ERR_BAD_INPUT;
  DSTREAM(s);
  const size_t nn = (size_t)n * n;
  DevBuf buf;
  if (buf.alloc(sizeof(double) * (3 * nn + 2 * (size_t)n))) return UNC_ERR_OOM;
  double *b = (double *)buf.p;
  Events ev;
  DC(hipEventCreate(&ev.a));
  DC(hipEventCreate(&ev.b));
  hipLaunchKernelGGL(k_fill_pattern, dim3(4096), dim3(256), 0, s, 2 * nn, b, 1.0);
  DC(residual_ex(n, b, b + nn, b + 2 * nn, b + 3 * nn, s, tile_cols));
  DC(hipEventRecord(ev.a, s));
  for (int r = 0; r < reps; ++r) DC(residual_ex(n, b, b + nn, b + 2 * nn, b + 3 * nn, s, tile_cols));
  DC(hipEventRecord(ev.b, s));
  DC(hipEventSynchronize(ev.b));
  float ms = 0.f;
  DC(hipEventElapsedTime(&ms, ev.a, ev.b));
  *avg_ms = ms / reps;
  return UNC_OK;
}

int unc_diag_leaf_bench(int reps, int variant, double *avg_us) {
  if (reps <= 0 || !avg_us || variant < 0 || variant > 1) return UNC_ERR_BAD_INPUT;
  DSTREAM(s);
  const int n = 64;
  std::vector<double> h((size_t)n * n);
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < n; ++r) h[r + (size_t)c * n] = (r == c) ? n + 1.0 : 1.0 / (1.0 + r + c);
  DevBuf buf;
  if (buf.alloc(sizeof(double) * 3 * n * n + 256)) return UNC_ERR_OOM;
  double *dA = (double *)buf.p, *dW = dA + n * n, *dX = dW + n * n;
  unsigned int *err = (unsigned int *)(dX + n * n);
  DC(hipMemcpyAsync(dA, h.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, s));
  DC(hipMemsetAsync(err, 0, sizeof(unsigned int), s));
  Events ev;
  DC(hipEventCreate(&ev.a));
  DC(hipEventCreate(&ev.b));
  auto one = [&]() -> hipError_t {
    if (variant == 1) return spd_inverse64(n, dA, n, dX, err, s);
    hipError_t e = hipMemcpyAsync(dW, dA, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s);
    return e != hipSuccess ? e : chol_inv_leaf(n, dW, n, dX, n, err, s);
  };
  DC(one());
  DC(hipEventRecord(ev.a, s));
  for (int r = 0; r < reps; ++r) DC(one());
  DC(hipEventRecord(ev.b, s));
  DC(hipEventSynchronize(ev.b));
  float ms = 0.f;
  DC(hipEventElapsedTime(&ms, ev.a, ev.b));
  *avg_us = 1e3 * ms / reps;
  return UNC_OK;
}

}  // extern "C"
