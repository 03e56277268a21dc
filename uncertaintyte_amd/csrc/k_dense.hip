// k_dense.hip -- (d) dense FP64 kernels of the Z inverse + Taylor chain, gfx950.
//
// Reference: teInverse (compute.cpp:121-166) with SDM::inv (LAPACK dgetrf +
// dgetri, ScaledDenseMatrix.cpp:232-279), SDM::operator*= (dgemm,
// ScaledDenseMatrix.cpp:314-353), absMax (:207-216) and iZupdate
// (compute.cpp:64-71).
//
// gemm_f64: C = alpha A B + beta C (column-major, no transposes) on the FP64
// matrix cores (v_mfma_f64_16x16x4_f64).  128x128 tiles, 8 waves of 64x32
// (k_gemm_f64w8), K staged through LDS 16 at a time with the next slab prefetched
// into registers.  Optional fused epilogue: max|C| folded into one 64-bit atomicMax
// per workgroup (non-negative doubles order like their bit patterns, so the
// result is order independent), and an early exit when the Taylor control
// block says the series has stopped (so the whole chain is enqueued at once).
#include "common.h"
#include "gemm_tile.h"

namespace unc {

constexpr int G_THREADS = 256;
// products with fewer 128 x 128 tiles than this take the 64 x 64 kernel
constexpr int SMALL_GEMM_TILES = 96;

// 8-wave kernel (production): 128x128 tile, 8 waves of 64x32 (two waves per
// SIMD, so one wave's LDS and barrier waits are covered by the other's MFMAs),
// K slabs of BK staged through LDS with the next slab prefetched into
// registers.  Same operands, flags and epilogue as k_gemm_f64, and the same
// per-element accumulation order (k ascending in steps of 4), so the two give
// identical results.
//
// SPLIT: the grid holds two workgroups per tile, each summing one half of the
// K range (split at a multiple of BK); the first half goes to C and the second
// to `part1` (M x N, leading dimension M), both unscaled, and
// k_split_reduce applies the epilogue.  Used where the tiles alone would leave
// CUs idle (N = 1800 gives 225 tiles for 256 CUs, two resident workgroups each).
//
// The tile body is gemm_w8_tile (gemm_tile.h).
template <int BK, bool TRANS_B = false, bool SPLIT = false>
__global__ __launch_bounds__(512, 1) void k_gemm_f64w8(int M, int N, int K, double alpha,
                                                      const double *__restrict__ A, int lda,
                                                      const double *__restrict__ B, int ldb,
                                                      double beta, double *__restrict__ C,
                                                      int ldc, unsigned long long *absmax,
                                                      const int *active, int tri,
                                                      double *__restrict__ part1, int parts) {
  if (active && *active == 0) return;
  gemm_w8_tile<BK, TRANS_B, SPLIT>(blockIdx.x, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                   absmax, tri, part1, parts, 0, 0);
}

// Epilogue of the split product: C = alpha (C + W), with the max|C| fold and
// the early exit of the unsplit kernel.  k_split_reduce_flat takes C stored
// densely (ldc == M, even size, 16-byte aligned) as one array and moves two
// double2 per thread per step; k_split_reduce goes column by column;
// k_split_reduce_sym is the symmetric form.
__device__ inline void fold_absmax(double lmax, unsigned long long *absmax, double *wmax) {
  const int tid = threadIdx.x;
  for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off));
  if ((tid & 63) == 0) wmax[tid >> 6] = lmax;
  __syncthreads();
  if (tid == 0) {
    const double m = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    atomicMax(absmax, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ __launch_bounds__(256) void k_split_reduce_flat(size_t n4, double alpha,
                                                           double *__restrict__ C,
                                                           const double *__restrict__ W,
                                                           unsigned long long *absmax,
                                                           const int *active) {
  if (active && *active == 0) return;
  __shared__ double wmax[4];
  double2 *c2 = (double2 *)C;
  const double2 *w2 = (const double2 *)W;
  double lmax = 0.0;
  // n4 = number of 4-element groups
  for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n4; q += (size_t)gridDim.x * 256) {
    const double2 a0 = c2[2 * q], a1 = c2[2 * q + 1];
    const double2 b0 = w2[2 * q], b1 = w2[2 * q + 1];
    double2 r0, r1;
    r0.x = alpha * (a0.x + b0.x);
    r0.y = alpha * (a0.y + b0.y);
    r1.x = alpha * (a1.x + b1.x);
    r1.y = alpha * (a1.y + b1.y);
    c2[2 * q] = r0;
    c2[2 * q + 1] = r1;
    lmax = fmax(lmax, fmax(fmax(fabs(r0.x), fabs(r0.y)), fmax(fabs(r1.x), fabs(r1.y))));
  }
  if (absmax) fold_absmax(lmax, absmax, wmax);
}

// Symmetric epilogue for an n x n product computed on the 128 x 128 tiles on
// and below the diagonal in `parts` K chunks (C and W_1 .. W_{parts-1}, each
// n x n, ld n): one workgroup per 64 x 64 block (bi >= bj) forms
// v = alpha (C + W_1 + ...) and writes it to block (bi, bj) and its transpose
// to block (bj, bi); in diagonal blocks the lower triangle is mirrored, so C is
// exactly symmetric.  max|C| is folded over the lower triangle.
template <int PARTS>
__global__ __launch_bounds__(256) void k_split_reduce_sym(int n, double alpha,
                                                          double *__restrict__ C,
                                                          const double *__restrict__ W,
                                                          unsigned long long *absmax,
                                                          const int *active) {
  if (active && *active == 0) return;
  __shared__ double V[64][65];
  __shared__ double wmax[4];
  // block index -> (bi, bj), bi >= bj, in row-major order of the lower triangle
  const int x = blockIdx.x;
  int bi = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= x) ++bi;
  while (bi * (bi + 1) / 2 > x) --bi;
  const int bj = x - bi * (bi + 1) / 2;
  const int i0 = bi * 64, j0 = bj * 64, tid = threadIdx.x;
  const size_t nn = (size_t)n * n;
  // every load of the block is issued before the first sum: 16 elements per
  // thread, chunk by chunk, each added in chunk order
  constexpr int E = 64 * 64 / 256;
  double v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = tid + 256 * q, gi = i0 + (e & 63), gj = j0 + (e >> 6);
    v[q] = (gi < n && gj < n) ? C[(size_t)gi + (size_t)gj * n] : 0.0;
  }
#pragma unroll
  for (int p = 1; p < PARTS; ++p)
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int e = tid + 256 * q, gi = i0 + (e & 63), gj = j0 + (e >> 6);
      if (gi < n && gj < n) v[q] += W[(size_t)gi + (size_t)gj * n + (size_t)(p - 1) * nn];
    }
  double lmax = 0.0;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = tid + 256 * q, r = e & 63, c = e >> 6;
    const double s = alpha * v[q];
    if (i0 + r < n && j0 + c < n && (bi != bj || r >= c)) lmax = fmax(lmax, fabs(s));
    V[r][c] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = tid + 256 * q, r = e & 63, c = e >> 6, gi = i0 + r, gj = j0 + c;
    if (bi == bj) {  // mirror the lower triangle of the diagonal block
      if (gi < n && gj < n) C[(size_t)gi + (size_t)gj * n] = r >= c ? V[r][c] : V[c][r];
    } else {
      if (gi < n && gj < n) C[(size_t)gi + (size_t)gj * n] = V[r][c];
      // block (bj, bi): element (j0 + r, i0 + c) = V[c][r]
      const int ti = j0 + r, tj = i0 + c;
      if (ti < n && tj < n) C[(size_t)ti + (size_t)tj * n] = V[c][r];
    }
  }
  if (absmax) fold_absmax(lmax, absmax, wmax);
}

__global__ __launch_bounds__(256) void k_split_reduce(int M, double alpha, double *__restrict__ C,
                                                      int ldc, const double *__restrict__ W,
                                                      unsigned long long *absmax,
                                                      const int *active) {
  if (active && *active == 0) return;
  __shared__ double wmax[4];
  const int col = blockIdx.x, tid = threadIdx.x;
  double *c = C + (size_t)col * ldc;
  const double *w = W + (size_t)col * M;
  double lmax = 0.0;
  for (int row = tid; row < M; row += 256) {
    const double v = alpha * (c[row] + w[row]);
    c[row] = v;
    lmax = fmax(lmax, fabs(v));
  }
  if (absmax) fold_absmax(lmax, absmax, wmax);
}


hipError_t gemm_f64(int M, int N, int K, double alpha, const double *A, int lda, const double *B,
                    int ldb, double beta, double *C, int ldc, unsigned long long *absmax,
                    const int *active, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  const int tiles = (int)cdiv(M, GB_M) * (int)cdiv(N, GB_N);
  if (!absmax && !active && tiles < SMALL_GEMM_TILES)
    return gemm_f64_ex(M, N, K, alpha, A, lda, B, ldb, false, beta, C, ldc, 0, st);
  // 8 waves, K slabs of 16
  hipLaunchKernelGGL((k_gemm_f64w8<16>), dim3(tiles), dim3(512), 0, st, M, N, K, alpha, A, lda, B,
                     ldb, beta, C, ldc, absmax, active, 0, nullptr, 1);
  return hipGetLastError();
}

// Split-K applies while the doubled grid still fits the resident slots (two
// 512-thread workgroups per CU on 256 CUs) and each half keeps >= 512 of K.
constexpr int SPLIT_MAX_TILES = 256, SPLIT_MIN_K = 1024;
// The symmetric form launches only the tiles on and below the diagonal, in
// 4 x 4 tile blocks with each XCD taking a contiguous range (so the
// workgroups that share A and B panels share an L2), and splits K into four
// chunks while that keeps the grid within the 512 resident slots (two per CU),
// else three.  N = 1800: 120 tiles, 480 workgroups; rings200 chain product
// 0.213 ms with all tiles launched and K in three chunks, 0.186 ms this way.
// The scratch is sized for up to SYM_PARTS chunks.
constexpr int SYM_PARTS = 4;

bool gemm_split_applies(int M, int N, int K) {
  const long long tiles = cdiv(M, GB_M) * cdiv(N, GB_N);
  return tiles >= SMALL_GEMM_TILES && tiles <= SPLIT_MAX_TILES && K >= SPLIT_MIN_K;
}

size_t gemm_sym_split_work_doubles(int n) {
  return gemm_split_applies(n, n, n) ? (size_t)(SYM_PARTS - 1) * n * n : 0;
}

hipError_t gemm_f64_sym(int n, double alpha, const double *A, const double *B, double *C,
                        double *work, size_t work_doubles, unsigned long long *absmax,
                        const int *active, hipStream_t st, int parts) {
  if (n <= 0) return hipSuccess;
  if (!work || work_doubles < gemm_sym_split_work_doubles(n) || !gemm_split_applies(n, n, n))
    return gemm_f64(n, n, n, alpha, A, n, B, n, 0.0, C, n, absmax, active, st);
  const int tm = (int)cdiv(n, GB_M), lower = tm * (tm + 1) / 2;
  if (parts <= 0 || parts > SYM_PARTS) parts = SYM_PARTS * lower <= 512 ? SYM_PARTS : 3;
  hipLaunchKernelGGL((k_gemm_f64w8<16, false, true>), dim3(parts * lower), dim3(512), 0, st, n,
                     n, n, 1.0, A, n, B, n, 0.0, C, n, nullptr, active,
                     GEMM_C_LOWER | GEMM_LOWER_LIST, work, parts);
  const int nb = (int)cdiv(n, 64);
  const dim3 rgrid(nb * (nb + 1) / 2);
  switch (parts) {
    case 1:
      hipLaunchKernelGGL(k_split_reduce_sym<1>, rgrid, dim3(256), 0, st, n, alpha, C, work, absmax,
                         active);
      break;
    case 2:
      hipLaunchKernelGGL(k_split_reduce_sym<2>, rgrid, dim3(256), 0, st, n, alpha, C, work, absmax,
                         active);
      break;
    case 3:
      hipLaunchKernelGGL(k_split_reduce_sym<3>, rgrid, dim3(256), 0, st, n, alpha, C, work, absmax,
                         active);
      break;
    default:
      hipLaunchKernelGGL(k_split_reduce_sym<4>, rgrid, dim3(256), 0, st, n, alpha, C, work, absmax,
                         active);
  }
  return hipGetLastError();
}

// Batched symmetric products (gemm_f64_sym_batch): the Taylor chain's powers
// X0^j are formed several at a time as products of one power with up to four
// lower ones, so that one launch holds nb x 120 workgroups at N = 1800 (two per
// CU at nb = 4) without splitting K: no partial sums are written or re-read,
// and the mirror and max|C| of k_split_reduce_sym move into the tile's own
// epilogue.  Products are dealt to the XCDs one after the other (each XCD works
// on one product's tiles and shares its operand panels in L2).  rings200
// (N = 1800): 0.50 ms for four products, against 0.69 ms for four split
// products with their epilogues.
__global__ __launch_bounds__(512, 4) void k_gemm_sym_batch(
    int n, int nb, const double *__restrict__ A, const double *B0, const double *B1,
    const double *B2, const double *B3, double *C0, double *C1, double *C2, double *C3,
    unsigned long long *absmax, const int *active) {
  if (active && *active == 0) return;
  const int tiles = (n + GB_M - 1) / GB_M, nl = tiles * (tiles + 1) / 2;
  const int w = xcd_contiguous(blockIdx.x, nb * nl);
  const int b = w / nl, t = w - b * nl;
  const double *B = b == 0 ? B0 : b == 1 ? B1 : b == 2 ? B2 : B3;
  double *C = b == 0 ? C0 : b == 1 ? C1 : b == 2 ? C2 : C3;
  gemm_w8_tile<16, false, false>(t, n, n, n, 1.0, A, n, B, n, 0.0, C, n, absmax + b,
                                 GEMM_C_LOWER | GEMM_LOWER_LIST | GEMM_TILE_DIRECT | GEMM_SYM_OUT,
                                 nullptr, 1, 0, 0);
}

hipError_t gemm_f64_sym_batch(int n, int nb, const double *A, const double *const B[4],
                              double *const C[4], unsigned long long *absmax, const int *active,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (nb < 1 || nb > 4 || !absmax) return hipErrorInvalidValue;
  const int tiles = (int)cdiv(n, GB_M), nl = tiles * (tiles + 1) / 2;
  const double *b[4];
  double *c[4];
  for (int q = 0; q < 4; ++q) {
    b[q] = B[q < nb ? q : 0];
    c[q] = C[q < nb ? q : 0];
  }
  hipLaunchKernelGGL(k_gemm_sym_batch, dim3(nb * nl), dim3(512), 0, st, n, nb, A, b[0], b[1], b[2],
                     b[3], c[0], c[1], c[2], c[3], absmax, active);
  return hipGetLastError();
}

hipError_t gemm_f64_split(int M, int N, int K, double alpha, const double *A, int lda,
                          const double *B, int ldb, double *C, int ldc, double *work,
                          unsigned long long *absmax, const int *active, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (!work || !gemm_split_applies(M, N, K))
    return gemm_f64(M, N, K, alpha, A, lda, B, ldb, 0.0, C, ldc, absmax, active, st);
  const int tiles = (int)cdiv(M, GB_M) * (int)cdiv(N, GB_N);
  hipLaunchKernelGGL((k_gemm_f64w8<16, false, true>), dim3(2 * tiles), dim3(512), 0, st, M, N, K,
                     1.0, A, lda, B, ldb, 0.0, C, ldc, nullptr, active, GEMM_BLOCKED, work, 2);
  const size_t mn = (size_t)M * N;
  if (ldc == M && mn % 4 == 0 && ((uintptr_t)C & 15) == 0 && ((uintptr_t)work & 15) == 0) {
    const size_t n4 = mn / 4;
    const int grid = (int)std::min<size_t>(cdiv((long long)n4, 256), 2048);
    hipLaunchKernelGGL(k_split_reduce_flat, dim3(grid), dim3(256), 0, st, n4, alpha, C, work,
                       absmax, active);
  } else {
    hipLaunchKernelGGL(k_split_reduce, dim3(N), dim3(256), 0, st, M, alpha, C, ldc, work, absmax,
                       active);
  }
  return hipGetLastError();
}

// Small-tile variant for the latency-bound products at the bottom of the
// Cholesky recursion and the panel products of the Gauss-Jordan inverse:
// 64 x 64 tiles, 4 waves of 32 x 32, K slabs of 32.  Same operand flags.
constexpr int GS_M = 64, GS_N = 64, GS_K = 32;

template <bool TRANS_B>
__global__ __launch_bounds__(256) void k_gemm_f64s(int M, int N, int K, double alpha,
                                                  const double *__restrict__ A, int lda,
                                                  const double *__restrict__ B, int ldb,
                                                  double beta, double *__restrict__ C, int ldc,
                                                  int tri) {
  __shared__ double As[GS_K][GS_M + G_PAD];
  __shared__ double Bs[GS_K][GS_N + G_PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int tiles_m = (M + GS_M - 1) / GS_M;
  const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
  const int m0 = tm * GS_M, n0 = tn * GS_N;
  if ((tri & GEMM_C_LOWER) && n0 >= m0 + GS_M) return;
  int kbeg = 0, kend = K;
  if (tri & GEMM_A_LOWER) kend = min(kend, m0 + GS_M);
  if (tri & GEMM_B_LOWER) kbeg = n0;
  if (tri & GEMM_BT_LOWER) kend = min(kend, n0 + GS_N);
  double4_t acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
  constexpr int NL = GS_K * 64 / 256;
  double ra[NL], rb[NL];
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      const int idx = tid + 256 * r;
      const int k = idx >> 6, m = idx & 63;
      const int gm = m0 + m, gk = k0 + k;
      ra[r] = (gm < M && gk < kend) ? A[(size_t)gm + (size_t)gk * lda] : 0.0;
      if (TRANS_B) {
        const int gn = n0 + m;
        rb[r] = (gn < N && gk < kend) ? B[(size_t)gn + (size_t)gk * ldb] : 0.0;
      } else {
        const int n = idx / GS_K, kk = idx % GS_K;
        const int gn = n0 + n, gk2 = k0 + kk;
        rb[r] = (gn < N && gk2 < kend) ? B[(size_t)gk2 + (size_t)gn * ldb] : 0.0;
      }
    }
  };
  load_tiles(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += GS_K) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      const int idx = tid + 256 * r;
      As[idx >> 6][idx & 63] = ra[r];
      if (TRANS_B) Bs[idx >> 6][idx & 63] = rb[r];
      else Bs[idx % GS_K][idx / GS_K] = rb[r];
    }
    __syncthreads();
    if (k0 + GS_K < kend) load_tiles(k0 + GS_K);
#pragma unroll
    for (int ks = 0; ks < GS_K; ks += 4) {
      const int kr = ks + (lane >> 4);
      double af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i] = As[kr][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j] = Bs[kr][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;
        if (row < M && col < N) {
          double v = alpha * acc[i][j][r];
          double *cp = C + (size_t)row + (size_t)col * ldc;
          if (beta != 0.0) v = v + beta * (*cp);
          *cp = v;
        }
      }
    }
}

// C = alpha A op(B) + beta C with op(B) = B or B^T and the triangular flags
// GEMM_* (lower-triangular A / B / B^T operands, lower-only C tiles)
hipError_t gemm_f64_ex(int M, int N, int K, double alpha, const double *A, int lda, const double *B,
                       int ldb, bool trans_b, double beta, double *C, int ldc, int tri,
                       hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  const int tiles = (int)cdiv(M, GB_M) * (int)cdiv(N, GB_N);
  if (tiles < SMALL_GEMM_TILES) {
    const int stiles = (int)cdiv(M, GS_M) * (int)cdiv(N, GS_N);
    if (trans_b)
      hipLaunchKernelGGL((k_gemm_f64s<true>), dim3(stiles), dim3(256), 0, st, M, N, K, alpha, A, lda,
                         B, ldb, beta, C, ldc, tri);
    else
      hipLaunchKernelGGL((k_gemm_f64s<false>), dim3(stiles), dim3(256), 0, st, M, N, K, alpha, A,
                         lda, B, ldb, beta, C, ldc, tri);
    return hipGetLastError();
  }
  if (trans_b)
    hipLaunchKernelGGL((k_gemm_f64w8<16, true>), dim3(tiles), dim3(512), 0, st, M, N, K, alpha, A,
                       lda, B, ldb, beta, C, ldc, nullptr, nullptr, tri, nullptr, 1);
  else
    hipLaunchKernelGGL((k_gemm_f64w8<16, false>), dim3(tiles), dim3(512), 0, st, M, N, K, alpha, A,
                       lda, B, ldb, beta, C, ldc, nullptr, nullptr, tri, nullptr, 1);
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// R = I - A X with error-free extraction (after Rump, Ogita, Oishi, "Accurate
// floating-point summation part I", SIAM J. Sci. Comput. 31 (2008),
// ExtractScalar): sigma_ij = 2^(e_i + f_j + M) with 2^e_i > max_k |A_ik|,
// 2^f_j > max_k |X_kj| and 2^M >= 2n, so every product a x = A_ik X_kj has
// |a x| <= sigma / (2n).  t = fma(a, x, sigma) then lies in [sigma/2, 2 sigma],
// so q = t - sigma is exact (Sterbenz) and a multiple of 2^-53 sigma; the n
// parts q sum exactly (every partial sum is such a multiple below sigma), and
// the rest a x - q, at most 2^-53 sigma, comes from one more fma and is summed
// in plain FP64.  Five operations per term against the ten of the compensated
// dot product; the error is below 16 n^3 eps^2 max|A_i.| max|X_.j| + 2 eps |R_ij|,
// the order of the compensated form's worst case.
// ---------------------------------------------------------------------------
// mx[i] (i < n) = max_k |A_ik| as IEEE bits (zeroed first); one workgroup per
// 64 rows x one of RM_CHUNKS column chunks
constexpr int RM_CHUNKS = 16;
constexpr int R_T = 64;  // rows per residual tile
__global__ __launch_bounds__(256) void k_row_absmax(int n, const double *__restrict__ A,
                                                    unsigned long long *mx) {
  __shared__ double part[4][64];
  const int r = threadIdx.x & 63, c0 = threadIdx.x >> 6, i = blockIdx.x * 64 + r;
  const int per = (n + RM_CHUNKS - 1) / RM_CHUNKS;
  const int kb = blockIdx.y * per, ke = min(n, kb + per);
  double m = 0.0;
  if (i < n)
    for (int k = kb + c0; k < ke; k += 4) m = fmax(m, fabs(A[(size_t)i + (size_t)k * n]));
  part[c0][r] = m;
  __syncthreads();
  if (c0 == 0 && i < n) {
    m = fmax(fmax(part[0][r], part[1][r]), fmax(part[2][r], part[3][r]));
    atomicMax(&mx[i], (unsigned long long)__double_as_longlong(m));
  }
}

// mx[j] = max_k |X_kj| as IEEE bits; one wave per column
__global__ __launch_bounds__(256) void k_col_absmax(int n, const double *__restrict__ X,
                                                    unsigned long long *mx) {
  const int lane = threadIdx.x & 63, j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  double m = 0.0;
  for (int k = lane; k < n; k += 64) m = fmax(m, fabs(X[(size_t)k + (size_t)j * n]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if (lane == 0) mx[j] = (unsigned long long)__double_as_longlong(m);
}

hipError_t col_absmax(int n, const double *A, unsigned long long *mx, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_col_absmax, dim3(cdiv(n, 4)), dim3(256), 0, st, n, A, mx);
  return hipGetLastError();
}

template <int QN>
__global__ __launch_bounds__(256, QN == 2 ? 5 : 7) void k_residual_ex(
    int n, int M, const double *__restrict__ A, const double *__restrict__ X,
    double *__restrict__ R, const unsigned long long *__restrict__ amax,
    const unsigned long long *__restrict__ xmax) {
  constexpr int RC = 16 * QN;
  // k-slab depth (two barriers per slab): 32 where the registers allow
  constexpr int RX_K = QN == 1 ? 32 : 16;
  __shared__ double As[RX_K][R_T + 1];
  __shared__ double Xs[RX_K][RC + 1];
  const int tid = threadIdx.x;
  const int tr = tid & 15, tc = tid >> 4;
  const int i0 = blockIdx.x * R_T, j0 = blockIdx.y * RC;
  double hi[4][QN], lo[4][QN], sg[4][QN];
  int ea[4], ex[QN];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int gi = i0 + tr + 16 * p;
    ea[p] = 0;
    if (gi < n) (void)frexp(__longlong_as_double((long long)amax[gi]), &ea[p]);
  }
#pragma unroll
  for (int q = 0; q < QN; ++q) {
    const int gj = j0 + tc + 16 * q;
    ex[q] = 0;
    if (gj < n) (void)frexp(__longlong_as_double((long long)xmax[gj]), &ex[q]);
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      hi[p][q] = 0.0;
      lo[p][q] = 0.0;
      sg[p][q] = ldexp(1.0, ea[p] + ex[q] + M);
    }
  for (int k0 = 0; k0 < n; k0 += RX_K) {
    __syncthreads();
    for (int x = tid; x < RX_K * R_T; x += 256) {
      const int m = x & 63, k = x >> 6;
      const int gi = i0 + m, gk = k0 + k;
      As[k][m] = (gi < n && gk < n) ? A[(size_t)gi + (size_t)gk * n] : 0.0;
    }
    for (int x = tid; x < RX_K * RC; x += 256) {
      const int kk = x % RX_K, jj = x / RX_K;
      const int gj = j0 + jj, gk = k0 + kk;
      Xs[kk][jj] = (gj < n && gk < n) ? X[(size_t)gk + (size_t)gj * n] : 0.0;
    }
    __syncthreads();
    // zeros past n give q = 0 and a zero remainder
#pragma unroll
    for (int k = 0; k < RX_K; ++k) {
      double av[4], xv[QN];
#pragma unroll
      for (int p = 0; p < 4; ++p) av[p] = As[k][tr + 16 * p];
#pragma unroll
      for (int q = 0; q < QN; ++q) xv[q] = Xs[k][tc + 16 * q];
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < QN; ++q) {
          const double t = fma(av[p], xv[q], sg[p][q]);
          const double e = t - sg[p][q];  // extracted part of a x: exact
          hi[p][q] = hi[p][q] + e;         // exact
          lo[p][q] = lo[p][q] + fma(av[p], xv[q], -e);
        }
    }
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const int gi = i0 + tr + 16 * p, gj = j0 + tc + 16 * q;
      if (gi < n && gj < n)
        R[(size_t)gi + (size_t)gj * n] = (((gi == gj) ? 1.0 : 0.0) - hi[p][q]) - lo[p][q];
    }
}

// Tile choice: 64 x 32 tiles unless their grid would fill fewer than three
// rounds of the device's resident workgroup slots (occupancy of the 64 x 32
// kernel times the CU count): then the last round runs at low occupancy for a
// large share of the time, and 64 x 16 tiles balance better (N = 1800:
// 1.28 ms with 64 x 16).  tile_cols 16 or 32 forces one form.
hipError_t residual_ex(int n, const double *A, const double *X, double *R, double *scratch,
                       hipStream_t st, int tile_cols) {
  if (n <= 0) return hipSuccess;
  if (!scratch) return hipErrorInvalidValue;
  unsigned long long *amax = (unsigned long long *)scratch, *xmax = amax + n;
  int M = 0;
  while ((1ll << M) < 2ll * n) ++M;
  if (tile_cols != 16 && tile_cols != 32) {
    int dev = 0, cus = 256, per_cu = 5;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_residual_ex<2>, 256, 0) !=
            hipSuccess) {
      (void)hipGetLastError();
      cus = 256;
      per_cu = 5;
    }
    const long long wg32 = (long long)cdiv(n, R_T) * cdiv(n, 32);
    tile_cols = wg32 < 3LL * per_cu * cus ? 16 : 32;
  }
  hipError_t e = hipMemsetAsync(amax, 0, sizeof(unsigned long long) * n, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_row_absmax, dim3(cdiv(n, 64), RM_CHUNKS), dim3(256), 0, st, n, A, amax);
  hipLaunchKernelGGL(k_col_absmax, dim3(cdiv(n, 4)), dim3(256), 0, st, n, X, xmax);
  if (tile_cols == 16)
    hipLaunchKernelGGL(k_residual_ex<1>, dim3(cdiv(n, R_T), cdiv(n, 16)), dim3(256), 0, st, n, M,
                       A, X, R, amax, xmax);
  else
    hipLaunchKernelGGL(k_residual_ex<2>, dim3(cdiv(n, R_T), cdiv(n, 32)), dim3(256), 0, st, n, M,
                       A, X, R, amax, xmax);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// small element-wise helpers
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_copy_block(int m, int n, const T *src, int lds, T *dst, int ldd, T scale) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (size_t)m * n) return;
  int i = (int)(x % m), j = (int)(x / m);
  T v = src[(size_t)i + (size_t)j * lds];
  dst[(size_t)i + (size_t)j * ldd] = (scale == (T)1) ? v : scale * v;
}

hipError_t copy_block(int m, int n, const double *src, int lds, double *dst, int ldd, double scale,
                      hipStream_t st) {
  size_t tot = (size_t)m * n;
  if (!tot) return hipSuccess;
  hipLaunchKernelGGL(k_copy_block<double>, dim3(cdiv(tot, 256)), dim3(256), 0, st, m, n, src, lds,
                     dst, ldd, scale);
  return hipGetLastError();
}

__global__ void k_add_diag(int n, double *A, int lda, double v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(size_t)i + (size_t)i * lda] = A[(size_t)i + (size_t)i * lda] + v;
}

hipError_t add_diag(int n, double *A, int lda, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_add_diag, dim3(cdiv(n, 256)), dim3(256), 0, st, n, A, lda, v);
  return hipGetLastError();
}

// out = sym(X0 + D): out[r,c] = 0.5 ((X0[r,c] + D[r,c]) + (X0[c,r] + D[c,r]))
__global__ void k_sym_update(int n, const double *X0, const double *D, double *out) {
  size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (size_t)n * n) return;
  int r = (int)(x % n), c = (int)(x / n);
  size_t t = (size_t)c + (size_t)r * n;
  out[x] = 0.5 * ((X0[x] + D[x]) + (X0[t] + D[t]));
}

// out = X0 + D for symmetric X0 and D (both exact mirror images): equal, bit
// for bit, to sym_update's 0.5 ((X0 + D) + (X0 + D)^T), read coalesced.
__global__ void k_add_sym(size_t n2, const double2 *__restrict__ X0, const double2 *__restrict__ D,
                          double2 *__restrict__ out) {
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < n2;
       x += (size_t)gridDim.x * blockDim.x) {
    const double2 a = X0[x], b = D[x];
    double2 r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    out[x] = r;
  }
}

hipError_t add_symmetric(int n, const double *X0, const double *D, double *out, hipStream_t st) {
  const size_t tot = (size_t)n * n;
  if ((tot & 1) || ((uintptr_t)X0 & 15) || ((uintptr_t)D & 15) || ((uintptr_t)out & 15))
    return sym_update(n, X0, D, out, st);
  const size_t n2 = tot / 2;
  hipLaunchKernelGGL(k_add_sym, dim3((unsigned)std::min<size_t>(cdiv((long long)n2, 256), 4096)),
                     dim3(256), 0, st, n2, (const double2 *)X0, (const double2 *)D, (double2 *)out);
  return hipGetLastError();
}

hipError_t sym_update(int n, const double *X0, const double *D, double *out, hipStream_t st) {
  size_t tot = (size_t)n * n;
  hipLaunchKernelGGL(k_sym_update, dim3(cdiv(tot, 256)), dim3(256), 0, st, n, X0, D, out);
  return hipGetLastError();
}

// One Taylor step: the stopping rule (compute.cpp:146-156) and, for an accepted
// term, iZupdate X = X + k T (compute.cpp:64-71), in one launch.  Every workgroup
// evaluates the rule itself; workgroup 0 alone writes the new state.  The state a
// workgroup reads is either the one before the step or the one workgroup 0 leaves,
// and both give the same decision: accepting writes old_change = change, which
// accepts again (change > change is false); stopping leaves old_change alone or
// clears `active`.
// Control block at the start of the chain: every max|T_i| at DBL_MIN (absMax's
// start value, ScaledDenseMatrix.cpp:208), the previous change at DBL_MAX, no
// term accepted yet.  Written on the device so that no host buffer has to
// outlive the call.
__global__ void k_te_init(TeCtrl *c) {
  const int t = threadIdx.x;
  if (t < 21) c->maxbits[t] = (unsigned long long)__double_as_longlong(2.2250738585072014e-308);
  if (t < 20) c->change[t] = 0.0;
  if (t == 0) {
    c->old_change = 1.7976931348623157e308;
    c->coef = 0.0;
    c->active = 1;
    c->nterms = 0;
    c->pad[0] = c->pad[1] = 0;
  }
}

hipError_t te_init(TeCtrl *ctrl, hipStream_t st) {
  hipLaunchKernelGGL(k_te_init, dim3(1), dim3(64), 0, st, ctrl);
  return hipGetLastError();
}

// A[0, n) = 0 with 16-byte stores (a fill kernel rather than a memset, so the
// call also records as an ordinary kernel node under stream capture)
__global__ void k_fill_zero(size_t n2, double2 *A) {
  const double2 z = {0.0, 0.0};
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < n2;
       x += (size_t)gridDim.x * blockDim.x)
    A[x] = z;
}

__global__ void k_fill_zero_tail(double *A) { A[0] = 0.0; }

hipError_t fill_zero(double *A, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const size_t n2 = n / 2;
  if (((uintptr_t)A & 15) != 0) return hipErrorInvalidValue;
  if (n2 > 0) {
    const int grid = (int)std::min<size_t>(cdiv((long long)n2, 256), 8192);
    hipLaunchKernelGGL(k_fill_zero, dim3(grid), dim3(256), 0, st, n2, (double2 *)A);
  }
  if (n & 1) hipLaunchKernelGGL(k_fill_zero_tail, dim3(1), dim3(1), 0, st, A + n - 1);
  return hipGetLastError();
}

__global__ void k_te_step(TeCtrl *c, int i, double k, size_t nn, double *X, const double *T) {
  __shared__ int accept_s;
  if (threadIdx.x == 0) {
    int accept = 0;
    if (c->active) {
      const double m = __longlong_as_double((long long)c->maxbits[i]);
      const double change = fabs(k * m);
      accept = !(change < 1e-5 || change > c->old_change);
      if (blockIdx.x == 0) {
        if (accept) {
          c->old_change = change;
          c->change[c->nterms] = change;
          c->nterms = c->nterms + 1;
          c->coef = k;
        } else {
          c->active = 0;
        }
      }
    }
    accept_s = accept;
  }
  __syncthreads();
  if (!accept_s) return;
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < nn;
       x += (size_t)gridDim.x * blockDim.x)
    X[x] = X[x] + k * T[x];
}

// te_steps: the decisions of up to four terms on one thread (the count of
// accepted ones left in pad[0]), then the sum over them in term order.
__global__ void k_te_decide(TeCtrl *c, int i0, int m, double k0, double k1, double k2, double k3) {
  const double kk[4] = {k0, k1, k2, k3};
  int accepted = 0;
  for (int j = 0; j < m && c->active; ++j) {
    const double mx = __longlong_as_double((long long)c->maxbits[i0 + j]);
    const double change = fabs(kk[j] * mx);
    if (change < 1e-5 || change > c->old_change) {
      c->active = 0;
      break;
    }
    c->old_change = change;
    c->change[c->nterms] = change;
    c->nterms = c->nterms + 1;
    c->coef = kk[j];
    ++accepted;
  }
  c->pad[0] = accepted;
}

__global__ void k_te_accum(const TeCtrl *c, size_t nn, double *X, double k0, double k1, double k2,
                           double k3, const double *T0, const double *T1, const double *T2,
                           const double *T3) {
  const int a = c->pad[0];
  if (a == 0) return;
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < nn;
       x += (size_t)gridDim.x * blockDim.x) {
    double v = X[x];
    v = v + k0 * T0[x];
    if (a > 1) v = v + k1 * T1[x];
    if (a > 2) v = v + k2 * T2[x];
    if (a > 3) v = v + k3 * T3[x];
    X[x] = v;
  }
}

hipError_t te_steps(int n, TeCtrl *ctrl, int i0, int m, const double k[4], double *X,
                    const double *const T[4], hipStream_t st) {
  if (m < 1 || m > 4 || i0 < 1 || i0 + m > 20) return hipErrorInvalidValue;
  const size_t nn = (size_t)n * n;
  double kk[4];
  const double *t[4];
  for (int q = 0; q < 4; ++q) {
    kk[q] = q < m ? k[q] : 0.0;
    t[q] = T[q < m ? q : 0];
  }
  hipLaunchKernelGGL(k_te_decide, dim3(1), dim3(1), 0, st, ctrl, i0, m, kk[0], kk[1], kk[2], kk[3]);
  unsigned int grid = cdiv(nn, 256);
  if (grid > 8192) grid = 8192;
  hipLaunchKernelGGL(k_te_accum, dim3(grid), dim3(256), 0, st, ctrl, nn, X, kk[0], kk[1], kk[2],
                     kk[3], t[0], t[1], t[2], t[3]);
  return hipGetLastError();
}

hipError_t te_step(int n, TeCtrl *ctrl, int i, double k, double *X, const double *T,
                   hipStream_t st) {
  size_t nn = (size_t)n * n;
  unsigned int grid = cdiv(nn, 256);
  if (grid > 8192) grid = 8192;
  hipLaunchKernelGGL(k_te_step, dim3(grid), dim3(256), 0, st, ctrl, i, k, nn, X, T);
  return hipGetLastError();
}

__global__ void k_absmax(size_t n, const double *A, unsigned long long *out) {
  double m = 0.0;
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < n;
       x += (size_t)gridDim.x * blockDim.x)
    m = fmax(m, fabs(A[x]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

hipError_t absmax_f64(size_t n, const double *A, unsigned long long *out, hipStream_t st) {
  unsigned int grid = cdiv(n, 256);
  if (grid > 4096) grid = 4096;
  hipLaunchKernelGGL(k_absmax, dim3(grid), dim3(256), 0, st, n, A, out);
  return hipGetLastError();
}

}  // namespace unc
