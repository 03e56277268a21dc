// k_dense.hip -- (d) dense FP64 kernels of the Z inverse + Taylor chain, gfx950.
//
// Reference: teInverse (compute.cpp:121-166) with SDM::inv (LAPACK dgetrf +
// dgetri, ScaledDenseMatrix.cpp:232-279), SDM::operator*= (dgemm,
// ScaledDenseMatrix.cpp:314-353), absMax (:207-216) and iZupdate
// (compute.cpp:64-71).
//
// gemm_f64: C = alpha A B + beta C (column-major, no transposes) on the FP64
// matrix cores (v_mfma_f64_16x16x4_f64).  128x128 tiles, 8 waves of 64x32
// (k_gemm_f64w8; the 4-wave k_gemm_f64 with 64x64 waves is kept as a tuning
// variant), K staged through LDS 16 at a time with the next slab prefetched
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

template <int GB_K, int MINB, bool TRANS_B = false>
__global__ __launch_bounds__(G_THREADS, MINB) void k_gemm_f64(int M, int N, int K, double alpha,
                                                       const double *__restrict__ A, int lda,
                                                       const double *__restrict__ B, int ldb,
                                                       double beta, double *__restrict__ C,
                                                       int ldc, unsigned long long *absmax,
                                                       const int *active, int tri = 0) {
  if (active && *active == 0) return;
  __shared__ double As[GB_K][GB_M + G_PAD];
  __shared__ double Bs[GB_K][GB_N + G_PAD];
  __shared__ double wmax[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  // tiles: 8-wide groups of tile-rows walk the columns so neighbouring
  // workgroups share A panels in L2
  const int tiles_m = (M + GB_M - 1) / GB_M, tiles_n = (N + GB_N - 1) / GB_N;
  const int bid = blockIdx.x;
  const int group = 8;
  const int per_group = group * tiles_n;
  const int gidx = bid / per_group;
  const int first_m = gidx * group;
  const int gsize = min(tiles_m - first_m, group);
  const int tm = first_m + (bid % per_group) % gsize;
  const int tn = (bid % per_group) / gsize;
  const int m0 = tm * GB_M, n0 = tn * GB_N;
  // triangular operands: only the K range that can be non-zero is visited
  if ((tri & GEMM_C_LOWER) && n0 >= m0 + GB_M) return;  // tile strictly above the diagonal
  int kbeg = 0, kend = K;
  if (tri & GEMM_A_LOWER) kend = min(kend, m0 + GB_M);
  if (tri & GEMM_B_LOWER) kbeg = n0;
  if (tri & GEMM_BT_LOWER) kend = min(kend, n0 + GB_N);

  double4_t acc[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};

  constexpr int NL = GB_K * 128 / G_THREADS;  // elements of each tile per thread
  double ra[NL], rb[NL];
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      int idx = tid + G_THREADS * r;
      int k = idx >> 7, m = idx & 127;  // A: GB_K columns of 128 rows
      int gm = m0 + m, gk = k0 + k;
      ra[r] = (gm < M && gk < kend) ? A[(size_t)gm + (size_t)gk * lda] : 0.0;
      if (TRANS_B) {  // B(k, n) = Bmem[n + k ldb]: 128 contiguous n per k
        int gn = n0 + m;
        rb[r] = (gn < N && gk < kend) ? B[(size_t)gn + (size_t)gk * ldb] : 0.0;
      } else {
        int n = idx / GB_K, kk = idx % GB_K;  // B: 128 columns of GB_K rows
        int gn = n0 + n, gk2 = k0 + kk;
        rb[r] = (gn < N && gk2 < kend) ? B[(size_t)gk2 + (size_t)gn * ldb] : 0.0;
      }
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      int idx = tid + G_THREADS * r;
      As[idx >> 7][idx & 127] = ra[r];
      if (TRANS_B) Bs[idx >> 7][idx & 127] = rb[r];
      else Bs[idx % GB_K][idx / GB_K] = rb[r];
    }
  };

  load_tiles(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += GB_K) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (k0 + GB_K < kend) load_tiles(k0 + GB_K);
    // fragments of k-substep ks+1 are read from LDS while the 16 MFMAs of
    // substep ks run (register double buffering)
    double af[2][4], bf[2][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) af[0][i] = As[lane >> 4][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
    for (int j = 0; j < 4; ++j) bf[0][j] = Bs[lane >> 4][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
    for (int ks = 0; ks < GB_K / 4; ++ks) {
      const int cur = ks & 1, nxt = cur ^ 1;
      if (ks + 1 < GB_K / 4) {
        const int kr = 4 * (ks + 1) + (lane >> 4);
#pragma unroll
        for (int i = 0; i < 4; ++i) af[nxt][i] = As[kr][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < 4; ++j) bf[nxt][j] = Bs[kr][wn * 64 + j * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[cur][i], bf[cur][j], acc[i][j], 0, 0, 0);
    }
  }

  // epilogue: the f64 16x16x4 result has one row group per register -- lane
  // holds rows lane/16 + 4*r (r = 0..3) of column lane%16 of each 16x16 block
  double lmax = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + wn * 64 + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 64 + i * 16 + (lane >> 4) + 4 * r;
        if (row < M && col < N) {
          double v = alpha * acc[i][j][r];
          double *cp = C + (size_t)row + (size_t)col * ldc;
          if (beta != 0.0) v = v + beta * (*cp);
          *cp = v;
          lmax = fmax(lmax, fabs(v));
        }
      }
    }
  if (absmax) {
    for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off));
    if (lane == 0) wmax[wave] = lmax;
    __syncthreads();
    if (tid == 0) {
      double m = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
      atomicMax(absmax, (unsigned long long)__double_as_longlong(m));
    }
  }
}

// Pipelined variant: K slabs of 32 in two LDS buffers (A as [k][m] with a row
// stride of 144 doubles, B as [n][k] with 34, both conflict-free for the
// fragment reads), one barrier per slab, global loads two slabs ahead.
constexpr int G3_K = 32, G3_AS = 144, G3_BS = 34;
constexpr size_t G3_LDS = sizeof(double) * 2 * (G3_K * G3_AS + 128 * G3_BS);

__global__ __launch_bounds__(G_THREADS, 1) void k_gemm_f64_p(int M, int N, int K, double alpha,
                                                            const double *__restrict__ A, int lda,
                                                            const double *__restrict__ B, int ldb,
                                                            double beta, double *__restrict__ C,
                                                            int ldc, unsigned long long *absmax,
                                                            const int *active) {
  if (active && *active == 0) return;
  extern __shared__ double g3_smem[];
  double *As = g3_smem;                      // [2][G3_K][G3_AS]
  double *Bs = g3_smem + 2 * G3_K * G3_AS;   // [2][128][G3_BS]
  __shared__ double wmax[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int tiles_m = (M + GB_M - 1) / GB_M, tiles_n = (N + GB_N - 1) / GB_N;
  const int bid = blockIdx.x;
  const int group = 8;
  const int per_group = group * tiles_n;
  const int first_m = (bid / per_group) * group;
  const int gsize = min(tiles_m - first_m, group);
  const int tm = first_m + (bid % per_group) % gsize;
  const int tn = (bid % per_group) / gsize;
  const int m0 = tm * GB_M, n0 = tn * GB_N;

  double4_t acc[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};

  double ra[16], rb[16];
  // every load address is clamped into the matrix and out-of-range values are
  // replaced by zero with a select: the slab loop is straight-line code, which
  // keeps the accumulators resident in AGPRs across iterations
  const int am = min(m0 + (tid & 127), M - 1);
  const bool am_ok = m0 + (tid & 127) < M;
  const int bn0 = n0 + (tid >> 5);
  // the zero for out-of-range elements is applied with a bit mask, so the load
  // itself is unconditional (a select would let the compiler sink the load
  // into a branch and wait on it)
  auto keep = [](bool c) { return c ? ~0ull : 0ull; };
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int gk = k0 + 2 * r + (tid >> 7);
      const double va = A[(size_t)am + (size_t)min(gk, K - 1) * lda];
      ra[r] = __longlong_as_double(__double_as_longlong(va) & keep(am_ok && gk < K));
      const int gn = bn0 + 8 * r, gk2 = k0 + (tid & 31);
      const double vb = B[(size_t)min(gk2, K - 1) + (size_t)min(gn, N - 1) * ldb];
      rb[r] = __longlong_as_double(__double_as_longlong(vb) & keep(gn < N && gk2 < K));
    }
  };
  auto store = [&](int buf) {
    double *a = As + buf * G3_K * G3_AS;
    double *b = Bs + buf * 128 * G3_BS;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int idx = tid + G_THREADS * r;
      a[(idx >> 7) * G3_AS + (idx & 127)] = ra[r];
      b[(idx >> 5) * G3_BS + (idx & 31)] = rb[r];
    }
  };

  const int nk = (K + G3_K - 1) / G3_K;
  load(0);
  store(0);
  load(G3_K);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const double *a = As + cur * G3_K * G3_AS + wm * 64 + (lane & 15);
    const double *b = Bs + cur * 128 * G3_BS + (wn * 64 + (lane & 15)) * G3_BS;
#pragma unroll
    for (int ks = 0; ks < G3_K; ks += 4) {
      const int kr = ks + (lane >> 4);
      double af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = a[kr * G3_AS + i * 16];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = b[j * 16 * G3_BS + kr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    store(cur ^ 1);            // slab kt+1 (past the end: unused zeros)
    load((kt + 2) * G3_K);     // slab kt+2 (clamped addresses past the end)
    __syncthreads();
  }

  double lmax = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + wn * 64 + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 64 + i * 16 + (lane >> 4) + 4 * r;
        if (row < M && col < N) {
          double v = alpha * acc[i][j][r];
          double *cp = C + (size_t)row + (size_t)col * ldc;
          if (beta != 0.0) v = v + beta * (*cp);
          *cp = v;
          lmax = fmax(lmax, fabs(v));
        }
      }
    }
  if (absmax) {
    for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off));
    if (lane == 0) wmax[wave] = lmax;
    __syncthreads();
    if (tid == 0) {
      double m = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
      atomicMax(absmax, (unsigned long long)__double_as_longlong(m));
    }
  }
}

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

// 8-wave variant with conflict-free LDS rows (stride 144 doubles: rows shift by
// 32 banks) and an interior fast path: tiles fully inside M x N load full K
// slabs without per-element bounds checks.
template <int BK>
__global__ __launch_bounds__(512, 1) void k_gemm_f64w8p(int M, int N, int K, double alpha,
                                                       const double *__restrict__ A, int lda,
                                                       const double *__restrict__ B, int ldb,
                                                       double beta, double *__restrict__ C,
                                                       int ldc) {
  constexpr int LS = 144;
  __shared__ double As[BK][LS];
  __shared__ double Bs[BK][LS];
  constexpr int T = 512;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int tiles_m = (M + GB_M - 1) / GB_M, tiles_n = (N + GB_N - 1) / GB_N;
  const int bid = blockIdx.x;
  const int group = 8;
  const int per_group = group * tiles_n;
  const int first_m = (bid / per_group) * group;
  const int gsize = min(tiles_m - first_m, group);
  const int tm = first_m + (bid % per_group) % gsize;
  const int tn = (bid % per_group) / gsize;
  const int m0 = tm * GB_M, n0 = tn * GB_N;
  const bool interior = (m0 + GB_M <= M) && (n0 + GB_N <= N);
  double4_t acc[4][2];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
  constexpr int NL = BK * 128 / T;
  double ra[NL], rb[NL];
  auto load_tiles = [&](int k0) {
    if (interior && k0 + BK <= K) {
#pragma unroll
      for (int r = 0; r < NL; ++r) {
        const int idx = tid + T * r;
        ra[r] = A[(size_t)(m0 + (idx & 127)) + (size_t)(k0 + (idx >> 7)) * lda];
        rb[r] = B[(size_t)(k0 + idx % BK) + (size_t)(n0 + idx / BK) * ldb];
      }
    } else {
#pragma unroll
      for (int r = 0; r < NL; ++r) {
        const int idx = tid + T * r;
        const int gm = m0 + (idx & 127), gk = k0 + (idx >> 7);
        ra[r] = (gm < M && gk < K) ? A[(size_t)gm + (size_t)gk * lda] : 0.0;
        const int gn = n0 + idx / BK, gk2 = k0 + idx % BK;
        rb[r] = (gn < N && gk2 < K) ? B[(size_t)gk2 + (size_t)gn * ldb] : 0.0;
      }
    }
  };
  load_tiles(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      const int idx = tid + T * r;
      As[idx >> 7][idx & 127] = ra[r];
      Bs[idx % BK][idx / BK] = rb[r];
    }
    __syncthreads();
    if (k0 + BK < K) load_tiles(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4);
      double af[4], bf[2];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = As[kr][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j] = Bs[kr][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 64 + i * 16 + (lane >> 4) + 4 * r;
        if (row < M && col < N) {
          double v = alpha * acc[i][j][r];
          double *cp = C + (size_t)row + (size_t)col * ldc;
          if (beta != 0.0) v = v + beta * (*cp);
          *cp = v;
        }
      }
    }
}

static int g_gemm_variant = 0;

void set_gemm_variant(int v) { g_gemm_variant = v; }

hipError_t gemm_f64(int M, int N, int K, double alpha, const double *A, int lda, const double *B,
                    int ldb, double beta, double *C, int ldc, unsigned long long *absmax,
                    const int *active, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  const int tiles = (int)cdiv(M, GB_M) * (int)cdiv(N, GB_N);
  if (g_gemm_variant == 0 && !absmax && !active && tiles < SMALL_GEMM_TILES)
    return gemm_f64_ex(M, N, K, alpha, A, lda, B, ldb, false, beta, C, ldc, 0, st);
  switch (g_gemm_variant) {
    case 1:
      hipLaunchKernelGGL((k_gemm_f64<16, 2>), dim3(tiles), dim3(G_THREADS), 0, st, M, N, K, alpha,
                         A, lda, B, ldb, beta, C, ldc, absmax, active, 0);
      break;
    case 2:
      hipLaunchKernelGGL((k_gemm_f64<16, 1>), dim3(tiles), dim3(G_THREADS), 0, st, M, N, K, alpha,
                         A, lda, B, ldb, beta, C, ldc, absmax, active, 0);
      break;
    case 3: {
      static bool attr = false;
      if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)k_gemm_f64_p,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)G3_LDS);
        if (e != hipSuccess) return e;
        attr = true;
      }
      hipLaunchKernelGGL(k_gemm_f64_p, dim3(tiles), dim3(G_THREADS), G3_LDS, st, M, N, K, alpha, A,
                         lda, B, ldb, beta, C, ldc, absmax, active);
      break;
    }
    case 4:
      hipLaunchKernelGGL((k_gemm_f64w8<32>), dim3(tiles), dim3(512), 0, st, M, N, K, alpha, A, lda, B,
                         ldb, beta, C, ldc, absmax, active, 0, nullptr, 1);
      break;
    case 5:
      hipLaunchKernelGGL((k_gemm_f64<32, 1>), dim3(tiles), dim3(G_THREADS), 0, st, M, N, K, alpha,
                         A, lda, B, ldb, beta, C, ldc, absmax, active, 0);
      break;
    case 6:
      if (absmax || active) return hipErrorInvalidValue;
      hipLaunchKernelGGL((k_gemm_f64w8p<16>), dim3(tiles), dim3(512), 0, st, M, N, K, alpha, A, lda,
                         B, ldb, beta, C, ldc);
      break;
    case 7:
      if (absmax || active) return hipErrorInvalidValue;
      hipLaunchKernelGGL((k_gemm_f64w8p<32>), dim3(tiles), dim3(512), 0, st, M, N, K, alpha, A, lda,
                         B, ldb, beta, C, ldc);
      break;
    default:  // production: 8 waves, K slabs of 16
      hipLaunchKernelGGL((k_gemm_f64w8<16>), dim3(tiles), dim3(512), 0, st, M, N, K, alpha, A, lda,
                         B, ldb, beta, C, ldc, absmax, active, 0, nullptr, 1);
  }
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
// The scratch is sized for up to SYM_PARTS chunks; both choices can be
// overridden per process for tuning.
constexpr int SYM_PARTS = 4;
static int g_sym_parts = 0;  // 0: by the rule above
void set_sym_parts(int p) { g_sym_parts = p < 0 ? 0 : (p > SYM_PARTS ? SYM_PARTS : p); }
static int g_sym_list = 1;
void set_sym_list(int on) { g_sym_list = on; }
// the full split product in XCD-contiguous 4 x 4 tile blocks
static int g_split_blocked = 1;
void set_split_blocked(int on) { g_split_blocked = on; }

bool gemm_split_applies(int M, int N, int K) {
  const long long tiles = cdiv(M, GB_M) * cdiv(N, GB_N);
  return g_gemm_variant == 0 && tiles >= SMALL_GEMM_TILES && tiles <= SPLIT_MAX_TILES &&
         K >= SPLIT_MIN_K;
}

size_t gemm_sym_split_work_doubles(int n) {
  return gemm_split_applies(n, n, n) ? (size_t)(SYM_PARTS - 1) * n * n : 0;
}

hipError_t gemm_f64_sym(int n, double alpha, const double *A, const double *B, double *C,
                        double *work, size_t work_doubles, unsigned long long *absmax,
                        const int *active, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (!work || work_doubles < gemm_sym_split_work_doubles(n) || !gemm_split_applies(n, n, n))
    return gemm_f64(n, n, n, alpha, A, n, B, n, 0.0, C, n, absmax, active, st);
  const int tiles = (int)cdiv(n, GB_M) * (int)cdiv(n, GB_N);
  const int tm = (int)cdiv(n, GB_M), lower = tm * (tm + 1) / 2;
  const int parts = g_sym_parts ? g_sym_parts : (SYM_PARTS * lower <= 512 ? SYM_PARTS : 3);
  if (g_sym_list)
    hipLaunchKernelGGL((k_gemm_f64w8<16, false, true>), dim3(parts * lower), dim3(512), 0, st, n,
                       n, n, 1.0, A, n, B, n, 0.0, C, n, nullptr, active,
                       GEMM_C_LOWER | GEMM_LOWER_LIST, work, parts);
  else
    hipLaunchKernelGGL((k_gemm_f64w8<16, false, true>), dim3(parts * tiles), dim3(512), 0, st,
                       n, n, n, 1.0, A, n, B, n, 0.0, C, n, nullptr, active, GEMM_C_LOWER, work,
                       parts);
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

hipError_t gemm_f64_split(int M, int N, int K, double alpha, const double *A, int lda,
                          const double *B, int ldb, double *C, int ldc, double *work,
                          unsigned long long *absmax, const int *active, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (!work || !gemm_split_applies(M, N, K))
    return gemm_f64(M, N, K, alpha, A, lda, B, ldb, 0.0, C, ldc, absmax, active, st);
  const int tiles = (int)cdiv(M, GB_M) * (int)cdiv(N, GB_N);
  hipLaunchKernelGGL((k_gemm_f64w8<16, false, true>), dim3(2 * tiles), dim3(512), 0, st, M, N, K,
                     1.0, A, lda, B, ldb, 0.0, C, ldc, nullptr, active, g_split_blocked ? GEMM_BLOCKED : 0,
                     work, 2);
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
// R = I - A X with compensated dot products (Ogita-Rump-Oishi Dot2): the
// product error comes from fma, the running sum uses TwoSum, the errors are
// summed separately and added once at the end.  Needs -ffp-contract=off.
// ---------------------------------------------------------------------------
constexpr int R_T = 64, R_K = 16;

// Tile 64 x 16*QN per 256 threads; thread (tr, tc) owns rows tr + 16p (p < 4)
// and columns tc + 16q (q < QN).  QN = 2 halves the work per wave against
// QN = 4, so the waves spread evenly over the SIMDs (N = 1800: 6.6 waves per
// SIMD rather than 3.3).
template <int QN>
__global__ __launch_bounds__(256, QN == 4 ? 4 : QN == 3 ? 5 : QN == 2 ? 6 : 8) void k_residual_dd(
    int n, const double *__restrict__ A, const double *__restrict__ X, double *__restrict__ R) {
  constexpr int RC = 16 * QN;
  __shared__ double As[R_K][R_T + 1];
  __shared__ double Xs[R_K][RC + 1];
  const int tid = threadIdx.x;
  const int tr = tid & 15, tc = tid >> 4;
  const int i0 = blockIdx.x * R_T, j0 = blockIdx.y * RC;
  double s[4][QN], c[4][QN];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const int gi = i0 + tr + 16 * p, gj = j0 + tc + 16 * q;
      s[p][q] = (gi == gj) ? 1.0 : 0.0;
      c[p][q] = 0.0;
    }
  for (int k0 = 0; k0 < n; k0 += R_K) {
    __syncthreads();
    for (int x = tid; x < R_K * R_T; x += 256) {
      const int m = x & 63, k = x >> 6;
      const int gi = i0 + m, gk = k0 + k;
      As[k][m] = (gi < n && gk < n) ? A[(size_t)gi + (size_t)gk * n] : 0.0;
    }
    for (int x = tid; x < R_K * RC; x += 256) {
      const int kk = x & 15, jj = x >> 4;
      const int gj = j0 + jj, gk = k0 + kk;
      Xs[kk][jj] = (gj < n && gk < n) ? X[(size_t)gk + (size_t)gj * n] : 0.0;
    }
    __syncthreads();
    // k beyond n reads zeros from LDS: their products are exact zeros and leave
    // (s, c) unchanged, so the chunk loop has a constant trip count
#pragma unroll
    for (int k = 0; k < R_K; ++k) {
      double av[4], xv[QN];
#pragma unroll
      for (int p = 0; p < 4; ++p) av[p] = As[k][tr + 16 * p];
#pragma unroll
      for (int q = 0; q < QN; ++q) xv[q] = Xs[k][tc + 16 * q];
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < QN; ++q) {
          const double P = av[p] * xv[q];
          const double E = fma(av[p], xv[q], -P);
          const double pm = -P;
          const double t = s[p][q] + pm;
          const double bb = t - s[p][q];
          const double err = (s[p][q] - (t - bb)) + (pm - bb);
          s[p][q] = t;
          c[p][q] = c[p][q] + (err - E);
        }
    }
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const int gi = i0 + tr + 16 * p, gj = j0 + tc + 16 * q;
      if (gi < n && gj < n) R[(size_t)gi + (size_t)gj * n] = s[p][q] + c[p][q];
    }
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

template <int QN>
__global__ __launch_bounds__(256, QN == 2 ? 5 : 8) void k_residual_ex(
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

static int g_residual_variant = 0;
void set_residual_variant(int v) { g_residual_variant = v; }

hipError_t residual_dd(int n, const double *A, const double *X, double *R, double *scratch,
                       hipStream_t st) {
  // 64 x 32 tiles at six workgroups per CU, unless they would need fewer than
  // three rounds of the 1536 resident slots: then the last round runs at low
  // occupancy for a large share of the time, and 64 x 16 tiles (eight per CU)
  // balance better (N = 1800: 2.14 ms against 2.47 ms for the compensated form).
  const long long wg32 = (long long)cdiv(n, R_T) * cdiv(n, 32);
  if (scratch && (g_residual_variant == 0 || g_residual_variant >= 6)) {  // error-free extraction
    unsigned long long *amax = (unsigned long long *)scratch, *xmax = amax + n;
    int M = 0;
    while ((1ll << M) < 2ll * n) ++M;
    hipError_t e = hipMemsetAsync(amax, 0, sizeof(unsigned long long) * n, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_row_absmax, dim3(cdiv(n, 64), RM_CHUNKS), dim3(256), 0, st, n, A, amax);
    hipLaunchKernelGGL(k_col_absmax, dim3(cdiv(n, 4)), dim3(256), 0, st, n, X, xmax);
    if (g_residual_variant == 7 || (g_residual_variant == 0 && wg32 < 3 * 1536))
      hipLaunchKernelGGL(k_residual_ex<1>, dim3(cdiv(n, R_T), cdiv(n, 16)), dim3(256), 0, st, n,
                         M, A, X, R, amax, xmax);
    else
      hipLaunchKernelGGL(k_residual_ex<2>, dim3(cdiv(n, R_T), cdiv(n, 32)), dim3(256), 0, st, n,
                         M, A, X, R, amax, xmax);
    return hipGetLastError();
  }
  if ((g_residual_variant == 0 || g_residual_variant == 5) && wg32 < 3 * 1536) {
    hipLaunchKernelGGL(k_residual_dd<1>, dim3(cdiv(n, R_T), cdiv(n, 16)), dim3(256), 0, st, n, A,
                       X, R);
    return hipGetLastError();
  }
  if (g_residual_variant == 1)
    hipLaunchKernelGGL(k_residual_dd<4>, dim3(cdiv(n, R_T), cdiv(n, 64)), dim3(256), 0, st, n, A,
                       X, R);
  else if (g_residual_variant == 2)
    hipLaunchKernelGGL(k_residual_dd<3>, dim3(cdiv(n, R_T), cdiv(n, 48)), dim3(256), 0, st, n, A,
                       X, R);
  else if (g_residual_variant == 3)
    hipLaunchKernelGGL(k_residual_dd<1>, dim3(cdiv(n, R_T), cdiv(n, 16)), dim3(256), 0, st, n, A,
                       X, R);
  else  // 0 and 5 (large n), 4: 64 x 32 tiles
    hipLaunchKernelGGL(k_residual_dd<2>, dim3(cdiv(n, R_T), cdiv(n, 32)), dim3(256), 0, st, n, A,
                       X, R);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// in-LDS Gauss-Jordan inverse of one b x b diagonal block (b <= 64), no
// pivoting (the matrix is symmetric positive definite)
// ---------------------------------------------------------------------------
constexpr int GJ_B = 64;

// 256 threads as a 16 x 16 grid; thread (ti, tj) keeps rows ti + 16u and
// columns tj + 16v (u, v < 4) of the block in registers.  Each pivot step
// publishes the pivot row and column through LDS (two barriers) and updates the
// 16 register elements.
template <typename T>
__global__ __launch_bounds__(256) void k_gj_small(int b, const T *W, int ldw, T *P,
                                                  unsigned int *err) {
  __shared__ T colp[GJ_B], rowp[GJ_B];
  const int tid = threadIdx.x, ti = tid & 15, tj = tid >> 4;
  T m[4][4];  // m[u][v] = element (ti + 16u, tj + 16v)
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = ti + 16 * u, j = tj + 16 * v;
      m[u][v] = (i < b && j < b) ? W[(size_t)i + (size_t)j * ldw] : (T)(i == j ? 1 : 0);
    }
  bool bad = false;
  for (int p = 0; p < b; ++p) {
    const int pu = p >> 4, pl = p & 15;
    // publish column p (owned by tj == pl) and row p (owned by ti == pl)
    if (tj == pl) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        T x = 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) x = (v == pu) ? m[u][v] : x;
        colp[ti + 16 * u] = x;
      }
    }
    if (ti == pl) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        T x = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) x = (u == pu) ? m[u][v] : x;
        rowp[tj + 16 * v] = x;
      }
    }
    __syncthreads();
    const T piv = colp[p];
    if (!(piv > (T)0)) bad = true;
    const T ip = (T)1 / piv;
    T ci[4], rj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) ci[u] = colp[ti + 16 * u];
#pragma unroll
    for (int v = 0; v < 4; ++v) rj[v] = rowp[tj + 16 * v] * ip;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = ti + 16 * u;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = tj + 16 * v;
        T x;
        if (i == p && j == p) x = ip;
        else if (i == p) x = rj[v];
        else if (j == p) x = -ci[u] * ip;
        else x = m[u][v] - ci[u] * rj[v];
        m[u][v] = x;
      }
    }
    __syncthreads();
  }
  if (bad && tid == 0) atomicOr(err, ERRBIT_GJ_PIVOT);
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = ti + 16 * u, j = tj + 16 * v;
      if (i < b && j < b) P[(size_t)i + (size_t)j * b] = m[u][v];
    }
}

hipError_t gj_small_inverse(int b, const double *W, int ldw, double *P, unsigned int *err,
                            hipStream_t st) {
  if (b > GJ_B) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gj_small<double>, dim3(1), dim3(256), 0, st, b, W, ldw, P, err);
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
