// gemm_tile.h -- the 128 x 128 FP64 MFMA tile of the 8-wave GEMM as a device
// function, shared by the GEMM kernels (k_dense.hip) and the fused
// Gauss-Jordan step (k_zplus.hip).
#pragma once
#include "common.h"

namespace unc {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int GB_M = 128, GB_N = 128;
constexpr int G_PAD = 4;

// Tile t of the lower triangle of a tiles x tiles grid (tn <= tm), ordered by
// 4 x 4 blocks of tiles (blocks row by row over the lower triangle, tiles row
// by row inside a block), so that a run of consecutive t covers a compact
// region whose workgroups share A and B panels.
__device__ __forceinline__ void lower_tile_blocked(int t, int tiles, int &tm, int &tn) {
  const int nsb = (tiles + 3) / 4;
  for (int sr = 0; sr < nsb; ++sr)
    for (int sc = 0; sc <= sr; ++sc) {
      const int r0 = 4 * sr, c0 = 4 * sc;
      const int rn = min(4, tiles - r0), cn = min(4, tiles - c0);
      for (int r = 0; r < rn; ++r) {
        const int len = sr == sc ? min(r + 1, cn) : cn;  // diagonal block: tn <= tm
        if (t < len) {
          tm = r0 + r;
          tn = c0 + t;
          return;
        }
        t -= len;
      }
    }
  tm = tn = 0;  // not reached for t below the tile count
}

// Tile t of a tiles_m x tiles_n grid in 4 x 4 blocks (blocks row by row, tiles
// row by row inside a block).
__device__ __forceinline__ void full_tile_blocked(int t, int tiles_m, int tiles_n, int &tm,
                                                  int &tn) {
  const int bcols = (tiles_n + 3) / 4;
  for (int sr = 0; 4 * sr < tiles_m; ++sr) {
    const int rn = min(4, tiles_m - 4 * sr);
    const int row_tiles = rn * tiles_n;  // tiles in this band of block rows
    if (t >= row_tiles) {
      t -= row_tiles;
      continue;
    }
    for (int sc = 0; sc < bcols; ++sc) {
      const int cn = min(4, tiles_n - 4 * sc);
      if (t < rn * cn) {
        tm = 4 * sr + t / cn;
        tn = 4 * sc + t % cn;
        return;
      }
      t -= rn * cn;
    }
  }
  tm = tn = 0;  // not reached for t below the tile count
}

// Workgroup id -> item of the launch such that each of the 8 XCDs (ids are
// dealt to them round robin) takes one contiguous range of items.
__device__ __forceinline__ int xcd_contiguous(int bid, int total) {
  const int x = bid & 7, slot = bid >> 3;
  const int base = total >> 3, extra = total & 7;
  return x * base + min(x, extra) + slot;
}

// One 128 x 128 tile of C = alpha op(A) B + beta C for the workgroup's tile
// index tile_bid (512 threads, 8 waves of 64 x 32; tiles in 8-row groups so
// neighbouring workgroups share A panels in L2).  tri: triangular-operand flags
// (GEMM_*); SPLIT: the grid holds `parts` workgroups per tile, part p summing
// the p-th BK-aligned chunk of K into C (p = 0) or part1 + (p - 1) M N
// (leading dimension M), unscaled.  Rows and columns of C in [skip0, skip1)
// are neither read nor written, which lets a Gauss-Jordan step leave the
// panel's band to other workgroups of the same launch.
template <int BK, bool TRANS_B, bool SPLIT>
__device__ __forceinline__ void gemm_w8_tile(int tile_bid, int M, int N, int K, double alpha,
                                             const double *__restrict__ A, int lda,
                                             const double *__restrict__ B, int ldb, double beta,
                                             double *__restrict__ C, int ldc,
                                             unsigned long long *absmax, int tri,
                                             double *__restrict__ part1, int parts, int skip0,
                                             int skip1) {
  __shared__ double As[BK][GB_M + G_PAD];
  __shared__ double Bs[BK][GB_N + G_PAD];
  __shared__ double wmax[8];
  constexpr int T = 512;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;  // 2 x 4 waves of 64 x 32
  const int tiles_m = (M + GB_M - 1) / GB_M, tiles_n = (N + GB_N - 1) / GB_N;
  int bid = tile_bid, part = 0, tm, tn;
  if (tri & GEMM_TILE_DIRECT) {  // bid is the lower-list position
    lower_tile_blocked(bid, tiles_m, tm, tn);
  } else if (tri & GEMM_LOWER_LIST) {  // grid = parts x lower tiles, chunk-major
    const int nl = tiles_m * (tiles_m + 1) / 2;
    const int w = xcd_contiguous(bid, parts * nl);
    part = w / nl;
    lower_tile_blocked(w - part * nl, tiles_m, tm, tn);
  } else if (tri & GEMM_BLOCKED) {  // grid = parts x all tiles, chunk-major
    const int nt = tiles_m * tiles_n;
    const int w = xcd_contiguous(bid, parts * nt);
    part = w / nt;
    full_tile_blocked(w - part * nt, tiles_m, tiles_n, tm, tn);
  } else {
    if (SPLIT) {
      const int nt = tiles_m * tiles_n;
      part = bid / nt;
      bid -= part * nt;
    }
    const int group = 8;
    const int per_group = group * tiles_n;
    const int first_m = (bid / per_group) * group;
    const int gsize = min(tiles_m - first_m, group);
    tm = first_m + (bid % per_group) % gsize;
    tn = (bid % per_group) / gsize;
  }
  const int m0 = tm * GB_M, n0 = tn * GB_N;
  if ((tri & GEMM_C_LOWER) && n0 >= m0 + GB_M) return;  // tile strictly above the diagonal
  int kbeg = 0, kend = K;
  if (tri & GEMM_A_LOWER) kend = min(kend, m0 + GB_M);
  if (tri & GEMM_B_LOWER) kbeg = n0;
  if (tri & GEMM_BT_LOWER) kend = min(kend, n0 + GB_N);
  if (SPLIT) {  // no triangular operand here: K is split into `parts` BK-aligned chunks
    const int chunk = ((K + parts - 1) / parts + BK - 1) / BK * BK;
    kbeg = min(K, part * chunk);
    kend = min(K, (part + 1) * chunk);
  }
  double4_t acc[4][2];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
  constexpr int NL = BK * 128 / T;
  double ra[NL], rb[NL];
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      const int idx = tid + T * r;
      const int k = idx >> 7, m = idx & 127;
      const int gm = m0 + m, gk = k0 + k;
      ra[r] = (gm < M && gk < kend) ? A[(size_t)gm + (size_t)gk * lda] : 0.0;
      if (TRANS_B) {  // B(k, n) = Bmem[n + k ldb]: 128 contiguous n per k
        const int gn = n0 + m;
        rb[r] = (gn < N && gk < kend) ? B[(size_t)gn + (size_t)gk * ldb] : 0.0;
      } else {
        const int n = idx / BK, kk = idx % BK;
        const int gn = n0 + n, gk2 = k0 + kk;
        rb[r] = (gn < N && gk2 < kend) ? B[(size_t)gk2 + (size_t)gn * ldb] : 0.0;
      }
    }
  };
  load_tiles(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      const int idx = tid + T * r;
      As[idx >> 7][idx & 127] = ra[r];
      if (TRANS_B) Bs[idx >> 7][idx & 127] = rb[r];
      else Bs[idx % BK][idx / BK] = rb[r];
    }
    __syncthreads();
    if (k0 + BK < kend) load_tiles(k0 + BK);
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
  if (SPLIT) {
    double *dst = part ? part1 + (size_t)(part - 1) * M * N : C;
    const int ld = part ? M : ldc;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int col = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * 64 + i * 16 + (lane >> 4) + 4 * r;
          if (row < M && col < N) dst[(size_t)row + (size_t)col * ld] = acc[i][j][r];
        }
      }
    return;
  }
  double lmax = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn * 32 + j * 16 + (lane & 15);
      const bool col_ok = col < N && (col < skip0 || col >= skip1);
      if (tri & GEMM_SYM_OUT) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * 64 + i * 16 + (lane >> 4) + 4 * r;
          if (row < M && col < N && row >= col) {
            const double v = alpha * acc[i][j][r];
            C[(size_t)row + (size_t)col * ldc] = v;
            if (row != col) C[(size_t)col + (size_t)row * ldc] = v;
            lmax = fmax(lmax, fabs(v));
          }
        }
        continue;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 64 + i * 16 + (lane >> 4) + 4 * r;
        if (row < M && col_ok && (row < skip0 || row >= skip1)) {
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
      double m = wmax[0];
      for (int w = 1; w < 8; ++w) m = fmax(m, wmax[w]);
      atomicMax(absmax, (unsigned long long)__double_as_longlong(m));
    }
  }
}

}  // namespace unc
