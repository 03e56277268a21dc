// k_zplus.hip -- kernels of the gauge-fixed Z^+ path (gfx950).
//
// Z_s has the 7-dimensional similarity gauge as its (numerical) null space:
// fixPts drops the fixed points' rows together with their columns
// (compute.cpp:337-364), so nothing pins the gauge.  The reference regularises
// with lambda I and runs its Taylor scheme (compute.cpp:121-166); this path
// computes the gauge-fixed pseudo-inverse Z^+ = Z_g^-1 - N N^T with
// Z_g = Z_s + N N^T and N an orthonormal basis of the 7 smallest eigenvectors,
// and only the parts the outputs need: the cameras' own 9x9 blocks.
//
//   1. Z~ = Z_s + N0 N0^T for a fixed pseudo-random orthonormal N0 (SPD for any
//      N0 that is not orthogonal to the null space);
//   2. Z~ = L L^T and X = L^-1 by a recursive blocked Cholesky whose leaves are
//      64x64 register-resident factorisations and whose inner nodes are FP64
//      MFMA GEMMs (pipeline.hip);
//   3. N = orth(Z~^-1 N0) (exact when the null space is exact), then Newton
//      steps N <- orth(N - P X^T X P (Z N - N N^T Z N)) with P = I - N N^T;
//   4. Z_g^-1 = Z~^-1 - M U S U^T M  (Woodbury, U = [N N0], C = diag(I, -I),
//      S = (C + U^T M U)^-1, M = X^T X), so camera block a of Z^+ is
//      X_a^T X_a - (MU)_a S (MU)_a^T - N_a N_a^T with X_a the a-th block column.
//
// Skinny products with 7 or 14 right-hand columns stream X or Z once from HBM.
#include "common.h"

#include <utility>
#include "gemm_tile.h"

namespace unc {

// Leaves are 64 x 64 blocks (one workgroup each).
constexpr int CL_B = 64;

// ---------------------------------------------------------------------------
// Panelled leaf: 256 threads, the block in LDS, 8-column panels, one
// barrier per step.  Step s:
//   wave 0 (lane = row) applies panel s - 1, still in its registers, to panel
//   s's 8 columns and factors them, right-looking inside the panel (each
//   finished column updates the later ones at once, so the dependent chain of
//   a column is one FMA, one broadcast and the reciprocal square root; the
//   pivot row's values are broadcast with readlane);
//   wave 1 (lane = column) does the same for row panel s - 1 of X = L^-1,
//   one panel behind: it needs L's column panels up to s - 1 only;
//   waves 2-3 apply panel s - 1's rank-8 update to the trailing columns beyond
//   panel s and X's row panel s - 2's update to X's rows beyond panel s - 1.
// Nine steps; every element receives the same updates in the same order as
// the plain column-by-column factorisation and forward substitution.
// The loops over panels stay rolled, so the code is small (no instruction-fetch
// stalls); every register array is indexed by unrolled constants.
// ---------------------------------------------------------------------------
constexpr int CP_W = 8;
constexpr int CP_THREADS = 256;

__device__ __forceinline__ double readlane_d(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Wave 0 (lane = row i): factor the 8 columns [p0, p0 + 8) held in r (all
// earlier panels applied) and store them to S; rows above the diagonal become
// zero.  True when a pivot was not positive.
__device__ __forceinline__ bool chol_panel8(double (*S)[CL_B + 1], int p0, int lane,
                                            double (&r)[CP_W]) {
  const int i = lane;
  bool bad = false;
#pragma unroll
  for (int jj = 0; jj < CP_W; ++jj) {
    const int j = p0 + jj;
    const double piv = readlane_d(r[jj], j);
    if (!(piv > 0.0)) bad = true;
    // 1 / L_jj from the refined reciprocal square root; it is kept in the
    // padding column for the triangular solve
    const double rd = rsqrt_normal(piv);
    const double d = piv * rd;
    r[jj] = i > j ? r[jj] * rd : (i == j ? d : 0.0);
    if (i == j) S[j][CL_B] = rd;
#pragma unroll
    for (int mm = jj + 1; mm < CP_W; ++mm) r[mm] = fma(-r[jj], readlane_d(r[jj], p0 + mm), r[mm]);
  }
#pragma unroll
  for (int jj = 0; jj < CP_W; ++jj) S[i][p0 + jj] = r[jj];
  return bad;
}

// Wave 1 (lane = column c): rows [p0, p0 + 8) of X = L^-1 from their updated
// right-hand sides in x and L's diagonal block; stored to Xs.
__device__ __forceinline__ void tri_panel8(double (*S)[CL_B + 1], double (*Xs)[CL_B + 1], int p0,
                                           int lane, double (&x)[CP_W]) {
#pragma unroll
  for (int rr = 0; rr < CP_W; ++rr) {
    double v = x[rr];
#pragma unroll
    for (int kk = 0; kk < rr; ++kk) v = fma(-S[p0 + rr][p0 + kk], x[kk], v);
    x[rr] = v * S[p0 + rr][CL_B];
  }
#pragma unroll
  for (int rr = 0; rr < CP_W; ++rr) Xs[p0 + rr][lane] = x[rr];
}

// Waves 2-3 (W = 0, 1) of the panelled leaf at step SS: the bulk rank-8 updates
// on the FP64 matrix cores, over 16 x 16 tiles (v_mfma_f64_16x16x4: lane l holds
// A's row l % 16 and B's column l % 16 at k = l / 16; accumulator r is the
// tile's row l / 16 + 4 r, column l % 16), the tile's old values as the
// accumulator and only the elements of the update's region written back:
//   Cholesky (SS < 8, panel p0 = 8 (SS - 1)): S[i][k] -= sum_jj S[i][p0 + jj]
//   S[k][p0 + jj] for p0 + 16 <= k <= i (the trailing triangle beyond panel SS);
//   inverse (SS >= 2, X's row panel xp = 8 (SS - 2)): Xs[r][c] -= sum_kk
//   S[r][xp + kk] Xs[xp + kk][c] for rows r >= xp + 16 (only columns c < xp + 8
//   change: X's rows of panel xp are zero beyond xp + 7).
// The tiles of both form one list dealt alternately to the two waves.  The
// list is fixed by SS, so each (SS, W) is its own code: tile positions are
// constants and a wave loads every operand of its tiles before its first MFMA.
struct BulkPlan {
  int nc, ni, ct0, cst, xt0, xnj, xst, p0, xp;
  constexpr BulkPlan(int ss)
      : nc(0), ni(0), ct0(0), cst(8 * ss + 8), xt0(0), xnj(0), xst(8 * ss), p0(8 * (ss - 1)),
        xp(8 * (ss - 2)) {
    ct0 = cst >> 4;
    const int cnt = (ss < CL_B / CP_W && cst < CL_B) ? 4 - ct0 : 0;
    nc = cnt * (cnt + 1) / 2;
    xt0 = xst >> 4;
    xnj = (xp >= 0 && xst < CL_B) ? ((xst - CP_W - 1) >> 4) + 1 : 0;
    ni = (4 - xt0) * xnj;
  }
  // tile t of the list: Cholesky tiles (lower, row by row from (ct0, ct0))
  // first, then the inverse's (rows from xt0, columns 0 .. xnj - 1)
  constexpr int row(int t) const {
    if (t >= nc) return 16 * (xt0 + (t - nc) / xnj);
    int ti = ct0, rem = t;
    while (rem > ti - ct0) rem -= ti++ - ct0 + 1;
    return 16 * ti;
  }
  constexpr int col(int t) const {
    if (t >= nc) return 16 * ((t - nc) % xnj);
    int ti = ct0, rem = t;
    while (rem > ti - ct0) rem -= ti++ - ct0 + 1;
    return 16 * (ct0 + rem);
  }
};

template <int SS, int W, int U>
struct BulkTile {
  static constexpr BulkPlan P{SS};
  static constexpr int t = W + 2 * U;
  static constexpr bool chol = t < P.nc;
  static constexpr int r0 = P.row(t), c0 = P.col(t), k0 = chol ? P.p0 : P.xp;
};

template <int SS, int W, int U>
__device__ __forceinline__ void bulk_load(double (*S)[CL_B + 1], double (*Xs)[CL_B + 1], int kq,
                                          int lc, double4_t &acc, double (&av)[2],
                                          double (&bv)[2]) {
  using T = BulkTile<SS, W, U>;
#pragma unroll
  for (int r = 0; r < 4; ++r)
    acc[r] = T::chol ? S[T::r0 + kq + 4 * r][T::c0 + lc] : Xs[T::r0 + kq + 4 * r][T::c0 + lc];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int k = T::k0 + 4 * h + kq;
    av[h] = -S[T::r0 + lc][k];
    bv[h] = T::chol ? S[T::c0 + lc][k] : Xs[k][T::c0 + lc];
  }
}

template <int SS, int W, int U>
__device__ __forceinline__ void bulk_store(double (*S)[CL_B + 1], double (*Xs)[CL_B + 1], int kq,
                                           int lc, const double4_t &acc) {
  using T = BulkTile<SS, W, U>;
  const int c = T::c0 + lc;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = T::r0 + kq + 4 * r;
    if (T::chol) {
      if (c >= T::P.cst && c <= i) S[i][c] = acc[r];
    } else if (i >= T::P.xst) {
      Xs[i][c] = acc[r];
    }
  }
}

template <int SS, int W, int... U>
__device__ __forceinline__ void bulk_tiles(double (*S)[CL_B + 1], double (*Xs)[CL_B + 1], int lane,
                                           std::integer_sequence<int, U...>) {
  const int kq = lane >> 4, lc = lane & 15;
  double4_t acc[sizeof...(U) + 1];
  double av[sizeof...(U) + 1][2], bv[sizeof...(U) + 1][2];
  (bulk_load<SS, W, U>(S, Xs, kq, lc, acc[U], av[U], bv[U]), ...);
  ((acc[U] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[U][0], bv[U][0], acc[U], 0, 0, 0)), ...);
  ((acc[U] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[U][1], bv[U][1], acc[U], 0, 0, 0)), ...);
  (bulk_store<SS, W, U>(S, Xs, kq, lc, acc[U]), ...);
}

template <int SS>
__device__ __forceinline__ void bulk_step(double (*S)[CL_B + 1], double (*Xs)[CL_B + 1], int widx,
                                          int lane) {
  constexpr BulkPlan P{SS};
  constexpr int NT = P.nc + P.ni;
  if (widx == 0)
    bulk_tiles<SS, 0>(S, Xs, lane, std::make_integer_sequence<int, (NT + 1) / 2>{});
  else
    bulk_tiles<SS, 1>(S, Xs, lane, std::make_integer_sequence<int, NT / 2>{});
}

// S: the SPD block (lower triangle valid, upper triangle zero) -> L with zeros
// above the diagonal; Xs -> L^-1 with zeros above.  Returns true in wave 0 when
// a pivot was not positive.  The first 256 threads work (act); in a larger
// workgroup the others pass only the barriers.
__device__ __forceinline__ bool chol_inv_lds(double (*S)[CL_B + 1], double (*Xs)[CL_B + 1],
                                             bool act = true) {
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);  // uniform: scalar tile logic
  constexpr int NP = CL_B / CP_W;  // panels
  bool bad = false;
  double lp[CP_W] = {};  // wave 0: L's panel s - 1 in this lane's row
  double xr[CP_W] = {};  // wave 1: X's row panel s - 2 in this lane's column
  for (int s = 0; s <= NP; ++s) {
    const int q0 = s * CP_W, p0 = q0 - CP_W;  // Cholesky: panel and its predecessor
    const int xq = p0, xp = p0 - CP_W;        // X: row panel and its predecessor
    if (wave == 0) {
      if (s < NP) {
        double r[CP_W];
#pragma unroll
        for (int jj = 0; jj < CP_W; ++jj) r[jj] = S[lane][q0 + jj];
        if (s >= 1) {
#pragma unroll
          for (int jj = 0; jj < CP_W; ++jj)
#pragma unroll
            for (int kk = 0; kk < CP_W; ++kk) r[jj] = fma(-lp[kk], S[q0 + jj][p0 + kk], r[jj]);
        }
        bad = chol_panel8(S, q0, lane, r) || bad;
#pragma unroll
        for (int jj = 0; jj < CP_W; ++jj) lp[jj] = r[jj];
      }
    } else if (act) {
      if (s == 0) {  // Xs = I meanwhile
        for (int x = t - 64; x < CL_B * CL_B; x += CP_THREADS - 64) {
          const int r = x >> 6, c = x & 63;
          Xs[r][c] = r == c ? 1.0 : 0.0;
        }
      } else if (wave == 1) {
        double x[CP_W];
#pragma unroll
        for (int rr = 0; rr < CP_W; ++rr) x[rr] = Xs[xq + rr][lane];
        if (xp >= 0) {
#pragma unroll
          for (int rr = 0; rr < CP_W; ++rr)
#pragma unroll
            for (int kk = 0; kk < CP_W; ++kk) x[rr] = fma(-S[xq + rr][xp + kk], xr[kk], x[rr]);
        }
        tri_panel8(S, Xs, xq, lane, x);
#pragma unroll
        for (int rr = 0; rr < CP_W; ++rr) xr[rr] = x[rr];
      } else {
        switch (s) {
          case 1: bulk_step<1>(S, Xs, wave - 2, lane); break;
          case 2: bulk_step<2>(S, Xs, wave - 2, lane); break;
          case 3: bulk_step<3>(S, Xs, wave - 2, lane); break;
          case 4: bulk_step<4>(S, Xs, wave - 2, lane); break;
          case 5: bulk_step<5>(S, Xs, wave - 2, lane); break;
          case 6: bulk_step<6>(S, Xs, wave - 2, lane); break;
          case 7: bulk_step<7>(S, Xs, wave - 2, lane); break;
          default: break;  // step 8: nothing left beyond the last panels
        }
      }
    }
    __syncthreads();
  }
  return bad;
}

// Dst[i][k] = f(i, k) over the 64 x 64 tile, (i, k) = (x & 63, x >> 6) for
// x = tid + 256 r: each thread issues its 16 loads before the first LDS store,
// so the global latency is paid once, not 16 times.  Only the first 256
// threads take part (act).
constexpr int STAGE_R = CL_B * CL_B / CP_THREADS;
template <typename F>
__device__ __forceinline__ void stage64(double (*Dst)[CL_B + 1], F f, bool act = true) {
  if (!act) return;
  double v[STAGE_R];
#pragma unroll
  for (int r = 0; r < STAGE_R; ++r) {
    const int x = threadIdx.x + CP_THREADS * r;
    v[r] = f(x & 63, x >> 6);
  }
#pragma unroll
  for (int r = 0; r < STAGE_R; ++r) {
    const int x = threadIdx.x + CP_THREADS * r;
    Dst[x & 63][x >> 6] = v[r];
  }
}

// b x b block of A (column-major, lower triangle read) -> S, identity padding
__device__ __forceinline__ void load_spd64(int b, const double *A, int lda,
                                           double (*S)[CL_B + 1]) {
  stage64(S, [&](int i, int k) {
    return (i < b && k < b) ? (k <= i ? A[(size_t)i + (size_t)k * lda] : 0.0)
                            : (i == k ? 1.0 : 0.0);
  });
  __syncthreads();
}

__global__ __launch_bounds__(CP_THREADS) void k_chol_inv64p(int b, double *A, int lda,
                                                            double *Xo, int ldx,
                                                            unsigned int *err) {
  __shared__ double S[CL_B][CL_B + 1];
  __shared__ double Xs[CL_B][CL_B + 1];
  load_spd64(b, A, lda, S);
  const bool bad = chol_inv_lds(S, Xs);
  for (int x = threadIdx.x; x < CL_B * CL_B; x += CP_THREADS) {
    const int i = x & 63, k = x >> 6;
    if (i < b && k < b) {
      A[(size_t)i + (size_t)k * lda] = S[i][k];
      Xo[(size_t)i + (size_t)k * ldx] = Xs[i][k];
    }
  }
  if (bad && threadIdx.x == 0) atomicOr(err, ERRBIT_GJ_PIVOT);
}

// 64 x 64 product on the FP64 matrix cores: wave w of the first four computes
// the 32 x 32 quadrant (rows 32 (w & 1), columns 32 (w >> 1)) of
// C = A[0:64][0:kk] B[0:kk][0:64] as 2 x 2 tiles of v_mfma_f64_16x16x4, k
// ascending in steps of 4; a_at(row, k) and b_at(k, col) read the operands
// (zero beyond kk up to the next multiple of 4).
template <typename AF, typename BF>
__device__ __forceinline__ void mfma_quadrant64(int kk, AF a_at, BF b_at, double4_t acc[2][2]) {
  const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) & 3;
  const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < kk; k0 += 4) {
    const int k = k0 + (lane >> 4);
    double av[2], bv[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) av[i] = a_at(wr + 16 * i + (lane & 15), k);
#pragma unroll
    for (int j = 0; j < 2; ++j) bv[j] = b_at(k, wc + 16 * j + (lane & 15));
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
  }
}

// f(row, col, value, q) for the 16 entries q of the quadrant a thread holds
template <typename F>
__device__ __forceinline__ void quadrant64_for_each(const double4_t acc[2][2], F f) {
  const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) & 3;
  const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        f(wr + 16 * i + (lane >> 4) + 4 * r, wc + 16 * j + (lane & 15), acc[i][j][r],
          8 * i + 4 * j + r);
}

// P = X^T X (X = L^-1 in Xs) -> P (b x b, leading dimension b) on the matrix
// cores, first four waves.  The entries on and below the diagonal are stored
// at both (i, j) and (j, i), so P is exactly symmetric.
__device__ __forceinline__ void store_xtx(int b, double (*Xs)[CL_B + 1], double *P) {
  const int wave = (threadIdx.x >> 6) & 3;
  if (wave == 2) return;  // the quadrant above the diagonal
  double4_t acc[2][2];
  mfma_quadrant64(CL_B, [&](int row, int k) { return Xs[k][row]; },
                  [&](int k, int col) { return Xs[k][col]; }, acc);
  quadrant64_for_each(acc, [&](int i, int j, double v, int) {
    if (i < b && j < b && i >= j) {
      P[(size_t)i + (size_t)j * b] = v;
      P[(size_t)j + (size_t)i * b] = v;
    }
  });
}

// A recursion node of n rows, 64 < n <= 128, in one workgroup: halves of 64
// and n2 = n - 64 rows, the four products on the FP64 matrix cores:
//   (L11, X11) = leaf(A11);  L21 = A21 X11^T;  W = L21 X11;
//   (L22, X22) = leaf(A22 - L21 L21^T);  X21 = -X22 W.
// Two 64 x 64 blocks of LDS only (S: A11 -> L11, then A21 -> L21, then
// A22 - P -> L22, then W; Xs: X11, then X22); W and P wait in registers.  At
// 66.5 KB the node fits on a CU beside one workgroup of the large products that
// may run on the other stream, so it does not wait for a CU to drain.
// The same results as the recursion's two leaves and four products (to
// rounding), in one launch instead of seven.
__global__ __launch_bounds__(CP_THREADS) void k_chol_inv128(int n, double *A, int lda, double *Xo,
                                                            int ldx, unsigned int *err) {
  __shared__ double S[CL_B][CL_B + 1];
  __shared__ double Xs[CL_B][CL_B + 1];
  const int n2 = n - CL_B;
  double *A21 = A + CL_B, *A22 = A + CL_B + (size_t)CL_B * lda;
  double *X21 = Xo + CL_B, *X22 = Xo + CL_B + (size_t)CL_B * ldx;
  // (L11, X11)
  load_spd64(CL_B, A, lda, S);
  bool bad = chol_inv_lds(S, Xs);
  for (int x = threadIdx.x; x < CL_B * CL_B; x += CP_THREADS) {
    const int i = x & 63, k = x >> 6;
    A[(size_t)i + (size_t)k * lda] = S[i][k];
    Xo[(size_t)i + (size_t)k * ldx] = Xs[i][k];
  }
  __syncthreads();  // S (L11) read by the stores above
  stage64(S, [&](int i, int k) { return i < n2 ? A21[(size_t)i + (size_t)k * lda] : 0.0; });
  __syncthreads();
  // L21 = A21 X11^T (X11 lower), over A21 in S
  double4_t acc[2][2];
  mfma_quadrant64(CL_B, [&](int row, int k) { return S[row][k]; },
                  [&](int k, int col) { return Xs[col][k]; }, acc);
  __syncthreads();
  quadrant64_for_each(acc, [&](int i, int j, double v, int) { S[i][j] = v; });
  __syncthreads();
  for (int x = threadIdx.x; x < CL_B * CL_B; x += CP_THREADS) {
    const int i = x & 63, k = x >> 6;
    if (i < n2) A21[(size_t)i + (size_t)k * lda] = S[i][k];
  }
  // W = L21 X11 and P = L21 L21^T, both kept in registers
  double4_t wacc[2][2], pacc[2][2];
  mfma_quadrant64(CL_B, [&](int row, int k) { return S[row][k]; },
                  [&](int k, int col) { return Xs[k][col]; }, wacc);
  mfma_quadrant64(CL_B, [&](int row, int k) { return S[row][k]; },
                  [&](int k, int col) { return S[col][k]; }, pacc);
  __syncthreads();  // L21 in S read by the products and the stores
  load_spd64(n2, A22, lda, S);
  quadrant64_for_each(pacc, [&](int i, int j, double v, int) {
    if (i < n2 && j <= i) S[i][j] -= v;
  });
  __syncthreads();
  // (L22, X22); Xs held X11, no longer needed
  bad = chol_inv_lds(S, Xs) || bad;
  for (int x = threadIdx.x; x < CL_B * CL_B; x += CP_THREADS) {
    const int i = x & 63, k = x >> 6;
    if (i < n2 && k < n2) {
      A22[(size_t)i + (size_t)k * lda] = S[i][k];
      X22[(size_t)i + (size_t)k * ldx] = Xs[i][k];
    }
  }
  __syncthreads();  // S (L22) read by the stores above
  quadrant64_for_each(wacc, [&](int i, int j, double v, int) { S[i][j] = v; });
  __syncthreads();
  // X21 = -X22 W (X22 lower), W in S
  mfma_quadrant64(CL_B, [&](int row, int k) { return Xs[row][k]; },
                  [&](int k, int col) { return S[k][col]; }, acc);
  quadrant64_for_each(acc, [&](int i, int j, double v, int) {
    if (i < n2) X21[(size_t)i + (size_t)j * ldx] = -v;
  });
  if (bad && threadIdx.x == 0) atomicOr(err, ERRBIT_GJ_PIVOT);
}

// P = W^-1 for the b x b SPD diagonal block of the blocked Gauss-Jordan
// inverse (Cholesky, L^-1, X^T X); P is b x b with leading dimension b.
__global__ __launch_bounds__(CP_THREADS) void k_spd_inv64(int b, const double *W, int ldw,
                                                          double *P, unsigned int *err) {
  __shared__ double S[CL_B][CL_B + 1];
  __shared__ double Xs[CL_B][CL_B + 1];
  load_spd64(b, W, ldw, S);
  const bool bad = chol_inv_lds(S, Xs);
  store_xtx(b, Xs, P);
  if (bad && threadIdx.x == 0) atomicOr(err, ERRBIT_GJ_PIVOT);
}

// First half of a Gauss-Jordan panel K = [k0, k0 + kb) on W (n x n, ld n),
// one workgroup per 64-row / 64-column block g:
//   T[rows_g, :] = W[rows_g, K] P      (T: n x kb, ld n)
//   R[:, cols_g] = W[K, cols_g]        (R: kb x n, ld kb)
// and block 0 copies the next diagonal block W[K1, K1] (kb1 x kb1) to Dsrc
// (ld 64) when kb1 > 0, before the trailing update changes it.
__global__ __launch_bounds__(CP_THREADS) void k_gj_panel_a(int n, int k0, int kb,
                                                           const double *__restrict__ W,
                                                           const double *__restrict__ P,
                                                           double *__restrict__ T,
                                                           double *__restrict__ R, int k1, int kb1,
                                                           double *__restrict__ Dsrc) {
  __shared__ double As[CL_B][CL_B + 1];
  __shared__ double Bs[CL_B][CL_B + 1];
  const int g0 = blockIdx.x * CL_B;
  // every load of the workgroup is issued before the first is used (one
  // memory round trip): R's block of W[K, :], the next diagonal block, and the
  // operands W[rows_g, K] and P of T's block
  double v[STAGE_R], w[STAGE_R], a[STAGE_R], pv[STAGE_R];
#pragma unroll
  for (int r = 0; r < STAGE_R; ++r) {
    const int x = threadIdx.x + CP_THREADS * r, i = x & 63, k = x >> 6;
    v[r] = (i < kb && g0 + k < n) ? W[(size_t)(k0 + i) + (size_t)(g0 + k) * n] : 0.0;
    w[r] = (blockIdx.x == 0 && i < kb1 && k < kb1) ? W[(size_t)(k1 + i) + (size_t)(k1 + k) * n]
                                                   : 0.0;
    a[r] = (g0 + i < n && k < kb) ? W[(size_t)(g0 + i) + (size_t)(k0 + k) * n] : 0.0;
    pv[r] = (i < kb && k < kb) ? P[i + k * kb] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < STAGE_R; ++r) {
    const int x = threadIdx.x + CP_THREADS * r, i = x & 63, k = x >> 6;
    if (i < kb && g0 + k < n) R[(size_t)i + (size_t)(g0 + k) * kb] = v[r];
    if (blockIdx.x == 0 && i < kb1 && k < kb1) Dsrc[i + k * CL_B] = w[r];
    As[i][k] = a[r];
    Bs[i][k] = pv[r];
  }
  __syncthreads();
  double4_t acc[2][2];
  mfma_quadrant64(kb, [&](int row, int k) { return As[row][k]; },
                  [&](int k, int col) { return Bs[k][col]; }, acc);
  quadrant64_for_each(acc, [&](int i, int j, double v, int) {
    if (g0 + i < n && j < kb) T[(size_t)(g0 + i) + (size_t)j * n] = v;
  });
}

// Second half of the panel, for the 64-column / 64-row block g:
//   g = K's block:  W[K, K] = P
//   otherwise:      W[K, cols_g] = P R[:, cols_g];  W[rows_g, K] = -T[rows_g, :]
// (k0 is a multiple of 64, so K is exactly block k0 / 64.)  Touches only the
// panel's rows and columns of W.
__device__ void gj_finish_block(double (*As)[CL_B + 1], double (*Bs)[CL_B + 1], bool act, int g,
                                int n, int k0, int kb, double *__restrict__ W,
                                const double *__restrict__ P, const double *__restrict__ T,
                                const double *__restrict__ R) {
  const int g0 = g * CL_B;
  if (g0 == k0) {
    if (!act) return;
    double v[STAGE_R];
#pragma unroll
    for (int r = 0; r < STAGE_R; ++r) {
      const int x = threadIdx.x + CP_THREADS * r, i = x & 63, k = x >> 6;
      v[r] = (i < kb && k < kb) ? P[i + k * kb] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < STAGE_R; ++r) {
      const int x = threadIdx.x + CP_THREADS * r, i = x & 63, k = x >> 6;
      if (i < kb && k < kb) W[(size_t)(k0 + i) + (size_t)(k0 + k) * n] = v[r];
    }
    return;
  }
  if (act) {  // W[rows_g, K] = -T[rows_g, :]
    double v[STAGE_R];
#pragma unroll
    for (int r = 0; r < STAGE_R; ++r) {
      const int x = threadIdx.x + CP_THREADS * r, i = x & 63, k = x >> 6;
      v[r] = (g0 + i < n && k < kb) ? T[(size_t)(g0 + i) + (size_t)k * n] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < STAGE_R; ++r) {
      const int x = threadIdx.x + CP_THREADS * r, i = x & 63, k = x >> 6;
      if (g0 + i < n && k < kb) W[(size_t)(g0 + i) + (size_t)(k0 + k) * n] = -v[r];
    }
  }
  stage64(As, [&](int i, int k) { return (i < kb && k < kb) ? P[i + k * kb] : 0.0; }, act);
  stage64(Bs, [&](int i, int k) {
    return (i < kb && g0 + k < n) ? R[(size_t)i + (size_t)(g0 + k) * kb] : 0.0;
  }, act);
  __syncthreads();
  if (!act) return;
  double4_t acc[2][2];
  mfma_quadrant64(kb, [&](int row, int k) { return As[row][k]; },
                  [&](int k, int col) { return Bs[k][col]; }, acc);
  quadrant64_for_each(acc, [&](int i, int j, double v, int) {
    if (i < kb && g0 + j < n) W[(size_t)(k0 + i) + (size_t)(g0 + j) * n] = v;
  });
}

// Look-ahead inverse of the next diagonal block: D = Dsrc - T1 R1 (Dsrc: b x b
// copy of the block taken before this panel's update, ld 64; T1: the b rows
// of T for the next panel, ld ldt; R1: the b columns of R, ld ldr; kb = |K|),
// then Pn = D^-1 (b x b, ld b).
__device__ void gj_next_inverse(double (*S)[CL_B + 1], double (*Xs)[CL_B + 1], bool act, int b,
                                const double *Dsrc, const double *T1, int ldt, const double *R1,
                                int ldr, int kb, double *Pn, unsigned int *err) {
  // S = T1 (b x kb), Xs = R1 (kb x b); every load (with the thread's entries of
  // Dsrc) is issued before the first is used
  double4_t acc[2][2];
  double d[16];  // D's lower triangle: Dsrc less the product
  if (act) {
    double tv[STAGE_R], rv[STAGE_R];
#pragma unroll
    for (int r = 0; r < STAGE_R; ++r) {
      const int x = threadIdx.x + CP_THREADS * r, i = x & 63, k = x >> 6;
      tv[r] = (i < b && k < kb) ? T1[(size_t)i + (size_t)k * ldt] : 0.0;
      rv[r] = (i < kb && k < b) ? R1[(size_t)i + (size_t)k * ldr] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
    quadrant64_for_each(acc, [&](int i, int k, double, int q) {
      d[q] = (i < b && k <= i) ? Dsrc[i + k * CL_B] : 0.0;
    });
#pragma unroll
    for (int r = 0; r < STAGE_R; ++r) {
      const int x = threadIdx.x + CP_THREADS * r;
      S[x & 63][x >> 6] = tv[r];
      Xs[x & 63][x >> 6] = rv[r];
    }
  }
  __syncthreads();
  if (act) {
    mfma_quadrant64(kb, [&](int row, int k) { return S[row][k]; },
                    [&](int k, int col) { return Xs[k][col]; }, acc);
    quadrant64_for_each(acc, [&](int i, int k, double v, int q) {
      d[q] = (i < b && k <= i) ? d[q] - v : 0.0;
    });
  }
  __syncthreads();  // S and Xs are read by the product until here
  if (act) {
    quadrant64_for_each(acc, [&](int i, int k, double, int q) {
      S[i][k] = (i < b && k < b) ? d[q] : (i == k ? 1.0 : 0.0);
    });
  }
  __syncthreads();
  const bool bad = chol_inv_lds(S, Xs, act);
  if (act) store_xtx(b, Xs, Pn);
  if (bad && threadIdx.x == 0) atomicOr(err, ERRBIT_GJ_PIVOT);
}

// One Gauss-Jordan step after k_gj_panel_a, as one launch of three kinds of
// workgroup that touch disjoint data:
//   [0, tiles)             W -= T R on the 128 x 128 MFMA tiles, leaving the
//                          panel's rows and columns K alone
//   [tiles, tiles + nblk)  gj_finish_block: the panel's rows and columns
//   tiles + nblk (kb1 > 0) gj_next_inverse: P of the next panel
// The next panel's P depends only on T, R and the copied block, so it is
// formed while the trailing update runs instead of after it.
__global__ __launch_bounds__(512, 1) void k_gj_update(int n, int k0, int kb, double *W,
                                                      const double *P, const double *T,
                                                      const double *R, int kb1, const double *Dsrc,
                                                      double *Pn, unsigned int *err) {
  __shared__ double S[CL_B][CL_B + 1];
  __shared__ double Xs[CL_B][CL_B + 1];
  const int tm = (n + GB_M - 1) / GB_M, tiles = tm * tm, nblk = (n + CL_B - 1) / CL_B;
  const int bid = blockIdx.x;
  if (bid < tiles) {
    gemm_w8_tile<16, false, false>(bid, n, n, kb, -1.0, T, n, R, kb, 1.0, W, n, nullptr, 0,
                                   nullptr, 1, k0, k0 + kb);
    return;
  }
  const bool act = threadIdx.x < CP_THREADS;
  if (bid < tiles + nblk) {
    gj_finish_block(S, Xs, act, bid - tiles, n, k0, kb, W, P, T, R);
    return;
  }
  const int k1 = k0 + kb;
  gj_next_inverse(S, Xs, act, kb1, Dsrc, T + k1, n, R + (size_t)k1 * kb, kb, kb, Pn, err);
}

hipError_t chol_inv_leaf(int b, double *A, int lda, double *X, int ldx, unsigned int *err,
                         hipStream_t st) {
  if (b < 1 || b > CL_B) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_chol_inv64p, dim3(1), dim3(CP_THREADS), 0, st, b, A, lda, X, ldx, err);
  return hipGetLastError();
}

hipError_t chol_inv_node128(int n, double *A, int lda, double *X, int ldx, unsigned int *err,
                            hipStream_t st) {
  if (n <= CL_B || n > 2 * CL_B) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_chol_inv128, dim3(1), dim3(CP_THREADS), 0, st, n, A, lda, X, ldx, err);
  return hipGetLastError();
}

hipError_t spd_inverse64(int b, const double *W, int ldw, double *P, unsigned int *err,
                         hipStream_t st) {
  if (b < 1 || b > CL_B) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_spd_inv64, dim3(1), dim3(CP_THREADS), 0, st, b, W, ldw, P, err);
  return hipGetLastError();
}

hipError_t gj_panel_a(int n, int k0, int kb, const double *W, const double *P, double *T,
                      double *R, int k1, int kb1, double *Dsrc, hipStream_t st) {
  if (kb < 1 || kb > CL_B || kb1 < 0 || kb1 > CL_B || (k0 % CL_B) != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gj_panel_a, dim3(cdiv(n, CL_B)), dim3(CP_THREADS), 0, st, n, k0, kb, W, P,
                     T, R, k1, kb1, Dsrc);
  return hipGetLastError();
}

hipError_t gj_update(int n, int k0, int kb, double *W, const double *P, const double *T,
                     const double *R, int kb1, const double *Dsrc, double *Pn, unsigned int *err,
                     hipStream_t st) {
  if (kb < 1 || kb > CL_B || kb1 < 0 || kb1 > CL_B || (k0 % CL_B) != 0) return hipErrorInvalidValue;
  const int grid = (int)(cdiv(n, GB_M) * cdiv(n, GB_N) + cdiv(n, CL_B)) + (kb1 > 0 ? 1 : 0);
  hipLaunchKernelGGL(k_gj_update, dim3(grid), dim3(512), 0, st, n, k0, kb, W, P, T, R, kb1, Dsrc,
                     Pn, err);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// skinny products with KC right-hand columns (n x KC, column-major, ld n)
// ---------------------------------------------------------------------------
// C = A^T B.  A wave owns tn_wave_cols(KC) output rows (columns of A, contiguous
// in memory) and streams them 64 rows per group, U groups at a time, every load
// of A and of B (the rows of B from L2, shared by the workgroup's waves) issued
// before the first is used; no LDS and no barriers.  Eight columns per wave for
// KC = 7 halve the B rows fetched per element of A.  The 64 lanes' partial sums
// are then folded by a fixed butterfly.  lower: A is lower triangular, so
// column i only has rows r >= i.
__host__ __device__ constexpr int tn_wave_cols(int kc) { return kc > 8 ? 4 : 8; }
__host__ __device__ constexpr int tn_cols(int kc) { return 4 * tn_wave_cols(kc); }

template <int KC>
__global__ __launch_bounds__(256) void k_skinny_tn(int n, const double *__restrict__ A, int lda,
                                                   const double *__restrict__ B,
                                                   double *__restrict__ C, int lower) {
  constexpr int CW = tn_wave_cols(KC), U = 2;  // columns per wave, 64-row groups in flight
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ibase = blockIdx.x * tn_cols(KC) + CW * wave;
  if (ibase >= n) return;
  double acc[CW][KC];
#pragma unroll
  for (int c = 0; c < CW; ++c)
#pragma unroll
    for (int j = 0; j < KC; ++j) acc[c][j] = 0.0;
  const double *col[CW];
#pragma unroll
  for (int c = 0; c < CW; ++c) col[c] = A + (size_t)min(ibase + c, n - 1) * lda;
  const int rstart = lower ? (ibase & ~63) : 0;
  for (int r0 = rstart; r0 < n; r0 += 64 * U) {
    double a[U][CW], b[U][KC];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = min(r0 + 64 * u + lane, n - 1);
#pragma unroll
      for (int c = 0; c < CW; ++c) a[u][c] = col[c][r];
#pragma unroll
      for (int j = 0; j < KC; ++j) b[u][j] = B[(size_t)r + (size_t)j * n];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + 64 * u + lane;
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        const int i = ibase + c;
        const double av = (r < n && i < n && (!lower || r >= i)) ? a[u][c] : 0.0;
#pragma unroll
        for (int j = 0; j < KC; ++j) acc[c][j] = fma(av, b[u][j], acc[c][j]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CW; ++c)
#pragma unroll
    for (int j = 0; j < KC; ++j) {
      double v = acc[c][j];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      acc[c][j] = v;
    }
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < CW; ++c) {
      const int i = ibase + c;
      if (i < n)
#pragma unroll
        for (int j = 0; j < KC; ++j) C[(size_t)i + (size_t)j * n] = acc[c][j];
    }
  }
}

// C = A B, split over the columns of A into NN_SPLIT partial sums (reduced by
// k_skinny_reduce).  Thread = output row; B's rows for 256 columns of A at a
// time are staged through LDS (two barriers per 256 columns), and the column
// loop is unrolled by 8 so that 8 loads of A (coalesced over the rows) are in
// flight.  lower: row r only has columns c <= r.
constexpr int NN_SPLIT = 16;
constexpr int NN_CHUNK = 256;

template <int KC>
__global__ __launch_bounds__(256) void k_skinny_nn(int n, const double *__restrict__ A, int lda,
                                                   const double *__restrict__ B,
                                                   double *__restrict__ part, int lower) {
  constexpr int UC = 8;
  __shared__ double Bs[NN_CHUNK][KC];
  const int tid = threadIdx.x;
  const int r0 = blockIdx.x * 256, r = r0 + tid, rc = min(r, n - 1);
  const int per = (n + NN_SPLIT - 1) / NN_SPLIT;
  const int cb = blockIdx.y * per;
  int ce = min(n, cb + per);
  if (lower) ce = min(ce, r0 + 256);
  double acc[KC];
#pragma unroll
  for (int j = 0; j < KC; ++j) acc[j] = 0.0;
  for (int k0 = cb; k0 < ce; k0 += NN_CHUNK) {
    const int cnt = min(NN_CHUNK, ce - k0);
    __syncthreads();
    for (int x = tid; x < NN_CHUNK * KC; x += 256) {
      const int cc = x % NN_CHUNK, j = x / NN_CHUNK;
      Bs[cc][j] = cc < cnt ? B[(size_t)(k0 + cc) + (size_t)j * n] : 0.0;
    }
    __syncthreads();
    for (int u0 = 0; u0 < cnt; u0 += UC) {
      double a[UC];
#pragma unroll
      for (int u = 0; u < UC; ++u)
        a[u] = A[(size_t)rc + (size_t)min(k0 + u0 + u, ce - 1) * lda];
#pragma unroll
      for (int u = 0; u < UC; ++u) {  // branch-free, so that the loads above stay ahead
        const int c = k0 + u0 + u;
        const double av = (c < ce && (!lower || c <= r)) ? a[u] : 0.0;
#pragma unroll
        for (int j = 0; j < KC; ++j) acc[j] = fma(av, Bs[min(u0 + u, NN_CHUNK - 1)][j], acc[j]);
      }
    }
  }
  if (r < n) {
#pragma unroll
    for (int j = 0; j < KC; ++j) part[((size_t)blockIdx.y * KC + j) * n + r] = acc[j];
  }
}

template <int KC>
__global__ void k_skinny_reduce(int n, const double *part, double *C) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
#pragma unroll
  for (int j = 0; j < KC; ++j) {
    double s = 0.0;
    for (int q = 0; q < NN_SPLIT; ++q) s += part[((size_t)q * KC + j) * n + r];
    C[(size_t)r + (size_t)j * n] = s;
  }
}

template <int KC>
static hipError_t skinny_tn_k(int n, const double *A, int lda, const double *B, double *C,
                              bool lower, hipStream_t st) {
  hipLaunchKernelGGL(k_skinny_tn<KC>, dim3(cdiv(n, tn_cols(KC))), dim3(256), 0, st, n, A, lda, B, C,
                     lower ? 1 : 0);
  return hipGetLastError();
}

template <int KC>
static hipError_t skinny_nn_k(int n, const double *A, int lda, const double *B, double *C,
                              double *part, bool lower, hipStream_t st) {
  hipLaunchKernelGGL(k_skinny_nn<KC>, dim3(cdiv(n, 256), NN_SPLIT), dim3(256), 0, st, n, A, lda, B,
                     part, lower ? 1 : 0);
  hipLaunchKernelGGL(k_skinny_reduce<KC>, dim3(cdiv(n, 256)), dim3(256), 0, st, n, part, C);
  return hipGetLastError();
}

hipError_t skinny_tn(int n, int kc, const double *A, int lda, const double *B, double *C,
                     bool lower, hipStream_t st) {
  if (kc == 7) return skinny_tn_k<7>(n, A, lda, B, C, lower, st);
  if (kc == 14) return skinny_tn_k<14>(n, A, lda, B, C, lower, st);
  return hipErrorInvalidValue;
}

hipError_t skinny_nn(int n, int kc, const double *A, int lda, const double *B, double *C,
                     double *part, bool lower, hipStream_t st) {
  if (kc == 7) return skinny_nn_k<7>(n, A, lda, B, C, part, lower, st);
  if (kc == 14) return skinny_nn_k<14>(n, A, lda, B, C, part, lower, st);
  return hipErrorInvalidValue;
}

size_t skinny_part_doubles(int n) { return (size_t)NN_SPLIT * 16 * (size_t)n; }

// ---------------------------------------------------------------------------
// small dense helpers on n x k blocks (k <= 16) and k x k matrices (row-major)
// ---------------------------------------------------------------------------
// G = V^T W (k1 x k2, row-major): 64 partial sums over contiguous row ranges,
// then one block adds them in a fixed order.
constexpr int GM_BLOCKS = 64, GM_ROWS = 64;

__global__ __launch_bounds__(256) void k_gram_part(int n, int k1, const double *V, int k2,
                                                   const double *W, double *partial) {
  __shared__ double sv[GM_ROWS][33];
  const int tid = threadIdx.x;
  const int kk = k1 + k2;
  const int per = (n + GM_BLOCKS - 1) / GM_BLOCKS;
  const int r0 = blockIdx.x * per, r1 = min(n, r0 + per);
  const int a = tid / k2, b = tid % k2;
  const bool act = tid < k1 * k2;
  double s = 0.0;
  for (int c0 = r0; c0 < r1; c0 += GM_ROWS) {
    const int cnt = min(GM_ROWS, r1 - c0);
    __syncthreads();
    for (int x = tid; x < GM_ROWS * kk; x += 256) {
      const int q = x / GM_ROWS, rr = x % GM_ROWS;
      double v = 0.0;
      if (rr < cnt) v = q < k1 ? V[(size_t)q * n + c0 + rr] : W[(size_t)(q - k1) * n + c0 + rr];
      sv[rr][q] = v;
    }
    __syncthreads();
    if (act)
      for (int rr = 0; rr < cnt; ++rr) s = fma(sv[rr][a], sv[rr][k1 + b], s);
  }
  if (act) partial[(size_t)blockIdx.x * 256 + tid] = s;
}

__global__ void k_gram_sum(int cnt, const double *partial, double *G) {
  const int e = threadIdx.x;
  if (e >= cnt) return;
  double s = 0.0;
  for (int q = 0; q < GM_BLOCKS; ++q) s += partial[(size_t)q * 256 + e];
  G[e] = s;
}

hipError_t gram(int n, int k1, const double *V, int k2, const double *W, double *partial,
                double *G, hipStream_t st) {
  if (k1 < 1 || k2 < 1 || k1 + k2 > 32 || k1 * k2 > 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gram_part, dim3(GM_BLOCKS), dim3(256), 0, st, n, k1, V, k2, W, partial);
  hipLaunchKernelGGL(k_gram_sum, dim3(1), dim3(256), 0, st, k1 * k2, partial, G);
  return hipGetLastError();
}

// T = L^-T for G = L L^T (K x K, row-major): V T has orthonormal columns.
// One thread; K is a template parameter so the factors stay in registers.
template <int K>
__global__ void k_small_orth(const double *G, double *T, unsigned int *err) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double L[K][K], Li[K][K];
  bool bad = false;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    double s = G[j * K + j];
#pragma unroll
    for (int q = 0; q < j; ++q) s -= L[j][q] * L[j][q];
    if (!(s > 0.0)) {
      bad = true;
      s = 1e-300;
    }
    L[j][j] = sqrt(s);
#pragma unroll
    for (int i = j + 1; i < K; ++i) {
      double t = G[i * K + j];
#pragma unroll
      for (int q = 0; q < j; ++q) t -= L[i][q] * L[j][q];
      L[i][j] = t / L[j][j];
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    Li[j][j] = 1.0 / L[j][j];
#pragma unroll
    for (int i = j + 1; i < K; ++i) {
      double s = 0.0;
#pragma unroll
      for (int q = j; q < i; ++q) s += L[i][q] * Li[q][j];
      Li[i][j] = -s / L[i][i];
    }
  }
#pragma unroll
  for (int i = 0; i < K; ++i)
#pragma unroll
    for (int j = 0; j < K; ++j) T[i * K + j] = (j >= i) ? Li[j][i] : 0.0;
  if (bad) atomicOr(err, ERRBIT_GJ_PIVOT);
}

// S = (C + G)^-1 with C = diag(I_h, -I_h), k = 2h <= 16 (row-major):
// Gauss-Jordan with partial pivoting on [C + G | I] held in LDS, thread per
// element of the augmented matrix.
__global__ __launch_bounds__(512) void k_small_woodbury(int k, const double *G, double *S,
                                                       unsigned int *err) {
  __shared__ double a[16][33];
  __shared__ int best_s;
  const int tid = threadIdx.x, i = tid >> 5, j = tid & 31;
  const int h = k / 2;
  const bool act = i < k && j < 2 * k;
  if (act)
    a[i][j] = j < k ? G[i * k + j] + (i == j ? (i < h ? 1.0 : -1.0) : 0.0) : (j - k == i ? 1.0 : 0.0);
  __syncthreads();
  bool bad = false;
  for (int p = 0; p < k; ++p) {
    if (tid == 0) {
      int best = p;
      for (int r = p + 1; r < k; ++r)
        if (fabs(a[r][p]) > fabs(a[best][p])) best = r;
      best_s = best;
    }
    __syncthreads();
    const int best = best_s;
    if (best != p && tid < 2 * k) {
      const double t0 = a[p][tid];
      a[p][tid] = a[best][tid];
      a[best][tid] = t0;
    }
    __syncthreads();
    const double piv = a[p][p];
    if (!(fabs(piv) > 0.0)) bad = true;
    const double ip = 1.0 / piv;
    const double f = act ? a[i][p] : 0.0, rp = act ? a[p][j] * ip : 0.0;
    __syncthreads();
    if (act) a[i][j] = (i == p) ? rp : a[i][j] - f * rp;
    __syncthreads();
  }
  if (act && j < k) S[i * k + j] = 0.5 * (a[i][k + j] + a[j][k + i]);
  if (bad && tid == 0) atomicOr(err, ERRBIT_GJ_PIVOT);
}

// out(n x k2) = alpha A + beta B T (A n x k2, B n x k1, T k1 x k2 row-major,
// T == nullptr: identity).  Row-local, so out may alias A or B.
__global__ void k_small_apply(int n, int k1, int k2, double alpha, const double *A, double beta,
                              const double *B, const double *T, double *out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double b[16], o[16];
  for (int q = 0; q < k1; ++q) b[q] = B[(size_t)r + (size_t)q * n];
  for (int j = 0; j < k2; ++j) {
    double s = 0.0;
    if (T) {
      for (int q = 0; q < k1; ++q) s = fma(b[q], T[q * k2 + j], s);
    } else {
      s = b[j];
    }
    o[j] = (alpha != 0.0 ? alpha * A[(size_t)r + (size_t)j * n] : 0.0) + beta * s;
  }
  for (int j = 0; j < k2; ++j) out[(size_t)r + (size_t)j * n] = o[j];
}

hipError_t small_orth(int k, const double *G, double *T, unsigned int *err, hipStream_t st) {
  if (k != 7) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_small_orth<7>, dim3(1), dim3(64), 0, st, G, T, err);
  return hipGetLastError();
}
hipError_t small_woodbury(int k, const double *G, double *S, unsigned int *err, hipStream_t st) {
  if (k < 2 || k > 16 || (k & 1)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_small_woodbury, dim3(1), dim3(512), 0, st, k, G, S, err);
  return hipGetLastError();
}
hipError_t small_apply(int n, int k1, int k2, double alpha, const double *A, double beta,
                       const double *B, const double *T, double *out, hipStream_t st) {
  if (k1 < 1 || k1 > 16 || k2 < 1 || k2 > 16 || (!T && k1 != k2)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_small_apply, dim3(cdiv(n, 256)), dim3(256), 0, st, n, k1, k2, alpha, A, beta,
                     B, T, out);
  return hipGetLastError();
}

// deterministic pseudo-random fill in [-0.5, 0.5)
__global__ void k_fill_hash(size_t cnt, double *V, unsigned long long seed) {
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < cnt;
       x += (size_t)gridDim.x * blockDim.x) {
    unsigned long long h = (x + 1 + seed) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 27;
    V[x] = (double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
}

hipError_t fill_hash(size_t cnt, double *V, unsigned long long seed, hipStream_t st) {
  size_t g = cdiv(cnt, 256);
  hipLaunchKernelGGL(k_fill_hash, dim3((unsigned int)(g > 4096 ? 4096 : g)), dim3(256), 0, st, cnt,
                     V, seed);
  return hipGetLastError();
}

// out = A + V V^T (n x n; V n x 7 column-major); block (x, c) covers 256 rows
// of column c
// out = A + V V^T on and below the diagonal (the Cholesky recursion reads
// nothing above it): a thread per row r over RC_COLS columns c <= r, V's row r
// in registers and V's rows c, the same for the whole workgroup, through the
// scalar cache.
constexpr int RC_COLS = 8;
__global__ __launch_bounds__(256) void k_rank_copy(int n, const double *__restrict__ A,
                                                   const double *__restrict__ V,
                                                   double *__restrict__ out) {
  const int c0 = blockIdx.y * RC_COLS, r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n || blockIdx.x * 256 + 255 < c0) return;
  double vr[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) vr[q] = V[r + (size_t)q * n];
#pragma unroll
  for (int cc = 0; cc < RC_COLS; ++cc) {
    const int c = c0 + cc;
    if (c >= n) break;  // uniform
    if (c > r) continue;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 7; ++q) s = fma(vr[q], V[c + (size_t)q * n], s);
    const size_t x = (size_t)r + (size_t)c * n;
    out[x] = A[x] + s;
  }
}

// max_i X_ii^2 in one workgroup: a strided pass, then a fixed tree in LDS
__global__ __launch_bounds__(256) void k_diag_max_sq(int n, const double *X, int ldx, double *out) {
  __shared__ double m[256];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double x = X[(size_t)i + (size_t)i * ldx];
    v = fmax(v, x * x);
  }
  m[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = m[0];
}

hipError_t diag_max_sq(int n, const double *X, int ldx, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_max_sq, dim3(1), dim3(256), 0, st, n, X, ldx, out);
  return hipGetLastError();
}

hipError_t rank_copy(int n, int k, const double *A, const double *V, double *out, hipStream_t st) {
  if (k != 7 || n > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_rank_copy, dim3(cdiv(n, 256), cdiv(n, RC_COLS)), dim3(256), 0, st, n, A, V,
                     out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// camera blocks of Z^+: Xd[a] = X_a^T X_a - (MU)_a S (MU)_a^T - N_a N_a^T,
// X_a = block column a of the lower-triangular X (rows >= a cp), row-major
// cp x cp output.  One workgroup per camera.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cam_gram(int nc, int cp, int n, const double *X,
                                                  const double *MU, const double *S,
                                                  const double *Nb, double *Xd) {
  // accumulators follow the 9x9 upper triangle (entry (p, q), p <= q, has the
  // compile-time index p*9 - p(p-1)/2 + q - p); parameters beyond cp stay 0
  constexpr int MAXE = 45;
  __shared__ double red[4][MAXE];
  __shared__ double g[MAXE];
  const int a = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c0 = a * cp;
  double acc[MAXE];
#pragma unroll
  for (int e = 0; e < MAXE; ++e) acc[e] = 0.0;
  for (int r = c0 + tid; r < n; r += 256) {
    double x[9];
#pragma unroll
    for (int p = 0; p < 9; ++p) x[p] = p < cp ? X[(size_t)r + (size_t)(c0 + p) * n] : 0.0;
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int q = p; q < 9; ++q) {
        const int e = p * 9 - p * (p - 1) / 2 + (q - p);
        acc[e] = fma(x[p], x[q], acc[e]);
      }
  }
#pragma unroll
  for (int e = 0; e < MAXE; ++e) {
    double v = acc[e];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red[wave][e] = v;
  }
  __syncthreads();
  if (tid < MAXE) g[tid] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
  __syncthreads();
  if (tid < MAXE) {
    int r = tid, p = 0;
    while (r >= 9 - p) {
      r -= 9 - p;
      ++p;
    }
    const int q = p + r;
    if (q < cp) {
      double w = 0.0;
      for (int u = 0; u < 14; ++u) {
        double t = 0.0;
        for (int v = 0; v < 14; ++v) t = fma(S[u * 14 + v], MU[(size_t)(c0 + q) + (size_t)v * n], t);
        w = fma(MU[(size_t)(c0 + p) + (size_t)u * n], t, w);
      }
      double nn = 0.0;
      for (int u = 0; u < 7; ++u)
        nn = fma(Nb[(size_t)(c0 + p) + (size_t)u * n], Nb[(size_t)(c0 + q) + (size_t)u * n], nn);
      const double val = (g[tid] - w) - nn;
      Xd[(size_t)a * cp * cp + p * cp + q] = val;
      Xd[(size_t)a * cp * cp + q * cp + p] = val;
    }
  }
}

hipError_t cam_gram(const Structure &s, const double *X, const double *MU, const double *S,
                    const double *Nb, double *Xd, hipStream_t st) {
  if (s.cp > 9) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_cam_gram, dim3(s.nc), dim3(256), 0, st, s.nc, s.cp, s.N, X, MU, S, Nb, Xd);
  return hipGetLastError();
}

// covariance output of the Z^+ path: upper triangle of D Z^+ D per camera
__global__ void k_cam_output_cov(int nc, int cp, const double *Xd, const double *d, double *out) {
  const int a = blockIdx.x;
  const int per = cp * (cp + 1) / 2;
  for (int l = threadIdx.x; l < per; l += blockDim.x) {
    int r = l, j = 0;
    while (r >= cp - j) {
      r -= cp - j;
      ++j;
    }
    const int k = j + r;
    out[(size_t)a * per + l] = (Xd[(size_t)a * cp * cp + j * cp + k] * d[a * cp + j]) * d[a * cp + k];
  }
}

hipError_t cam_output_cov(const Structure &s, const double *Xd, const double *d, double *out,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_cam_output_cov, dim3(s.nc), dim3(64), 0, st, s.nc, s.cp, Xd, d, out);
  return hipGetLastError();
}

}  // namespace unc
