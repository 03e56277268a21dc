// k_svd.hip -- the SVD algorithm modes (SVD_QR_ITERATION, SVD_DEVIDE_AND_CONQUER)
// of computeCovariances (src/uncertaintyTE/compute.cpp:483-495 -> svdInverse
// :185-249) on the device, for problems small enough to hold the dense
// information matrix.
//
// The reference forms the unscaled M = J^T J over all cameras and points (no
// fixed points), runs LAPACK dgesvd / dgesdd, and keeps 1/sigma for the
// largest N - svdRemoveN singular values (or for sigma > epsilon), then writes
// the camera blocks of the pseudo-inverse with its (X[j,k] + X[k,j]/2) packing
// and no column scale (removeScaleJ4Z with a NULL scale, :168-183).  M is
// symmetric positive semi-definite, so its SVD is its eigen-decomposition
// (sigma = |lambda|, U = V up to the signs of negative rounding-level
// eigenvalues) and U diag(1/sigma) V^T = V diag(1/lambda) V^T: both LAPACK
// drivers compute the same pseudo-inverse, which is formed here from a
// two-sided parallel Jacobi eigen-decomposition:
//   rounds of a round-robin tournament pair every index with every other once
//   per sweep; the n/2 disjoint rotations of a round are applied together as
//   A <- J^T A J (each thread owns one 2x2 block of A, read from the previous
//   round's copy) and V <- V J; sweeps repeat until no pair passes the
//   rotation test |a_pq| > eps sqrt(|a_pp a_qq|).
// The reference's point step then dereferences the NULL scale and crashes
// (compute.cpp:402,413); here it is completed with unit scales and every point:
//   Sigma_p = V_p^-1 + V_p^-1 (sum_o W_o^T Sigma_c(o) W_o) V_p^-1,
// the closed form of pinv(A) B Sigma B^T pinv(A)^T (compute.cpp:516-527) for
// full-column-rank A, with Sigma_c the cameras' own blocks of the
// pseudo-inverse.
#include "common.h"

namespace unc {
namespace {

constexpr double kSvdEps = 2.220446049250313e-16;

// J^T (npar x nrows, column-major) from the CRS rows; one thread per row, so a
// row's repeated columns add in order.
__global__ void k_jt_scatter(int nrows, int npar, const int *__restrict__ rows,
                             const int *__restrict__ cols, const double *__restrict__ vals,
                             double *__restrict__ Jt, unsigned int *err) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  for (int e = rows[r]; e < rows[r + 1]; ++e) {
    const int c = cols[e];
    if (c < 0 || c >= npar) {
      atomicOr(err, ERRBIT_LAYOUT);
      return;
    }
    Jt[(size_t)c + (size_t)r * npar] += vals[e];
  }
}

// M <- (M + M^T) / 2, each off-diagonal pair set to 0.5 (lower + upper).  One
// workgroup per 64 x 64 tile (bi, bj), bi >= bj, of the lower triangle: it stages
// that tile and its mirror (bj, bi) through LDS, so that both are read and
// written along columns.
constexpr int SYM_T = 64;
__global__ __launch_bounds__(256) void k_symmetrize(int n, double *M) {
  __shared__ double L[SYM_T][SYM_T + 1], U[SYM_T][SYM_T + 1];
  const int t = blockIdx.x;
  int bi = 0;
  while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
  const int bj = t - bi * (bi + 1) / 2;
  const int i0 = bi * SYM_T, j0 = bj * SYM_T;
  const int r = threadIdx.x & 63, c0 = threadIdx.x >> 6;
  constexpr int PER = SYM_T * SYM_T / 256;
  double lv[PER], uv[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int c = c0 + 4 * k;
    lv[k] = (i0 + r < n && j0 + c < n) ? M[(size_t)(i0 + r) + (size_t)(j0 + c) * n] : 0.0;
    uv[k] = (j0 + r < n && i0 + c < n) ? M[(size_t)(j0 + r) + (size_t)(i0 + c) * n] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    L[r][c0 + 4 * k] = lv[k];
    U[r][c0 + 4 * k] = uv[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int c = c0 + 4 * k;
    const int li = i0 + r, lj = j0 + c;  // lower-tile entry (li, lj), mirror (lj, li) = U[c][r]
    if (li < n && lj < n && li > lj)
      M[(size_t)li + (size_t)lj * n] = 0.5 * (L[r][c] + U[c][r]);
    const int ui = j0 + r, uj = i0 + c;  // upper-tile entry (ui, uj), mirror (uj, ui) = L[c][r]
    if (ui < n && uj < n && uj > ui)
      M[(size_t)ui + (size_t)uj * n] = 0.5 * (L[c][r] + U[r][c]);
  }
}

// player at tournament position k in round r (m players, position 0 fixed)
__device__ __forceinline__ int player(int k, int r, int m) {
  return k == 0 ? 0 : ((k - 1 + r) % (m - 1)) + 1;
}

struct Rot {
  int p, q;   // p < q; q >= n marks the padding player (identity)
  double c, s, t;
  bool rot;
};

__device__ __forceinline__ Rot pair_rotation(int a, int r, int m, int n, const double *A) {
  Rot R;
  const int x = player(a, r, m), y = player(m - 1 - a, r, m);
  R.p = min(x, y);
  R.q = max(x, y);
  R.c = 1.0;
  R.s = 0.0;
  R.t = 0.0;
  R.rot = false;
  if (R.q >= n) return R;
  const double app = A[(size_t)R.p + (size_t)R.p * n], aqq = A[(size_t)R.q + (size_t)R.q * n];
  const double apq = A[(size_t)R.p + (size_t)R.q * n];
  R.rot = fabs(apq) > kSvdEps * sqrt(fabs(app)) * sqrt(fabs(aqq)) && apq != 0.0;
  if (R.rot) {
    const double tau = (aqq - app) / (2.0 * apq);
    double t = 1.0 / (fabs(tau) + sqrt(1.0 + tau * tau));
    R.t = tau < 0.0 ? -t : t;
    R.c = 1.0 / sqrt(1.0 + R.t * R.t);
    R.s = R.t * R.c;
  }
  return R;
}

// one round: Anew = J^T Aold J over all (pair a, pair b) 2x2 blocks
__global__ void k_jacobi_round_a(int n, int m, int r, const double *__restrict__ Aold,
                                 double *__restrict__ Anew, int *any_rot) {
  const int h = m / 2;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)h * h) return;
  const int a = (int)(idx % h), b = (int)(idx / h);
  const Rot Ra = pair_rotation(a, r, m, n, Aold);
  const Rot Rb = a == b ? Ra : pair_rotation(b, r, m, n, Aold);
  const int ia[2] = {Ra.p, Ra.q}, jb[2] = {Rb.p, Rb.q};
  double blk[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
      blk[u][v] = (ia[u] < n && jb[v] < n) ? Aold[(size_t)ia[u] + (size_t)jb[v] * n] : 0.0;
  double out[2][2];
  if (a == b) {
    // the pair's own 2x2 block: Rutishauser's update, a_pq set to zero
    const double app = blk[0][0], apq = blk[0][1], aqq = blk[1][1];
    out[0][0] = Ra.rot ? app - Ra.t * apq : app;
    out[1][1] = Ra.rot ? aqq + Ra.t * apq : aqq;
    out[0][1] = Ra.rot ? 0.0 : apq;
    out[1][0] = Ra.rot ? 0.0 : blk[1][0];
    if (Ra.rot) atomicOr(any_rot, 1);
  } else {
    // rows: J_a^T blk, then columns: (.) J_b, with J = [[c, s], [-s, c]]
    double t0[2], t1[2];
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      t0[v] = Ra.c * blk[0][v] - Ra.s * blk[1][v];
      t1[v] = Ra.s * blk[0][v] + Ra.c * blk[1][v];
    }
    out[0][0] = Rb.c * t0[0] - Rb.s * t0[1];
    out[0][1] = Rb.s * t0[0] + Rb.c * t0[1];
    out[1][0] = Rb.c * t1[0] - Rb.s * t1[1];
    out[1][1] = Rb.s * t1[0] + Rb.c * t1[1];
  }
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
      if (ia[u] < n && jb[v] < n) Anew[(size_t)ia[u] + (size_t)jb[v] * n] = out[u][v];
}

// V <- V J for the round's rotations (parameters from the same Aold)
__global__ void k_jacobi_round_v(int n, int m, int r, const double *__restrict__ Aold,
                                 double *__restrict__ V) {
  const int h = m / 2;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * h) return;
  const int i = (int)(idx % n), b = (int)(idx / n);
  const Rot R = pair_rotation(b, r, m, n, Aold);
  if (!R.rot) return;
  double *vp = V + (size_t)i + (size_t)R.p * n, *vq = V + (size_t)i + (size_t)R.q * n;
  const double x = *vp, y = *vq;
  *vp = R.c * x - R.s * y;
  *vq = R.s * x + R.c * y;
}

__global__ void k_identity(int n, double *V) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * n) return;
  V[idx] = (idx % n) == (idx / n) ? 1.0 : 0.0;
}

__global__ void k_diag(int n, const double *A, double *d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = A[(size_t)i + (size_t)i * n];
}

// camera blocks of the pseudo-inverse: Zc[a][j][k] = sum_i w_i V[a cp + j][i] V[a cp + k][i]
// (row-major cp x cp per camera), and the packed output (Z[j,k] + Z[k,j]/2), k >= j
__global__ void k_svd_cam_blocks(int nc, int cp, int n, const double *__restrict__ V,
                                 const double *__restrict__ w, double *__restrict__ Zc,
                                 double *__restrict__ cam_out) {
  const int a = blockIdx.x;
  __shared__ double blk[81];
  for (int e = threadIdx.x; e < cp * cp; e += blockDim.x) {
    const int j = e / cp, k = e % cp;
    const size_t rj = (size_t)a * cp + j, rk = (size_t)a * cp + k;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = fma(w[i] * V[rj + (size_t)i * n], V[rk + (size_t)i * n], s);
    blk[e] = s;
    Zc[(size_t)a * cp * cp + e] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double *o = cam_out + (size_t)a * (cp * (cp + 1) / 2);
    int l = 0;
    for (int j = 0; j < cp; ++j)
      for (int k = j; k < cp; ++k) o[l++] = blk[j * cp + k] + blk[k * cp + j] / 2;
  }
}

// point step with unit scales, every point; one thread per point scanning the
// observations (these modes only run on small problems)
__global__ void k_svd_points(int nobs, int cp, int nc, int np, const int *__restrict__ cols,
                             const double *__restrict__ vals, const double *__restrict__ Zc,
                             double *__restrict__ pt_out, unsigned int *err) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int rl = cp + 3, N = nc * cp;
  double V[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int o = 0; o < nobs; ++o) {
    const size_t e0 = (size_t)2 * o * rl;
    if ((cols[e0 + cp] - N) / 3 != p) continue;
    const int c = cols[e0] / cp;
    const double *r0 = vals + e0, *r1 = vals + e0 + rl;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) V[i][j] += r0[cp + i] * r0[cp + j] + r1[cp + i] * r1[cp + j];
    // W = J_c^T J_p (cp x 3), then S += W^T Zc W
    double W[9][3];
    for (int k = 0; k < cp; ++k)
      for (int j = 0; j < 3; ++j) W[k][j] = r0[k] * r0[cp + j] + r1[k] * r1[cp + j];
    const double *Z = Zc + (size_t)c * cp * cp;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < cp; ++k) {
          double t = 0.0;
          for (int l = 0; l < cp; ++l) t += Z[k * cp + l] * W[l][j];
          s += W[k][i] * t;
        }
        S[i][j] += s;
      }
  }
  // V^-1 by cofactors
  const double c00 = V[1][1] * V[2][2] - V[1][2] * V[2][1];
  const double c01 = V[1][2] * V[2][0] - V[1][0] * V[2][2];
  const double c02 = V[1][0] * V[2][1] - V[1][1] * V[2][0];
  const double det = V[0][0] * c00 + V[0][1] * c01 + V[0][2] * c02;
  if (!(fabs(det) > 0.0) || !isfinite(det)) {
    atomicOr(err, ERRBIT_SINGULAR_V);
    return;
  }
  const double id = 1.0 / det;
  double Vi[3][3];
  Vi[0][0] = c00 * id;
  Vi[1][0] = c01 * id;
  Vi[2][0] = c02 * id;
  Vi[0][1] = (V[0][2] * V[2][1] - V[0][1] * V[2][2]) * id;
  Vi[1][1] = (V[0][0] * V[2][2] - V[0][2] * V[2][0]) * id;
  Vi[2][1] = (V[0][1] * V[2][0] - V[0][0] * V[2][1]) * id;
  Vi[0][2] = (V[0][1] * V[1][2] - V[0][2] * V[1][1]) * id;
  Vi[1][2] = (V[0][2] * V[1][0] - V[0][0] * V[1][2]) * id;
  Vi[2][2] = (V[0][0] * V[1][1] - V[0][1] * V[1][0]) * id;
  double T[3][3], P[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T[i][j] = Vi[i][0] * S[0][j] + Vi[i][1] * S[1][j] + Vi[i][2] * S[2][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      P[i][j] = Vi[i][j] + T[i][0] * Vi[0][j] + T[i][1] * Vi[1][j] + T[i][2] * Vi[2][j];
  double *out = pt_out + (size_t)p * 6;
  out[0] = P[0][0];
  out[1] = P[0][1];
  out[2] = P[0][2];
  out[3] = P[1][1];
  out[4] = P[1][2];
  out[5] = P[2][2];
}

}  // namespace

hipError_t svd_jt_scatter(int nrows, int npar, const int *rows, const int *cols,
                          const double *vals, double *Jt, unsigned int *err, hipStream_t st) {
  hipLaunchKernelGGL(k_jt_scatter, dim3(cdiv(nrows, 256)), dim3(256), 0, st, nrows, npar, rows,
                     cols, vals, Jt, err);
  return hipGetLastError();
}

hipError_t svd_symmetrize(int n, double *M, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int nt = (n + SYM_T - 1) / SYM_T;
  hipLaunchKernelGGL(k_symmetrize, dim3(nt * (nt + 1) / 2), dim3(256), 0, st, n, M);
  return hipGetLastError();
}

hipError_t svd_identity(int n, double *V, hipStream_t st) {
  hipLaunchKernelGGL(k_identity, dim3(cdiv((long long)n * n, 256)), dim3(256), 0, st, n, V);
  return hipGetLastError();
}

hipError_t svd_jacobi_round(int n, int m, int r, const double *Aold, double *Anew, double *V,
                            int *any_rot, hipStream_t st) {
  const long long h = m / 2;
  hipLaunchKernelGGL(k_jacobi_round_v, dim3(cdiv((long long)n * h, 256)), dim3(256), 0, st, n, m, r,
                     Aold, V);
  hipLaunchKernelGGL(k_jacobi_round_a, dim3(cdiv(h * h, 256)), dim3(256), 0, st, n, m, r, Aold,
                     Anew, any_rot);
  return hipGetLastError();
}

hipError_t svd_diag(int n, const double *A, double *d, hipStream_t st) {
  hipLaunchKernelGGL(k_diag, dim3(cdiv(n, 256)), dim3(256), 0, st, n, A, d);
  return hipGetLastError();
}

hipError_t svd_cam_blocks(int nc, int cp, int n, const double *V, const double *w, double *Zc,
                          double *cam_out, hipStream_t st) {
  hipLaunchKernelGGL(k_svd_cam_blocks, dim3(nc), dim3(128), 0, st, nc, cp, n, V, w, Zc, cam_out);
  return hipGetLastError();
}

hipError_t svd_points(int nobs, int cp, int nc, int np, const int *cols, const double *vals,
                      const double *Zc, double *pt_out, unsigned int *err, hipStream_t st) {
  hipLaunchKernelGGL(k_svd_points, dim3(cdiv(np, 128)), dim3(128), 0, st, nobs, cp, nc, np, cols,
                     vals, Zc, pt_out, err);
  return hipGetLastError();
}

}  // namespace unc
