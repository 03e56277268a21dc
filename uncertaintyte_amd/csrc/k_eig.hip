// k_eig.hip -- batched symmetric eigen-decompositions of the output blocks:
// the eigen accessors of cov::Uncertainty (src/uncertaintyTE/uncertainty.cpp:
// 74-115 for the cameras, 146-185 for the points) on the device.
//
// The reference unpacks each block and calls Eigen's general (non-symmetric)
// EigenSolver, keeps the real parts and sorts the eigenvalues descending.  The
// unpacked blocks are exactly symmetric (both triangles come from the same
// packed value, uncertainty.cpp:54-58 and :123-131), so their eigenvalues are
// real and a symmetric method computes the same set.  Here each block is
// diagonalised by the cyclic Jacobi method with Rutishauser's rotation
// (Golub & Van Loan, Matrix Computations, 4th ed., Alg. 8.5.2-8.5.3):
//   tau = (a_qq - a_pp) / (2 a_pq),  t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
//   c = 1 / sqrt(1 + t^2),  s = t c,   (t evaluated in a division-light form)
//   a_pp -= t a_pq,  a_qq += t a_pq,  a_pq = 0,
//   a_rp, a_rq <- c a_rp - s a_rq, s a_rp + c a_rq        (r != p, q)
//   V_rp, V_rq <- c V_rp - s V_rq, s V_rp + c V_rq
// A rotation is applied when |a_pq| > eps sqrt(|a_pp a_qq|) (the
// relative-accuracy test of Demmel and Veselic, SIAM J. Matrix Anal. Appl. 13,
// 1992); the iteration ends when no pair passes that test.  Rotations are
// branch-free (t = 0 when skipped, which makes the update an exact identity),
// so the lanes of a wave stay converged.
//
// One lane owns one block; the block lives in registers (packed upper
// triangle, every index a compile-time constant after unrolling).  The point
// kernel is HBM-bound (6 doubles in, 3 + 9 out per point): it stages its
// tile through LDS so every global access is a contiguous wave-wide transfer.
#include "common.h"

namespace unc {
namespace {

// offset of (i, j), i <= j, in a row-major packed upper triangle of order n
__host__ __device__ constexpr int tri_index(int n, int i, int j) {
  return i * n - i * (i - 1) / 2 + (j - i);
}

constexpr double kJacobiEps = 2.220446049250313e-16;

template <int n>
struct SymBlock {
  double a[n * (n + 1) / 2];
  __device__ __forceinline__ double &at(int i, int j) {
    return i <= j ? a[tri_index(n, i, j)] : a[tri_index(n, j, i)];
  }
};

template <int n>
struct Vectors {
  double v[n][n];
};

// One cyclic sweep over every (p, q), p < q.
// With d = a_qq - a_pp and e = 2 a_pq, Rutishauser's t = sign(tau) / (|tau| +
// sqrt(1 + tau^2)), tau = d / e, equals sign(d) e / (|d| + sqrt(d^2 + e^2)):
// one square root and one reciprocal per rotation.  The block is scaled to
// max |a| in [1, 2) beforehand (exactly, by a power of two), so d^2 + e^2
// cannot overflow, and an a_pq whose square underflows is below 1e-154 of the
// block's largest entry.
// The rotation test: |a_pq| > eps sqrt(|a_pp a_qq|), squared.  An a_pq whose
// square is below DBL_MIN (under 1.5e-154 of the scaled block's largest entry)
// counts as zero, which keeps every operand of the rotation a normal number.
__device__ __forceinline__ bool needs_rotation(double apq, double app, double aqq) {
  return apq * apq > fmax(kJacobiEps * kJacobiEps * fabs(app * aqq), 2.2250738585072014e-308);
}

// True when no pair would rotate: the sweep that would find that is skipped.
template <int n>
__device__ __forceinline__ bool converged(SymBlock<n> &A) {
  bool any = false;
#pragma unroll
  for (int p = 0; p < n - 1; ++p)
#pragma unroll
    for (int q = p + 1; q < n; ++q) any = any || needs_rotation(A.at(p, q), A.at(p, p), A.at(q, q));
  return !any;
}

template <int n, bool VEC>
__device__ __forceinline__ void jacobi_sweep(SymBlock<n> &A, Vectors<n> &V) {
#pragma unroll
  for (int p = 0; p < n - 1; ++p) {
#pragma unroll
    for (int q = p + 1; q < n; ++q) {
      const double apq = A.at(p, q);
      const double app = A.at(p, p), aqq = A.at(q, q);
      const bool rot = needs_rotation(apq, app, aqq);
      const double d = aqq - app, e = 2.0 * apq;
      const double h = d * d + e * e;        // >= 4 DBL_MIN when rot
      const double r = h * rsqrt_normal(h);  // sqrt(h)
      double t = (d < 0.0 ? -e : e) * rcp_normal(fabs(d) + r);
      t = rot ? t : 0.0;  // a skipped pair gets the identity (and drops any NaN above)
      const double c = rsqrt_normal(1.0 + t * t);
      const double s = t * c;
      A.at(p, p) = fma(-t, apq, app);
      A.at(q, q) = fma(t, apq, aqq);
      A.at(p, q) = rot ? 0.0 : apq;
#pragma unroll
      for (int r2 = 0; r2 < n; ++r2) {
        if (r2 == p || r2 == q) continue;
        const double arp = A.at(r2, p), arq = A.at(r2, q);
        A.at(r2, p) = fma(c, arp, -s * arq);
        A.at(r2, q) = fma(s, arp, c * arq);
      }
      if (VEC) {
#pragma unroll
        for (int r2 = 0; r2 < n; ++r2) {
          const double vrp = V.v[r2][p], vrq = V.v[r2][q];
          V.v[r2][p] = fma(c, vrp, -s * vrq);
          V.v[r2][q] = fma(s, vrp, c * vrq);
        }
      }
    }
  }
}

// Diagonalise A in place (eigenvalues on its diagonal, V its eigenvectors when
// VEC), then sort descending into d (and the columns of V with them).
template <int n, bool VEC>
__device__ __forceinline__ void sym_eig(SymBlock<n> &A, Vectors<n> &V, double (&d)[n]) {
  if (VEC) {
#pragma unroll
    for (int i = 0; i < n; ++i)
#pragma unroll
      for (int j = 0; j < n; ++j) V.v[i][j] = i == j ? 1.0 : 0.0;
  }
  // exact power-of-two scaling to max |a| in [1, 2)
  double m = 0.0;
#pragma unroll
  for (int i = 0; i < n * (n + 1) / 2; ++i) m = fmax(m, fabs(A.a[i]));
  const int ex = (m > 0.0 && m <= 1.7976931348623157e308) ? ilogb(m) : 0;
#pragma unroll
  for (int i = 0; i < n * (n + 1) / 2; ++i) A.a[i] = ldexp(A.a[i], -ex);
  constexpr int kMaxSweeps = n <= 3 ? 12 : 24;
  for (int sweep = 0; sweep < kMaxSweeps && !converged(A); ++sweep) jacobi_sweep<n, VEC>(A, V);
#pragma unroll
  for (int i = 0; i < n; ++i) d[i] = ldexp(A.at(i, i), ex);
  // descending bubble network on register values (constant indices)
#pragma unroll
  for (int i = 0; i < n - 1; ++i) {
#pragma unroll
    for (int j = 0; j < n - 1 - i; ++j) {
      const bool sw = d[j] < d[j + 1];
      const double x = d[j], y = d[j + 1];
      d[j] = sw ? y : x;
      d[j + 1] = sw ? x : y;
      if (VEC) {
#pragma unroll
        for (int r = 0; r < n; ++r) {
          const double u = V.v[r][j], w = V.v[r][j + 1];
          V.v[r][j] = sw ? w : u;
          V.v[r][j + 1] = sw ? u : w;
        }
      }
    }
  }
  if (VEC) {
    // sign convention: the largest-magnitude component (first on ties) is positive
#pragma unroll
    for (int k = 0; k < n; ++k) {
      double mx = V.v[0][k];
#pragma unroll
      for (int r = 1; r < n; ++r) mx = fabs(V.v[r][k]) > fabs(mx) ? V.v[r][k] : mx;
      const double sg = mx < 0.0 ? -1.0 : 1.0;
#pragma unroll
      for (int r = 0; r < n; ++r) V.v[r][k] *= sg;
    }
  }
}

// ---- cameras: cp x cp eigenvalues (uncertainty.cpp:74-89) and the
// eigenvectors of the camera-centre block A(3:5, 3:5) (uncertainty.cpp:91-115)
constexpr int CAM_THREADS = 64;

template <int CP>
__global__ __launch_bounds__(CAM_THREADS) void k_eig_cameras(const double *__restrict__ cam,
                                                            int nc, double *__restrict__ vals,
                                                            double *__restrict__ cvecs) {
  constexpr int PER = CP * (CP + 1) / 2;
  const int c = blockIdx.x * CAM_THREADS + threadIdx.x;
  if (c >= nc) return;
  const double *src = cam + (size_t)c * PER;
  SymBlock<CP> A;
#pragma unroll
  for (int i = 0; i < PER; ++i) A.a[i] = src[i];
  if (CP >= 6 && cvecs) {
    SymBlock<3> B;
    Vectors<3> V;
    double d3[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) B.at(i, j) = A.at(3 + i, 3 + j);
    sym_eig<3, true>(B, V, d3);
    double *o = cvecs + (size_t)c * 9;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int r = 0; r < 3; ++r) o[3 * k + r] = V.v[r][k];
  }
  if (vals) {
    Vectors<CP> none;
    double d[CP];
    sym_eig<CP, false>(A, none, d);
    double *o = vals + (size_t)c * CP;
#pragma unroll
    for (int i = 0; i < CP; ++i) o[i] = d[i];
  }
}

// ---- points: 3x3 eigenvalues (uncertainty.cpp:146-162) and eigenvectors
// (uncertainty.cpp:164-185), packed input (xx, xy, xz, yy, yz, zz)
constexpr int PT_THREADS = 256;

template <bool VALS, bool VECS>
__global__ __launch_bounds__(PT_THREADS) void k_eig_points(const double *__restrict__ pts,
                                                          long long np,
                                                          double *__restrict__ vals,
                                                          double *__restrict__ vecs) {
  __shared__ double tile[PT_THREADS * 9];
  const long long base = (long long)blockIdx.x * PT_THREADS;
  const int cnt = (int)(np - base < PT_THREADS ? np - base : PT_THREADS);
  const int t = threadIdx.x;
  const double *src = pts + base * 6;
  for (int i = t; i < cnt * 6; i += PT_THREADS) tile[i] = src[i];
  __syncthreads();
  SymBlock<3> A;
#pragma unroll
  for (int i = 0; i < 6; ++i) A.a[i] = t < cnt ? tile[t * 6 + i] : 0.0;
  __syncthreads();
  Vectors<3> V;
  double d[3];
  sym_eig<3, VECS>(A, V, d);
  if (VECS) {
    if (t < cnt) {
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int r = 0; r < 3; ++r) tile[t * 9 + 3 * k + r] = V.v[r][k];
    }
    __syncthreads();
    double *dst = vecs + base * 9;
    for (int i = t; i < cnt * 9; i += PT_THREADS) dst[i] = tile[i];
    __syncthreads();
  }
  if (VALS) {
    if (t < cnt) {
#pragma unroll
      for (int k = 0; k < 3; ++k) tile[t * 3 + k] = d[k];
    }
    __syncthreads();
    double *dst = vals + base * 3;
    for (int i = t; i < cnt * 3; i += PT_THREADS) dst[i] = tile[i];
  }
}

}  // namespace

hipError_t launch_eig_cameras(const double *cam, int nc, int cp, double *vals, double *cvecs,
                              hipStream_t st) {
  if (nc <= 0 || (!vals && !cvecs)) return hipSuccess;
  const dim3 grid(cdiv(nc, CAM_THREADS)), block(CAM_THREADS);
  switch (cp) {
#define UNC_EIG_CASE(K)                                                          \
  case K:                                                                        \
    hipLaunchKernelGGL(k_eig_cameras<K>, grid, block, 0, st, cam, nc, vals, cvecs); \
    break;
    UNC_EIG_CASE(1) UNC_EIG_CASE(2) UNC_EIG_CASE(3) UNC_EIG_CASE(4) UNC_EIG_CASE(5)
    UNC_EIG_CASE(6) UNC_EIG_CASE(7) UNC_EIG_CASE(8) UNC_EIG_CASE(9)
#undef UNC_EIG_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_eig_points(const double *pts, long long np, double *vals, double *vecs,
                             hipStream_t st) {
  if (np <= 0 || (!vals && !vecs)) return hipSuccess;
  const dim3 grid((unsigned int)((np + PT_THREADS - 1) / PT_THREADS)), block(PT_THREADS);
  if (vals && vecs)
    hipLaunchKernelGGL((k_eig_points<true, true>), grid, block, 0, st, pts, np, vals, vecs);
  else if (vals)
    hipLaunchKernelGGL((k_eig_points<true, false>), grid, block, 0, st, pts, np, vals, vecs);
  else
    hipLaunchKernelGGL((k_eig_points<false, true>), grid, block, 0, st, pts, np, vals, vecs);
  return hipGetLastError();
}

}  // namespace unc
