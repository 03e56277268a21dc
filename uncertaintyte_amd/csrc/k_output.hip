// k_output.hip -- camera output packing and (e) point-covariance
// back-propagation, gfx950.
//
// Camera output: removeScaleJ4Z (compute.cpp:168-183), upper triangle row-major
// per camera of (X[j,k] + X[k,j]/2) d_j d_k -- the reference's operator
// precedence, kept for drop-in output parity.
//
// Points: the reference builds, per point, A = J_p, B = [J_c, -I2] per
// observation and Sigma = blockdiag(camera's own 9x9 block, I2), then
// pinv(A) B Sigma B^T pinv(A)^T with an Eigen COD (compute.cpp:382-441, 516-527).
// For full-column-rank A this equals
//     Sigma_p = D_p [Vi + Vi (sum_o W_s,o^T X_cc W_s,o) Vi] D_p
// with the scaled blocks already on the device, so one thread per point does
// ~300 FMA per observation and no pseudo-inverse.  Same fma order as the
// oracle.  Output slot: compact id mapped back past the partially sorted
// fixed triple (compute.cpp:443-447); fixed points keep the zero written first.
#include <algorithm>

#include "common.h"

namespace unc {

__global__ void k_cam_output(int nc, int cp, int N, const double *__restrict__ X,
                             const double *__restrict__ d, double *__restrict__ out) {
  const int a = blockIdx.x * blockDim.y + threadIdx.y;
  if (a >= nc) return;
  const int per = cp * (cp + 1) / 2;
  for (int l = threadIdx.x; l < per; l += blockDim.x) {
    int r = l, j = 0;
    while (r >= cp - j) { r -= cp - j; ++j; }
    const int k = j + r;
    const int J = a * cp + j, Kc = a * cp + k;
    out[(size_t)a * per + l] =
        ((X[(size_t)J + (size_t)Kc * N] + X[(size_t)Kc + (size_t)J * N] / 2) * d[J]) * d[Kc];
  }
}

hipError_t launch_cam_output(const Structure &s, const Blocks &b, const double *X,
                             double *cam_out, hipStream_t st) {
  dim3 blk(64, 4);
  hipLaunchKernelGGL(k_cam_output, dim3(cdiv(s.nc, 4)), blk, 0, st, s.nc, s.cp, s.N, X, b.d,
                     cam_out);
  return hipGetLastError();
}

// compact copy of the cameras' own 9x9 blocks of X (row-major per block)
__global__ void k_extract_diag(int nc, int cp, int N, const double *__restrict__ X,
                               double *__restrict__ Xd) {
  const int a = blockIdx.x;
  const int ne = cp * cp;
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    const int i = e / cp, j = e % cp;
    Xd[(size_t)a * ne + e] = X[(size_t)(a * cp + i) + (size_t)(a * cp + j) * N];
  }
}

hipError_t launch_extract_diag(const Structure &s, const double *X, double *Xd, hipStream_t st) {
  hipLaunchKernelGGL(k_extract_diag, dim3(s.nc), dim3(128), 0, st, s.nc, s.cp, s.N, X, Xd);
  return hipGetLastError();
}

// per point: G = sum over its observations in list order of
// M_o = W_o^T X_aa W_o (W_o = W_s of the observation, X_aa its camera's block
// of X, compact and row-major), then Sigma_p = D_p (V^-1 + V^-1 G V^-1) D_p
// at the original point id.  One thread per point; X blocks come from L2.
template <int CP>
__global__ __launch_bounds__(256) void k_point_cov(Structure s, Blocks b,
                                                   const double *__restrict__ Xd,
                                                   double *__restrict__ out, int npts_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= s.npc || !b.pt_has[p]) return;
  const int N = s.N;
  double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = s.pt_off[p]; t < s.pt_off[p + 1]; ++t) {
    const int o = s.pt_list[t];
    const int a = s.obs_cam[o];
    const double *Ws = b.Ws + (size_t)o * CP * 3;
    const double *Xa = Xd + (size_t)a * CP * CP;  // row-major block
    double w[CP * 3];
#pragma unroll
    for (int q = 0; q < CP * 3; ++q) w[q] = Ws[q];
    double H[CP * 3];
#pragma unroll
    for (int i = 0; i < CP; ++i) {
      const double x0 = Xa[i * CP];
      double h0 = x0 * w[0], h1 = x0 * w[1], h2 = x0 * w[2];
#pragma unroll
      for (int j = 1; j < CP; ++j) {
        const double x = Xa[i * CP + j];
        h0 = fma(x, w[j * 3 + 0], h0);
        h1 = fma(x, w[j * 3 + 1], h1);
        h2 = fma(x, w[j * 3 + 2], h2);
      }
      H[i * 3 + 0] = h0;
      H[i * 3 + 1] = h1;
      H[i * 3 + 2] = h2;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        double m = w[k] * H[l];
#pragma unroll
        for (int i = 1; i < CP; ++i) m = fma(w[i * 3 + k], H[i * 3 + l], m);
        G[k * 3 + l] = G[k * 3 + l] + m;
      }
  }
  const double *Vi = b.Vi + (size_t)p * 9;
  double v[9];
  for (int k = 0; k < 9; ++k) v[k] = Vi[k];
  double R[9], Q[9];
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l) {
      double r = v[k * 3 + 0] * G[0 * 3 + l];
      r = fma(v[k * 3 + 1], G[1 * 3 + l], r);
      r = fma(v[k * 3 + 2], G[2 * 3 + l], r);
      R[k * 3 + l] = r;
    }
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l) {
      double q = R[k * 3 + 0] * v[0 * 3 + l];
      q = fma(R[k * 3 + 1], v[1 * 3 + l], q);
      q = fma(R[k * 3 + 2], v[2 * 3 + l], q);
      Q[k * 3 + l] = v[k * 3 + l] + q;
    }
  const double *dp = b.d + N + 3 * p;
  int id = p;
  for (int i = 0; i < 3; ++i)
    if (id >= s.fix[i]) ++id;
  if (id < 0 || id >= npts_out) return;
  double *o6 = out + (size_t)id * 6;
  o6[0] = (Q[0] * dp[0]) * dp[0];
  o6[1] = (Q[1] * dp[0]) * dp[1];
  o6[2] = (Q[2] * dp[0]) * dp[2];
  o6[3] = (Q[4] * dp[1]) * dp[1];
  o6[4] = (Q[5] * dp[1]) * dp[2];
  o6[5] = (Q[8] * dp[2]) * dp[2];
}

template <int CP>
static void launch_point_cov_cp(const Structure &s, const Blocks &b, const double *Xd,
                                double *pt_out, int npts_out, hipStream_t st) {
  hipLaunchKernelGGL(k_point_cov<CP>, dim3(cdiv(s.npc, 256)), dim3(256), 0, st, s, b, Xd, pt_out,
                     npts_out);
}

hipError_t launch_point_output(const Structure &s, const Blocks &b, const double *Xd,
                               double *pt_out, int npts_out, hipStream_t st) {
  if (s.npc <= 0) return hipSuccess;
  switch (s.cp) {
    case 1: launch_point_cov_cp<1>(s, b, Xd, pt_out, npts_out, st); break;
    case 2: launch_point_cov_cp<2>(s, b, Xd, pt_out, npts_out, st); break;
    case 3: launch_point_cov_cp<3>(s, b, Xd, pt_out, npts_out, st); break;
    case 4: launch_point_cov_cp<4>(s, b, Xd, pt_out, npts_out, st); break;
    case 5: launch_point_cov_cp<5>(s, b, Xd, pt_out, npts_out, st); break;
    case 6: launch_point_cov_cp<6>(s, b, Xd, pt_out, npts_out, st); break;
    case 7: launch_point_cov_cp<7>(s, b, Xd, pt_out, npts_out, st); break;
    case 8: launch_point_cov_cp<8>(s, b, Xd, pt_out, npts_out, st); break;
    case 9: launch_point_cov_cp<9>(s, b, Xd, pt_out, npts_out, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Output check of the host API: one pass over the packed outputs on the
// device instead of a scan on the host after the copy.
__global__ __launch_bounds__(256) void k_flag_nonfinite(size_t n, const double *__restrict__ a,
                                                        unsigned int *flag) {
  bool bad = false;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    bad |= !isfinite(a[i]);
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

hipError_t flag_nonfinite(size_t n, const double *a, unsigned int *flag, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const int grid = (int)std::min<size_t>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_flag_nonfinite, dim3(grid), dim3(256), 0, st, n, a, flag);
  return hipGetLastError();
}

}  // namespace unc
