// k_scene.hip -- the Jacobian of a reconstruction on the device:
// JacobianComposer::scene2Jacobian (src/uncertaintyTE/JacobianComposer.cpp:
// 84-194) with analytic derivatives in place of Ceres' automatic ones.
//
// Residual of one observation (snavely_reprojection_error.h:148-302):
//   p = R(aa) (X - C)                   (AngleAxisRotatePointRC, :110-140)
//   x = p0 / p2,  y = p1 / p2,  r2 = x^2 + y^2
//   g = 1 + k1 r2 + k2 r2^2
//   res = f g (x, y) - observation
// Camera models and their free parameters (columns, in this order):
//   SIMPLE_PINHOLE  aa, C            (f, k1, k2 held at the camera's values)
//   PINHOLE         aa, C, f         (k1, k2 held)
//   SIMPLE_RADIAL   aa, C, f, k1     (no k2 term)
//   RADIAL          aa, C, f, k1, k2
// then the point's X (3 columns).  The CRS layout is the .jacob one: two rows
// per observation, cp + 3 nonzeros per row, camera columns cam*cp + j, point
// columns num_cams*cp + 3*pt + j.
//
// Derivatives (one lane per observation):
//   d(x,y)/dp = [1/p2 0 -p0/p2^2; 0 1/p2 -p1/p2^2]
//   d(res)/d(x,y) = f [g + x gx, x gy; y gx, g + y gy],  gx = 2x dg/dr2, ...
//   dp/dX = R,  dp/dC = -R,  dp/daa = -R [X - C]x Jr(aa)
// with Jr the right Jacobian of SO(3): I - (1 - cos t)/t^2 [aa]x +
// (t - sin t)/t^3 [aa]x^2.  The reference's first-order branch for t^2 <= DBL_EPSILON
// (snavely_reprojection_error.h:134-139) differs from the exact rotation used
// here by O(t^2) <= 2.2e-16.
#include "common.h"

namespace unc {
namespace {

constexpr int SJ_THREADS = 256;

// CP = 6, 7, 8, 9 selects the model (SIMPLE_PINHOLE, PINHOLE, SIMPLE_RADIAL,
// RADIAL); cams hold aa, C, f, k1, k2 for every model.
template <int CP>
__global__ __launch_bounds__(SJ_THREADS) void k_scene_jacobian(
    int nc, int np, int nobs, const double *__restrict__ cams, const double *__restrict__ pts,
    const int *__restrict__ obs_cam, const int *__restrict__ obs_pt, int *__restrict__ rows,
    int *__restrict__ cols, double *__restrict__ vals, unsigned int *err) {
  constexpr int rl = CP + 3;
  const int o = blockIdx.x * SJ_THREADS + threadIdx.x;
  if (o == 0) rows[2 * nobs] = 2 * nobs * rl;
  if (o >= nobs) return;
  const int c = obs_cam[o], q = obs_pt[o];
  rows[2 * o] = 2 * o * rl;
  rows[2 * o + 1] = (2 * o + 1) * rl;
  if (c < 0 || c >= nc || q < 0 || q >= np) {
    atomicOr(err, ERRBIT_RANGE);
    return;
  }
  const double *cm = cams + (size_t)c * 9;
  const double w0 = cm[0], w1 = cm[1], w2 = cm[2];
  const double f = cm[6], k1 = cm[7], k2 = CP == 8 ? 0.0 : cm[8];
  const double d0 = pts[(size_t)q * 3 + 0] - cm[3];
  const double d1 = pts[(size_t)q * 3 + 1] - cm[4];
  const double d2 = pts[(size_t)q * 3 + 2] - cm[5];
  // Rodrigues: R = I + a K + b K^2, K = [w]x
  const double th2 = w0 * w0 + w1 * w1 + w2 * w2;
  const double th = sqrt(th2);
  // below t = 0.01 the series (truncation < 3e-16) replace the cancelling forms
  const bool small = th < 1e-2;
  const double th4 = th2 * th2;
  const double a = small ? 1.0 - th2 / 6.0 + th4 / 120.0 : sin(th) / th;
  const double b = small ? 0.5 - th2 / 24.0 + th4 / 720.0 : (1.0 - cos(th)) / th2;
  const double c1 = b;  // (1 - cos t) / t^2
  const double c2 = small ? 1.0 / 6.0 - th2 / 120.0 + th4 / 5040.0 : (th - sin(th)) / (th2 * th);
  double K[3][3] = {{0.0, -w2, w1}, {w2, 0.0, -w0}, {-w1, w0, 0.0}};
  double K2[3][3], R[3][3], Jr[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) s += K[i][k] * K[k][j];
      K2[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double id = i == j ? 1.0 : 0.0;
      R[i][j] = id + a * K[i][j] + b * K2[i][j];
      Jr[i][j] = id - c1 * K[i][j] + c2 * K2[i][j];
    }
  const double p0 = R[0][0] * d0 + R[0][1] * d1 + R[0][2] * d2;
  const double p1 = R[1][0] * d0 + R[1][1] * d1 + R[1][2] * d2;
  const double p2 = R[2][0] * d0 + R[2][1] * d1 + R[2][2] * d2;
  if (!(p2 != 0.0)) atomicOr(err, ERRBIT_PROJECTION);
  const double iz = 1.0 / p2;
  const double x = p0 * iz, y = p1 * iz;
  const double r2 = x * x + y * y;
  const double g = 1.0 + k1 * r2 + k2 * r2 * r2;
  const double gr = k1 + 2.0 * k2 * r2;
  const double gx = 2.0 * x * gr, gy = 2.0 * y * gr;
  // d(res)/dp = d(res)/d(x,y) . d(x,y)/dp
  const double u00 = f * (g + x * gx), u01 = f * x * gy;
  const double u10 = f * y * gx, u11 = f * (g + y * gy);
  double dp[2][3];
  dp[0][0] = u00 * iz;
  dp[0][1] = u01 * iz;
  dp[0][2] = -(u00 * x + u01 * y) * iz;
  dp[1][0] = u10 * iz;
  dp[1][1] = u11 * iz;
  dp[1][2] = -(u10 * x + u11 * y) * iz;
  // dp/dX = R
  double JX[2][3];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      JX[r][j] = dp[r][0] * R[0][j] + dp[r][1] * R[1][j] + dp[r][2] * R[2][j];
  // dp/daa = -R [d]x Jr
  const double D[3][3] = {{0.0, -d2, d1}, {d2, 0.0, -d0}, {-d1, d0, 0.0}};
  double M[3][3];  // [d]x Jr
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M[i][j] = D[i][0] * Jr[0][j] + D[i][1] * Jr[1][j] + D[i][2] * Jr[2][j];
  double row[2][12];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      row[r][j] = -(JX[r][0] * M[0][j] + JX[r][1] * M[1][j] + JX[r][2] * M[2][j]);  // aa
      row[r][3 + j] = -JX[r][j];                                                   // C
    }
    const double xr = r == 0 ? x : y;
    row[r][6] = xr * g;                // f
    row[r][7] = f * xr * r2;           // k1
    row[r][8] = f * xr * r2 * r2;      // k2
  }
  const int N = nc * CP;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const size_t base = (size_t)(2 * o + r) * rl;
#pragma unroll
    for (int j = 0; j < CP; ++j) {
      cols[base + j] = c * CP + j;
      vals[base + j] = row[r][j];
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cols[base + CP + j] = N + 3 * q + j;
      vals[base + CP + j] = JX[r][j];
    }
  }
}

}  // namespace

int scene_cam_params(int model) {
  switch (model) {
    case UNC_CAM_SIMPLE_PINHOLE: return 6;
    case UNC_CAM_PINHOLE: return 7;
    case UNC_CAM_SIMPLE_RADIAL: return 8;
    case UNC_CAM_RADIAL: return 9;
    default: return -1;
  }
}

hipError_t launch_scene_jacobian(int model, int nc, int np, int nobs, const double *cams,
                                 const double *pts, const int *obs_cam, const int *obs_pt,
                                 int *rows, int *cols, double *vals, unsigned int *err,
                                 hipStream_t st) {
  const dim3 grid(cdiv(nobs > 0 ? nobs : 1, SJ_THREADS)), block(SJ_THREADS);
  switch (scene_cam_params(model)) {
    case 6:
      hipLaunchKernelGGL(k_scene_jacobian<6>, grid, block, 0, st, nc, np, nobs, cams, pts, obs_cam,
                         obs_pt, rows, cols, vals, err);
      break;
    case 7:
      hipLaunchKernelGGL(k_scene_jacobian<7>, grid, block, 0, st, nc, np, nobs, cams, pts, obs_cam,
                         obs_pt, rows, cols, vals, err);
      break;
    case 8:
      hipLaunchKernelGGL(k_scene_jacobian<8>, grid, block, 0, st, nc, np, nobs, cams, pts, obs_cam,
                         obs_pt, rows, cols, vals, err);
      break;
    case 9:
      hipLaunchKernelGGL(k_scene_jacobian<9>, grid, block, 0, st, nc, np, nobs, cams, pts, obs_cam,
                         obs_pt, rows, cols, vals, err);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace unc
