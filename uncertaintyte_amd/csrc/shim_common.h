// shim_common.h -- the body shared by the two C++ getCovariances entry points
// (include/uncertainty.h and include/uncertaintyTE/uncertainty.h): options in,
// unc_compute, statistics out.  The two headers declare cov::Options and
// cov::Statistic with different layouts, so everything here has internal
// linkage and is instantiated once per translation unit.
#pragma once
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "../../include/uncertainty_amd.h"

void unc_set_last_error(int code);  // uncertainty_api.cpp (thread-local)

namespace {

// computeCovariances (compute.cpp:458-540) over the C ABI, host buffers
template <class Options, class Statistic, class CRS>
int shim_compute(Options &options, Statistic &statistic, CRS &jacobian, double *h_camUnc,
                 double *h_ptUnc) {
  // compute.cpp:459: nothing to do without output buffers
  if (h_camUnc == NULL || h_ptUnc == NULL) return UNC_ERR_BAD_INPUT;
  unc_options o;
  unc_default_options(&o);
  o.algorithm = (int)options._algorithm;
  o.epsilon = options._epsilon;
  o.lambda = options._lambda;
  o.num_cams = options._numCams;
  o.cam_params = options._camParams;
  o.num_points = options._numPoints;
  o.num_obs = options._numObs;
  o.svd_remove_n = options._svdRemoveN;
  o.max_iter_te = options._maxIterTE;
  if (options._pts2fix) {
    o.pts2fix[0] = options._pts2fix[0];
    o.pts2fix[1] = options._pts2fix[1];
    o.pts2fix[2] = options._pts2fix[2];
  }
  if (jacobian.num_rows <= 0 || (long long)jacobian.rows.size() != jacobian.num_rows + 1LL)
    return UNC_ERR_BAD_INPUT;
  const long long nnz = jacobian.rows.back();
  if (nnz < 0 || (long long)jacobian.cols.size() < nnz || (long long)jacobian.values.size() < nnz)
    return UNC_ERR_BAD_INPUT;
  unc_stats st;
  std::memset(&st, 0, sizeof(st));
  int rc = unc_compute(NULL, &o, jacobian.num_rows, jacobian.num_cols, jacobian.rows.data(),
                       jacobian.cols.data(), jacobian.values.data(), h_camUnc, h_ptUnc, &st);
  if (rc != UNC_OK) return rc;
  // the reference sorts the first two ids of the caller's array in place (compute.cpp:285)
  if (options._pts2fix) {
    options._pts2fix[0] = st.fixed_pts[0];
    options._pts2fix[1] = st.fixed_pts[1];
  }
  statistic.timeCreateJ = st.time_create_j;
  statistic.timeFixJ = st.time_fix_j;
  statistic.timeNormJ = st.time_norm_j;
  statistic.timeMultiplyJJ = st.time_multiply_jj;
  statistic.timeSplitJJ = st.time_split_jj;
  statistic.timeInvV = st.time_inv_v;
  statistic.timeComposeZ = st.time_compose_z;
  statistic.timeInvZ = st.time_inv_z;
  statistic.timeTE = st.time_te;
  statistic.timePtsUnc = st.time_pts_unc;
  statistic.timeAll = st.time_all;
  statistic.lambda = st.lambda;
  free(statistic.fixedPts);
  statistic.fixedPts = (int *)malloc(3 * sizeof(int));
  if (statistic.fixedPts)
    for (int i = 0; i < 3; ++i) statistic.fixedPts[i] = st.fixed_pts[i];
  statistic.cycle_change.assign(st.cycle_change, st.cycle_change + st.num_terms);
  return UNC_OK;
}

}  // namespace
