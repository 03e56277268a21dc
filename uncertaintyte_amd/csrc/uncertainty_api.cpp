// uncertainty_api.cpp -- the reference's C++ entry point (include/uncertainty.h:
// 48-61, semantics of computeCovariances, compute.cpp:458-540) over the C ABI.
#include "../../include/uncertainty.h"

#include "../../include/uncertainty_amd.h"
#include "shim_common.h"

namespace {
thread_local int g_last_error = UNC_OK;
}

void unc_set_last_error(int code) { g_last_error = code; }

int getCovariancesLastError() { return g_last_error; }

void getCovariances(cov::Options &options, cov::Statistic &statistic, ceres::CRSMatrix &jacobian,
                    double *h_camUnc, double *h_ptUnc) {
  const int rc = shim_compute(options, statistic, jacobian, h_camUnc, h_ptUnc);
  g_last_error = rc;
  if (rc != UNC_OK) std::cerr << "getCovariances: " << unc_error_string(rc) << "\n";
}
