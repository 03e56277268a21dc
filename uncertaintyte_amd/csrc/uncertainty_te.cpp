// uncertainty_te.cpp -- the reference's newer C++ interface
// (src/uncertaintyTE/uncertainty.h:24-158, uncertainty.cpp:20-185) over the C
// ABI: getCovariances with fixed-point selection, and cov::Uncertainty with its
// unpack (host) and eigen (device) accessors.
#include "../../include/uncertaintyTE/uncertainty.h"

#include "../../include/uncertainty_amd.h"
#include "shim_common.h"

namespace {

[[noreturn]] void throw_unc(const char *what, int rc) {
  unc_set_last_error(rc);
  throw std::runtime_error(std::string(what) + ": " + unc_error_string(rc));
}

}  // namespace

void getCovariances(cov::Options &options, cov::Statistic &statistic, ceres::CRSMatrix &jacobian,
                    double *points3D, cov::Uncertainty &uncertainties) {
  int fix[3] = {-1, -1, -1};
  // JacobianComposer::findPts2Fix (JacobianComposer.cpp:63-83), seeded
  int rc = unc_find_pts2fix(options._numPoints, points3D, fix);
  if (rc == UNC_OK) {
    free(options._pts2fix);
    options._pts2fix = (int *)malloc(3 * sizeof(int));
    if (!options._pts2fix) {
      rc = UNC_ERR_OOM;
    } else {
      for (int i = 0; i < 3; ++i) options._pts2fix[i] = fix[i];
    }
  }
  if (rc == UNC_OK) {
    uncertainties.init(options);
    rc = shim_compute(options, statistic, jacobian, uncertainties._camerasUnc.data(),
                      uncertainties._pointsUnc.data());
  }
  unc_set_last_error(rc);
  if (rc != UNC_OK) std::cerr << "getCovariances: " << unc_error_string(rc) << "\n";
}

namespace cov {

void Uncertainty::init(const Options &options) {
  _camParams = options._camParams;
  _numCams = options._numCams;
  _numPoints = options._numPoints;
  _nbCovarianceValuePerCam = _camParams * (_camParams + 1) / 2;
  _camerasUnc.assign(_numCams * _nbCovarianceValuePerCam, 0.0);
  _pointsUnc.assign(_numPoints * 6, 0.0);
}

const std::vector<double> Uncertainty::getCamerasUncRaw() const { return _camerasUnc; }

const std::vector<double> Uncertainty::getCameraUncMatrix(int id) const {
  // uncertainty.cpp:48-61: the packed row-major upper triangle, mirrored
  const std::size_t n = _camParams;
  std::vector<double> M(n * n);
  const double *src = _camerasUnc.data() + (std::size_t)id * _nbCovarianceValuePerCam;
  for (std::size_t i = 0, t = 0; i < n; ++i)
    for (std::size_t j = i; j < n; ++j, ++t) M[i * n + j] = M[j * n + i] = src[t];
  return M;
}

const std::vector<double> Uncertainty::getCamerasUncMatrices() const {
  const std::size_t nn = _camParams * _camParams;
  std::vector<double> out(_numCams * nn);
  for (std::size_t k = 0; k < _numCams; ++k) {
    const std::vector<double> M = getCameraUncMatrix((int)k);
    std::copy(M.begin(), M.end(), out.begin() + k * nn);
  }
  return out;
}

const std::vector<double> Uncertainty::getCamerasUncEigenValues() const {
  std::vector<double> ev(_numCams * _camParams);
  const int rc = unc_eig_cameras(NULL, (int)_numCams, (int)_camParams, _camerasUnc.data(),
                                 ev.data(), NULL);
  if (rc != UNC_OK) throw_unc("getCamerasUncEigenValues", rc);
  return ev;
}

const std::vector<double> Uncertainty::getCamerasCentersUncEigVec() const {
  std::vector<double> vec(_numCams * 9);
  const int rc = unc_eig_cameras(NULL, (int)_numCams, (int)_camParams, _camerasUnc.data(), NULL,
                                 vec.data());
  if (rc != UNC_OK) throw_unc("getCamerasCentersUncEigVec", rc);
  return vec;
}

const std::vector<double> Uncertainty::getPointsUncRaw() const { return _pointsUnc; }

const std::vector<double> Uncertainty::getPointUncMatrix(int id) const {
  // uncertainty.cpp:120-134
  const double *a = _pointsUnc.data() + 6 * (std::size_t)id;
  return std::vector<double>{a[0], a[1], a[2], a[1], a[3], a[4], a[2], a[4], a[5]};
}

const std::vector<double> Uncertainty::getPointsUncMatrices() const {
  std::vector<double> out;
  out.reserve(_numPoints * 9);
  for (std::size_t k = 0; k < _numPoints; ++k) {
    const std::vector<double> M = getPointUncMatrix((int)k);
    out.insert(out.end(), M.begin(), M.end());
  }
  return out;
}

const std::vector<double> Uncertainty::getPointsUncEigenValues() const {
  std::vector<double> ev(_numPoints * 3);
  const int rc = unc_eig_points(NULL, (int)_numPoints, _pointsUnc.data(), ev.data(), NULL);
  if (rc != UNC_OK) throw_unc("getPointsUncEigenValues", rc);
  return ev;
}

const std::vector<double> Uncertainty::getPointsUncEigVectors() const {
  std::vector<double> vec(_numPoints * 9);
  const int rc = unc_eig_points(NULL, (int)_numPoints, _pointsUnc.data(), NULL, vec.data());
  if (rc != UNC_OK) throw_unc("getPointsUncEigVectors", rc);
  return vec;
}

}  // namespace cov
