/*
 * uncertaintyTE/uncertainty.h -- the reference's newer C++ interface
 * (src/uncertaintyTE/uncertainty.h:24-158 of alicevision/uncertaintyTE), kept
 * source-compatible and backed by the MI355X pipeline of uncertainty_amd.h.
 *
 * This header and include/uncertainty.h both declare cov::Options and
 * cov::Statistic, with different layouts (this one has EAlgorithm and the
 * extra bool _debug), exactly as in the reference: a translation unit
 * includes one of them.  The library exports both getCovariances overloads.
 *
 * Differences from the reference, all documented where they occur:
 *  - getCovariances selects the three fixed points with the reference's
 *    randomised "most distant triple" search (JacobianComposer.cpp:63-83),
 *    seeded with a fixed value instead of std::random_device, so a call is
 *    reproducible; Options::_pts2fix is allocated with malloc() (the reference
 *    uses new[] and frees with free()) and a previous array is freed first.
 *  - Statistic times are true seconds (the reference stores ten times the
 *    seconds, compute.cpp:33); Statistic::fixedPts is allocated with malloc().
 *  - The eigen accessors run on the GPU (unc_eig_cameras / unc_eig_points);
 *    they throw std::runtime_error when that fails (for instance without a
 *    device).  Eigenvalues are those of the reference (real, descending);
 *    eigenvectors are ordered by descending eigenvalue with their
 *    largest-magnitude component positive, where the reference returns Eigen
 *    EigenSolver's unspecified order and signs.
 *  - On an error getCovariances prints the reason to std::cerr and leaves the
 *    outputs zero; getCovariancesLastError() returns the UNC_* code.
 */
#pragma once
#include <cstdlib>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#define _USE_MATH_DEFINES
#include <cmath>

#if defined(__has_include)
#if __has_include("ceres/crs_matrix.h")
#include "ceres/crs_matrix.h"
#define UNC_HAVE_CERES_CRS 1
#endif
#endif

#ifndef UNC_HAVE_CERES_CRS
namespace ceres {
struct CRSMatrix {
  CRSMatrix() : num_cols(0), num_rows(0) {}
  int num_cols;
  int num_rows;
  std::vector<int> cols;
  std::vector<int> rows;
  std::vector<double> values;
};
}  // namespace ceres
#endif

namespace cov {

enum EAlgorithm {
  eAlgorithmSvdQrIteration = 0,
  eAlgorithmSvdDeviceAndconquer = 1,
  eAlgorithmSvdTaylorExpansion = 2
};

// names as the reference spells them (uncertainty.h:33-51)
inline std::string EAlgorithm_enumToString(EAlgorithm alg) {
  if (alg == eAlgorithmSvdQrIteration) return "SVD_QR_ITERATION";
  if (alg == eAlgorithmSvdDeviceAndconquer) return "SVD_DEVIDE_AND_CONQUER";
  if (alg == eAlgorithmSvdTaylorExpansion) return "TAYLOR_EXPANSION";
  return "not defined";
}

inline EAlgorithm EAlgorithm_stringToEnum(const std::string &name) {
  for (int a = 0; a < 3; ++a)
    if (name == EAlgorithm_enumToString((EAlgorithm)a)) return (EAlgorithm)a;
  throw std::runtime_error(std::string("Unrecognized algorithm: ") + name);
}

struct Options {
 public:
  double _epsilon, _lambda;
  EAlgorithm _algorithm;
  int _numCams, _camParams, _numPoints, _numObs, _svdRemoveN, _maxIterTE;
  int *_pts2fix = NULL;
  bool _debug = false;

  Options() : _lambda(-1), _svdRemoveN(-1), _maxIterTE(-1) {}

  Options(int numCams, int camParams, int numPoints, int numObs)
      : _epsilon(1e-10), _lambda(-1), _algorithm(eAlgorithmSvdTaylorExpansion),
        _numCams(numCams), _camParams(camParams), _numPoints(numPoints), _numObs(numObs),
        _svdRemoveN(-1), _maxIterTE(-1) {}

  Options(EAlgorithm algorithm, double eps_or_lamb, int numCams, int camParams, int numPoints,
          int numObs)
      : _epsilon(eps_or_lamb), _lambda(eps_or_lamb), _algorithm(algorithm), _numCams(numCams),
        _camParams(camParams), _numPoints(numPoints), _numObs(numObs), _svdRemoveN(-1),
        _maxIterTE(-1) {}

  ~Options() { free(_pts2fix); }
};

struct Statistic {
  double timeCreateJ, timeFixJ, timeNormJ, timeMultiplyJJ, timeSplitJJ, timeInvV, timeComposeZ,
      timeInvZ, timeTE, timePtsUnc, timeAll;
  double lambda;
  int *fixedPts = NULL;
  std::vector<double> cycle_change;

  ~Statistic() { free(fixedPts); }
};

/* Output container (uncertainty.h:84-143): cp(cp+1)/2 upper-triangle values
 * per camera, 6 per point (xx, xy, xz, yy, yz, zz). */
struct Uncertainty {
  std::size_t _camParams = 0;
  std::size_t _numCams = 0;
  std::size_t _numPoints = 0;
  std::size_t _nbCovarianceValuePerCam = 0;
  std::vector<double> _camerasUnc;
  std::vector<double> _pointsUnc;

  void init(const Options &options);

  /** upper triangle of each camera covariance: cp(cp+1)/2 values per camera */
  const std::vector<double> getCamerasUncRaw() const;
  /** cp x cp covariance of camera id (symmetric) */
  const std::vector<double> getCameraUncMatrix(int id) const;
  /** cp x cp covariance of every camera */
  const std::vector<double> getCamerasUncMatrices() const;
  /** cp eigenvalues per camera, descending (GPU) */
  const std::vector<double> getCamerasUncEigenValues() const;
  /** 9 values per camera: eigenvectors of the camera-centre block (GPU, cp >= 6) */
  const std::vector<double> getCamerasCentersUncEigVec() const;
  /** 6 values per point */
  const std::vector<double> getPointsUncRaw() const;
  /** 3 x 3 covariance of point id */
  const std::vector<double> getPointUncMatrix(int id) const;
  /** 3 x 3 covariance of every point */
  const std::vector<double> getPointsUncMatrices() const;
  /** 3 eigenvalues per point, descending (GPU) */
  const std::vector<double> getPointsUncEigenValues() const;
  /** 9 values per point: its eigenvectors (GPU) */
  const std::vector<double> getPointsUncEigVectors() const;
};
}  // namespace cov

/**
 * uncertainty.h:146-158 / uncertainty.cpp:20-30: choose three fixed points
 * among points3D (3 doubles per point) as the most distant triple found by a
 * seeded random search, size the output container, and compute the camera and
 * point covariances (semantics of computeCovariances, compute.cpp:458-540).
 */
void getCovariances(cov::Options &options, cov::Statistic &statistic,
                    ceres::CRSMatrix &jacobian, double *points3D,
                    cov::Uncertainty &uncertainties);

/* Last error code of getCovariances on this thread (UNC_* of uncertainty_amd.h). */
int getCovariancesLastError();
