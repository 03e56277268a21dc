/*
 * uncertainty.h -- the reference's public C++ header, kept source-compatible
 * (include/uncertainty.h:1-62 of alicevision/uncertaintyTE), backed by the
 * MI355X pipeline of uncertainty_amd.h.
 *
 * Layout notes:
 *  - cov::Options / cov::Statistic keep the reference's member order and
 *    constructors; default member initialisers were added for the pointers
 *    (they do not change the layout) so a default-constructed Statistic can be
 *    destroyed safely.
 *  - Both destructors call free(), as the reference does; this implementation
 *    allocates Statistic::fixedPts with malloc(), so the pairing is correct
 *    here (the reference allocates it with new[], compute.cpp:290).  Callers
 *    that fill Options::_pts2fix themselves must use malloc() for the same
 *    reason (the reference's JacobianIO uses new[], JacobianIO.cpp:48).
 *  - ceres::CRSMatrix comes from Ceres when its header is on the include path;
 *    otherwise a definition with Ceres' member names is provided (only
 *    num_rows, num_cols, rows, cols and values are used, compute.cpp:474).
 *  - Statistic times are true seconds (the reference stores ten times the
 *    seconds, compute.cpp:33).
 */
#pragma once
#include <cstdlib>
#include <iostream>
#include <memory>
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

#define SVD_QR_ITERATION 0
#define SVD_DEVIDE_AND_CONQUER 1
#define TAYLOR_EXPANSION 2

namespace cov {
struct Options {
 public:
  double _epsilon, _lambda;
  int _algorithm, _numCams, _camParams, _numPoints, _numObs, _svdRemoveN, _maxIterTE;
  int *_pts2fix = NULL;

  Options() : _lambda(-1), _svdRemoveN(-1), _maxIterTE(-1) {}

  Options(int numCams, int camParams, int numPoints, int numObs)
      : _epsilon(1e-10), _lambda(-1), _algorithm(TAYLOR_EXPANSION), _numCams(numCams),
        _camParams(camParams), _numPoints(numPoints), _numObs(numObs), _svdRemoveN(-1),
        _maxIterTE(-1) {}

  Options(int algorithm, double eps_or_lamb, int numCams, int camParams, int numPoints, int numObs)
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
}  // namespace cov

/* The reference declares this for Linux without a definition
 * (include/uncertainty.h:55-61); here it runs the MI355X pipeline.  Returns
 * nothing, as in the reference: on error it prints the reason to std::cerr and
 * leaves the outputs zero.  unc_compute() in uncertainty_amd.h reports codes. */
#ifdef _WIN32
extern "C" __declspec(dllexport) void getCovariances(cov::Options &options,
                                                     cov::Statistic &statistic,
                                                     ceres::CRSMatrix &jacobian,
                                                     double *h_camUnc, double *h_ptUnc);
#else
void getCovariances(cov::Options &options, cov::Statistic &statistic,
                    ceres::CRSMatrix &jacobian, double *h_camUnc, double *h_ptUnc);
#endif

/* Last error code of getCovariances on this thread (UNC_* of uncertainty_amd.h). */
int getCovariancesLastError();
