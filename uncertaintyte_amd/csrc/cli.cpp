// cli.cpp -- `uncertainty` command line tool with the reference's flags
// (src/software/main_uncertainty.cpp:33-93): read a .jacob, a COLMAP text
// model or a BAL problem file, compute the covariances on the MI355X, write a .cov.
//
//   uncertainty -in=scene.jacob [-in_form=JACOBIAN] -out=scene.cov
//   uncertainty -in=colmap_dir [-in_form=COLMAP] [-cam=SIMPLE_RADIAL] -out=scene.cov
//   uncertainty -in=problem.txt -in_form=BAL -out=scene.cov
//               [-alg=TAYLOR_EXPANSION] [-debug] [-legacy_time_units]
//
// As in the reference, a .jacob input always runs TAYLOR_EXPANSION
// (JacobianIO.cpp:25 ignores -alg); a path ending in .jacob is read as one
// whatever -in_form says.  -legacy_time_units prints stage times the way the
// reference does (ten times the seconds, compute.cpp:33).  -debug writes the
// dense matrices the reference prints on the TE path (compute.cpp:505,508):
// Z.txt (the scaled Schur complement) and iZ.txt (the Taylor result), each as
// `name = [a b ... ;...];` (ScaledDenseMatrix::printBlock2Matlab3,
// ScaledDenseMatrix.cpp:133-143); the sparse J, M, U, W, V dumps are not
// written (the device never forms them as matrices).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/uncertainty.h"
#include "../../include/uncertainty_amd.h"

namespace {

// JacobianIO::loadJacobian (JacobianIO.cpp:32-63), token for token, with the
// stream state checked so a truncated file is an error instead of garbage.
bool load_jacobian(const std::string &path, ceres::CRSMatrix &J, cov::Options &opt) {
  std::ifstream f(path);
  if (!f.good()) {
    std::cerr << "\nThe input file doesn't exist.\n";
    return false;
  }
  int numJ;
  int fx[3];
  opt._algorithm = TAYLOR_EXPANSION;
  opt._epsilon = 1e-10;
  opt._svdRemoveN = 7;
  opt._maxIterTE = -1;
  f >> opt._lambda >> opt._numCams >> opt._camParams >> opt._numPoints >> opt._numObs;
  f >> fx[0] >> fx[1] >> fx[2];
  if (fx[0] >= 0 && fx[1] >= 0 && fx[2] >= 0) {
    opt._pts2fix = (int *)malloc(3 * sizeof(int));
    for (int i = 0; i < 3; ++i) opt._pts2fix[i] = fx[i];
  }
  f >> J.num_rows >> J.num_cols >> numJ;
  if (!f || J.num_rows < 0 || numJ < 0) return false;
  J.rows.assign(J.num_rows + 1, 0);
  J.cols.assign(numJ, 0);
  J.values.assign(numJ, 0.0);
  for (int i = 0; i <= J.num_rows; ++i) f >> J.rows[i];
  for (int i = 0; i < numJ; ++i) f >> J.cols[i];
  for (int i = 0; i < numJ; ++i) f >> J.values[i];
  if (!f) {
    std::cerr << "\nTruncated .jacob file.\n";
    return false;
  }
  return true;
}

// saveResults (IO.cpp:15-60)
bool save_results(const std::string &path, const cov::Options &o, const cov::Statistic &s,
                  int per, const double *cam, const double *pt, double tscale) {
  std::ofstream out(path);
  if (!out.good()) return false;
  out << "# ---- Covariance v0.1 ----\n";
  out << "# Number of cameras: " << o._numCams << "\n";
  out << "# Number of camera parameters: " << o._camParams << "\n";
  out << "# Number of points in 3D: " << o._numPoints << "\n";
  out << "# Number of observations: " << o._numObs << "\n";
  static const char *kAlg[3] = {"SVD_QR_ITERATION", "SVD_DEVIDE_AND_CONQUER", "TAYLOR_EXPANSION"};
  const bool te = o._algorithm == TAYLOR_EXPANSION;
  out << "# Used algorithm: " << (o._algorithm >= 0 && o._algorithm < 3 ? kAlg[o._algorithm] : "not defined")
      << "\n";
  if (te) {  // the TE lines only, as the reference writes them (IO.cpp:26-44)
    if (s.fixedPts != NULL)
      out << "# Fixed points: " << s.fixedPts[0] << ", " << s.fixedPts[1] << ", " << s.fixedPts[2] << "\n";
    out << "# Used lambda: " << s.lambda << "\n";
    out << "# Loading jacobian time: " << s.timeCreateJ * tscale << "s\n";
    out << "# Normalization of jacobian time: " << s.timeNormJ * tscale << "s\n";
    out << "# Compose information matrix time: " << s.timeMultiplyJJ * tscale << "s\n";
    out << "# Split infromation matrix time: " << s.timeSplitJJ * tscale << "s\n";
    out << "# Inverse V time: " << s.timeInvV * tscale << "s\n";
    out << "# Compose Z time: " << s.timeComposeZ * tscale << "s\n";
    out << "# Inverse Z time: " << s.timeInvZ * tscale << "s\n";
    out << "# Taylor expansion time: " << s.timeTE * tscale << "s\n";
    out << "# Point uncertainty time: " << s.timePtsUnc * tscale << "s\n";
    out << "# TE number of iterations: " << s.cycle_change.size() << "\n";
    out << "# TE cycle change: ";
    for (size_t i = 0; i < s.cycle_change.size(); ++i) out << s.cycle_change[i] << " ";
    out << "\n";
  }
  out << "# Time of the algorithm: " << s.timeAll * tscale << "s\n";
  out << "# -----------------------\n";
  for (int i = 0; i < o._numCams; ++i) {
    for (int j = 0; j < per; ++j) out << cam[(size_t)i * per + j] << " ";
    out << "\n";
  }
  for (int i = 0; i < o._numPoints; ++i) {
    for (int j = 0; j < 6; ++j) out << pt[(size_t)i * 6 + j] << " ";
    out << "\n";
  }
  return true;
}

// printBlock2Matlab3 of a dense n x n column-major matrix into <name>.txt
bool write_matlab(const std::string &name, int n, const std::vector<double> &A) {
  std::ofstream f(name + ".txt");
  if (!f.good()) return false;
  f << name << " = [";
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) f << A[(size_t)i + (size_t)j * n] << " ";
    f << (i == n - 1 ? "" : ";");
  }
  f << "];";
  return f.good();
}

// the TE path once more through the C ABI, keeping Z and the Taylor result
bool dump_debug(const cov::Options &opt, const ceres::CRSMatrix &J) {
  unc_options o;
  unc_default_options(&o);
  o.lambda = opt._lambda;
  o.num_cams = opt._numCams;
  o.cam_params = opt._camParams;
  o.num_points = opt._numPoints;
  o.num_obs = opt._numObs;
  for (int i = 0; i < 3; ++i) o.pts2fix[i] = opt._pts2fix ? opt._pts2fix[i] : -1;
  const int N = opt._numCams * opt._camParams;
  const int per = opt._camParams * (opt._camParams + 1) / 2;
  std::vector<double> Z((size_t)N * N), X((size_t)N * N);
  std::vector<double> cam((size_t)opt._numCams * per), pt((size_t)opt._numPoints * 6);
  unc_debug dbg;
  std::memset(&dbg, 0, sizeof(dbg));
  dbg.Z = Z.data();
  dbg.X = X.data();
  const int rc = unc_compute_debug(NULL, &o, J.num_rows, J.num_cols, J.rows.data(), J.cols.data(),
                                   J.values.data(), cam.data(), pt.data(), NULL, &dbg);
  if (rc != UNC_OK) {
    std::cerr << "uncertainty: debug dump: " << unc_error_string(rc) << "\n";
    return false;
  }
  return write_matlab("Z", N, Z) && write_matlab("iZ", N, X);
}

bool flag(const char *arg, const char *name, std::string *val) {
  // gflags accepts -name=value and --name=value
  const char *a = arg;
  while (*a == '-') ++a;
  size_t n = std::strlen(name);
  if (std::strncmp(a, name, n) != 0) return false;
  if (a[n] == '=') {
    *val = a + n + 1;
    return true;
  }
  if (a[n] == 0) {
    *val = "true";
    return true;
  }
  return false;
}

// COLMAP text model -> Jacobian (ColmapIO::read + JacobianComposer::scene2Jacobian,
// JacobianComposer.cpp:123-194): cp from the camera model, fixed points by
// unc_find_pts2fix, the Jacobian by the device kernel.
bool load_colmap(const std::string &dir, const std::string &cam, const std::string &alg,
                 ceres::CRSMatrix &J, cov::Options &opt) {
  static const char *kModels[4] = {"SIMPLE_PINHOLE", "PINHOLE", "SIMPLE_RADIAL", "RADIAL"};
  int model = -1;
  for (int m = 0; m < 4; ++m)
    if (cam == kModels[m]) model = m;
  if (model < 0) {
    std::cerr << "unknown camera model " << cam << "\n";
    return false;
  }
  int algorithm = -1;
  static const char *kAlgs[3] = {"SVD_QR_ITERATION", "SVD_DEVIDE_AND_CONQUER", "TAYLOR_EXPANSION"};
  for (int a = 0; a < 3; ++a)
    if (alg == kAlgs[a]) algorithm = a;
  if (algorithm < 0) {
    std::cerr << "Unrecognized algorithm: " << alg << "\n";
    return false;
  }
  unc_scene *sc = nullptr;
  int rc = unc_scene_read_colmap(dir.c_str(), &sc);
  if (rc != UNC_OK) {
    std::cerr << "cannot read the COLMAP model in " << dir << ": " << unc_error_string(rc) << "\n";
    return false;
  }
  int nc = 0, np = 0, nobs = 0;
  unc_scene_counts(sc, &nc, &np, &nobs);
  std::vector<double> cams((size_t)nc * 9), pts((size_t)np * 3);
  std::vector<int> oc(nobs), op(nobs);
  unc_scene_arrays(sc, cams.data(), pts.data(), oc.data(), op.data());
  unc_scene_destroy(sc);
  int fix[3];
  rc = unc_find_pts2fix(np, pts.data(), fix);
  const int cp = 6 + model;
  if (rc == UNC_OK) {
    J.num_rows = 2 * nobs;
    J.num_cols = nc * cp + 3 * np;
    J.rows.resize(2 * (size_t)nobs + 1);
    J.cols.resize(2 * (size_t)nobs * (cp + 3));
    J.values.resize(J.cols.size());
    rc = unc_scene_jacobian(NULL, model, nc, np, nobs, cams.data(), pts.data(), oc.data(),
                            op.data(), J.rows.data(), J.cols.data(), J.values.data());
  }
  if (rc != UNC_OK) {
    std::cerr << "cannot compose the Jacobian: " << unc_error_string(rc) << "\n";
    return false;
  }
  opt._algorithm = algorithm;
  opt._epsilon = 1e-10;
  opt._lambda = -1;
  opt._numCams = nc;
  opt._camParams = cp;
  opt._numPoints = np;
  opt._numObs = nobs;
  opt._svdRemoveN = 7;
  opt._maxIterTE = -1;
  opt._pts2fix = (int *)malloc(3 * sizeof(int));
  for (int i = 0; i < 3; ++i) opt._pts2fix[i] = fix[i];
  return true;
}

// BAL problem file -> Jacobian (the reference's Matlab route: BALProblem,
// BuildProblem with bundle type 0 and unit weights, buildJacobian,
// matlabInterface.cpp:19-213): all nine camera parameters free
// (UNC_CAM_RADIAL), fixed points by unc_find_pts2fix (setPts2Fix,
// auxCmd.cpp:43-66), TAYLOR_EXPANSION whatever -alg says and no SVD mode count,
// as buildJacobian sets them (matlabInterface.cpp:203-212).
bool load_bal(const std::string &path, ceres::CRSMatrix &J, cov::Options &opt) {
  unc_scene *sc = nullptr;
  int rc = unc_scene_read_bal(path.c_str(), &sc);
  if (rc != UNC_OK) {
    std::cerr << "cannot read the BAL problem " << path << ": " << unc_error_string(rc) << "\n";
    return false;
  }
  int nc = 0, np = 0, nobs = 0;
  unc_scene_counts(sc, &nc, &np, &nobs);
  std::vector<double> cams((size_t)nc * 9), pts((size_t)np * 3);
  std::vector<int> oc(nobs), op(nobs);
  unc_scene_arrays(sc, cams.data(), pts.data(), oc.data(), op.data());
  unc_scene_destroy(sc);
  int fix[3];
  rc = unc_find_pts2fix(np, pts.data(), fix);
  const int cp = 9;
  if (rc == UNC_OK) {
    J.num_rows = 2 * nobs;
    J.num_cols = nc * cp + 3 * np;
    J.rows.resize(2 * (size_t)nobs + 1);
    J.cols.resize(2 * (size_t)nobs * (cp + 3));
    J.values.resize(J.cols.size());
    rc = unc_scene_jacobian(NULL, UNC_CAM_RADIAL, nc, np, nobs, cams.data(), pts.data(),
                            oc.data(), op.data(), J.rows.data(), J.cols.data(), J.values.data());
  }
  if (rc != UNC_OK) {
    std::cerr << "cannot compose the Jacobian: " << unc_error_string(rc) << "\n";
    return false;
  }
  opt._algorithm = TAYLOR_EXPANSION;
  opt._epsilon = 1e-10;
  opt._lambda = -1;
  opt._numCams = nc;
  opt._camParams = cp;
  opt._numPoints = np;
  opt._numObs = nobs;
  opt._svdRemoveN = -1;
  opt._maxIterTE = -1;
  opt._pts2fix = (int *)malloc(3 * sizeof(int));
  for (int i = 0; i < 3; ++i) opt._pts2fix[i] = fix[i];
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  // defaults as in the reference (main_uncertainty.cpp:33-49)
  std::string in = ".", in_form = "COLMAP", out = ".", alg = "TAYLOR_EXPANSION",
              cam = "SIMPLE_RADIAL";
  std::string v;
  bool legacy_units = false, debug = false;
  for (int i = 1; i < argc; ++i) {
    if (flag(argv[i], "in_form", &v)) in_form = v;
    else if (flag(argv[i], "in", &v)) in = v;
    else if (flag(argv[i], "out", &v)) out = v;
    else if (flag(argv[i], "alg", &v)) alg = v;
    else if (flag(argv[i], "cam", &v)) cam = v;
    else if (flag(argv[i], "debug", &v)) debug = (v != "false");
    else if (flag(argv[i], "legacy_time_units", &v)) legacy_units = (v != "false");
    else {
      std::cerr << "unknown flag " << argv[i] << "\n";
      return 2;
    }
  }
  const bool is_jacob = in.size() > 6 && in.substr(in.size() - 6) == ".jacob";
  cov::Options opt;
  cov::Statistic stat = cov::Statistic();
  ceres::CRSMatrix J;
  if (in_form == "JACOBIAN" || is_jacob) {
    if (!load_jacobian(in, J, opt)) return 1;
  } else if (in_form == "COLMAP") {
    if (!load_colmap(in, cam, alg, J, opt)) return 1;
  } else if (in_form == "BAL") {
    if (!load_bal(in, J, opt)) return 1;
  } else {
    std::cerr << "unsupported input format " << in_form << " (JACOBIAN, COLMAP or BAL)\n";
    return 2;
  }
  const int per = opt._camParams * (opt._camParams + 1) / 2;
  std::vector<double> camu((size_t)opt._numCams * per, 0.0), ptu((size_t)opt._numPoints * 6, 0.0);
  getCovariances(opt, stat, J, camu.data(), ptu.data());
  int rc = getCovariancesLastError();
  if (rc != UNC_OK) {
    std::cerr << "uncertainty: " << unc_error_string(rc) << " (code " << rc << ")\n";
    return 1;
  }
  if (debug && opt._algorithm == TAYLOR_EXPANSION && !dump_debug(opt, J)) return 1;
  if (!save_results(out, opt, stat, per, camu.data(), ptu.data(), legacy_units ? 10.0 : 1.0)) {
    std::cerr << "cannot write " << out << "\n";
    return 1;
  }
  std::cout << "Main function... [done]\n";
  return 0;
}
