// Test driver for the newer C++ interface (include/uncertaintyTE/uncertainty.h),
// written the way a caller of the reference's src/uncertaintyTE/uncertainty.h
// uses it.  Modes:
//   layout                      print sizeof/offsets of the structs
//   fix  <in.bin>               fixed-point choice only (no device needed):
//                               getCovariances on a Jacobian with no rows
//   run  <in.bin> <out.bin>     getCovariances + every accessor
//   unpack <cov.bin> <out.bin>  unpack accessors on given packed outputs
// in.bin: int32 nc, cp, np, nobs, nrows, ncols; int32 rows[nrows+1];
//         int32 cols[nnz]; f64 vals[nnz]; f64 points3D[3 np]
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <vector>

#include "uncertaintyTE/uncertainty.h"

namespace {
bool read_input(const char *path, cov::Options &opt, ceres::CRSMatrix &J, std::vector<double> &P) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  int h[6];
  bool ok = std::fread(h, 4, 6, f) == 6;
  if (ok) {
    opt._numCams = h[0];
    opt._camParams = h[1];
    opt._numPoints = h[2];
    opt._numObs = h[3];
    J.num_rows = h[4];
    J.num_cols = h[5];
    J.rows.resize(h[4] + 1);
    ok = std::fread(J.rows.data(), 4, J.rows.size(), f) == J.rows.size();
  }
  if (ok) {
    const size_t nnz = (size_t)J.rows.back();
    J.cols.resize(nnz);
    J.values.resize(nnz);
    P.resize(3 * (size_t)h[2]);
    ok = std::fread(J.cols.data(), 4, nnz, f) == nnz &&
         std::fread(J.values.data(), 8, nnz, f) == nnz &&
         std::fread(P.data(), 8, P.size(), f) == P.size();
  }
  std::fclose(f);
  return ok;
}

void put(FILE *g, const std::vector<double> &v) {
  const long long n = (long long)v.size();
  std::fwrite(&n, 8, 1, g);
  std::fwrite(v.data(), 8, v.size(), g);
}
}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string mode = argv[1];
  if (mode == "layout") {
    std::printf("%zu %zu %zu %zu %zu %zu\n", sizeof(cov::Options), offsetof(cov::Options, _algorithm),
                offsetof(cov::Options, _pts2fix), offsetof(cov::Options, _debug),
                sizeof(cov::Statistic), sizeof(cov::Uncertainty));
    std::printf("%s %d\n", cov::EAlgorithm_enumToString(cov::eAlgorithmSvdDeviceAndconquer).c_str(),
                (int)cov::EAlgorithm_stringToEnum("TAYLOR_EXPANSION"));
    try {
      cov::EAlgorithm_stringToEnum("bogus");
    } catch (const std::runtime_error &) {
      std::printf("throws\n");
    }
    return 0;
  }
  if (mode == "unpack" && argc >= 4) {
    FILE *f = std::fopen(argv[2], "rb");
    int h[3];
    if (!f || std::fread(h, 4, 3, f) != 3) return 3;
    cov::Options opt(h[0], h[1], h[2], 0);
    cov::Uncertainty unc;
    unc.init(opt);
    if (std::fread(unc._camerasUnc.data(), 8, unc._camerasUnc.size(), f) != unc._camerasUnc.size())
      return 3;
    if (std::fread(unc._pointsUnc.data(), 8, unc._pointsUnc.size(), f) != unc._pointsUnc.size())
      return 3;
    std::fclose(f);
    FILE *g = std::fopen(argv[3], "wb");
    put(g, unc.getCameraUncMatrix(4));
    put(g, unc.getPointUncMatrix(142));
    put(g, unc.getCamerasUncMatrices());
    put(g, unc.getPointsUncMatrices());
    put(g, unc.getCamerasUncRaw());
    put(g, unc.getPointsUncRaw());
    std::fclose(g);
    bool threw = false;
    try {
      unc.getPointsUncEigenValues();
    } catch (const std::runtime_error &e) {
      threw = true;
      std::printf("eigen: %s\n", e.what());
    }
    std::printf("threw %d\n", (int)threw);
    return 0;
  }
  cov::Options opt(1, 9, 0, 0);
  cov::Statistic stat = cov::Statistic();
  ceres::CRSMatrix J;
  std::vector<double> P;
  if (argc < 3 || !read_input(argv[2], opt, J, P)) return 3;
  cov::Uncertainty unc;
  if (mode == "fix") {
    ceres::CRSMatrix empty;  // no rows: the computation is refused after the choice
    getCovariances(opt, stat, empty, P.data(), unc);
    std::printf("fix %d %d %d rc %d\n", opt._pts2fix ? opt._pts2fix[0] : -1,
                opt._pts2fix ? opt._pts2fix[1] : -1, opt._pts2fix ? opt._pts2fix[2] : -1,
                getCovariancesLastError());
    return 0;
  }
  if (mode != "run" || argc < 4) return 2;
  getCovariances(opt, stat, J, P.data(), unc);
  const int rc = getCovariancesLastError();
  std::printf("rc %d fix %d %d %d stat %d %d %d lambda %.17g terms %zu\n", rc, opt._pts2fix[0],
              opt._pts2fix[1], opt._pts2fix[2], stat.fixedPts ? stat.fixedPts[0] : -1,
              stat.fixedPts ? stat.fixedPts[1] : -1, stat.fixedPts ? stat.fixedPts[2] : -1,
              stat.lambda, stat.cycle_change.size());
  if (rc != 0) return 1;
  FILE *g = std::fopen(argv[3], "wb");
  put(g, unc.getCamerasUncRaw());
  put(g, unc.getPointsUncRaw());
  put(g, unc.getCamerasUncEigenValues());
  put(g, unc.getCamerasCentersUncEigVec());
  put(g, unc.getPointsUncEigenValues());
  put(g, unc.getPointsUncEigVectors());
  std::fclose(g);
  return 0;
}
