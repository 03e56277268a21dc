// scene_io.cpp -- reconstruction front-end on the host: the COLMAP text
// reader (src/uncertaintyTE/ColmapIO.cpp:24-146, Image::qt2aaRC Image.cpp:
// 21-44) and the fixed-point choice (JacobianComposer::findPts2Fix,
// JacobianComposer.cpp:63-83), ordered as JacobianComposer::scene2Jacobian
// lays out its parameter blocks (JacobianComposer.cpp:123-157): images by id
// (those with no observation dropped), points by id, observations image by
// image with each image's sorted by point id.
//
// Where the reference's reader is broken this one follows the COLMAP format:
// camera parameters are read per model (the reference's inverted compare at
// ColmapIO.cpp:56-59 reads radial terms for the wrong models), an observation
// of an unknown point is an error (the reference creates a point at the
// origin), and an identity rotation gives aa = 0 (the reference divides by 0).
//
// The BAL text reader (BALProblem::BALProblem(filename), bal_problem.cpp:
// 111-172) keeps the file's order: cameras, points and observations as they
// stand.  Its nine camera numbers are taken as the reference's functor takes
// them (SnavelyReprojectionError, snavely_reprojection_error.h:309-347, the
// bundle type 0 of BuildProblem, matlabInterface.cpp:66-68): angle-axis,
// camera centre C, f, k1, k2, with unit weights since a BAL file carries none.
// Unlike the reference (which dies on a short file, fscanf in FscanfOrDie) a
// malformed or truncated file, or an index out of range, is an error code.
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/uncertainty_amd.h"

struct unc_scene {
  std::vector<double> cams;   // 9 per camera: aa, C, f, k1, k2
  std::vector<double> points; // 3 per point
  std::vector<int> obs_cam, obs_pt;
};

namespace {

struct ColmapCamera {
  double f = 0, k1 = 0, k2 = 0;
};

bool content_line(const std::string &l) { return l.size() > 1 && l[0] != '#'; }

int read_cameras(const std::string &path, std::map<long long, ColmapCamera> &out) {
  std::ifstream in(path);
  if (!in.good()) return UNC_ERR_BAD_INPUT;
  std::string line;
  while (std::getline(in, line)) {
    if (!content_line(line)) continue;
    std::istringstream ls(line);
    long long id;
    std::string model;
    double w, h;
    if (!(ls >> id >> model >> w >> h)) return UNC_ERR_BAD_INPUT;
    std::vector<double> p;
    double v;
    while (ls >> v) p.push_back(v);
    ColmapCamera c;
    // COLMAP parameter lists: SIMPLE_PINHOLE f cx cy; PINHOLE fx fy cx cy;
    // SIMPLE_RADIAL f cx cy k; RADIAL f cx cy k1 k2
    if (model == "SIMPLE_PINHOLE" && p.size() >= 3) {
      c.f = p[0];
    } else if (model == "PINHOLE" && p.size() >= 4) {
      if (p[0] != p[1]) return UNC_ERR_UNSUPPORTED;  // one focal length per camera
      c.f = p[0];
    } else if (model == "SIMPLE_RADIAL" && p.size() >= 4) {
      c.f = p[0];
      c.k1 = p[3];
    } else if (model == "RADIAL" && p.size() >= 5) {
      c.f = p[0];
      c.k1 = p[3];
      c.k2 = p[4];
    } else {
      return UNC_ERR_UNSUPPORTED;
    }
    out[id] = c;
  }
  return UNC_OK;
}

// Image::qt2aaRC: rotation (world to camera) from the unit quaternion, its
// angle-axis vector, and the centre C = -R^T t
void quat_to_aa_center(const double q[4], const double t[3], double aa[3], double C[3]) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  const double R[9] = {1 - 2 * y * y - 2 * z * z, 2 * x * y - 2 * w * z, 2 * z * x + 2 * w * y,
                       2 * x * y + 2 * w * z, 1 - 2 * x * x - 2 * z * z, 2 * y * z - 2 * w * x,
                       2 * z * x - 2 * w * y, 2 * y * z + 2 * w * x, 1 - 2 * x * x - 2 * y * y};
  const double s = std::sqrt(x * x + y * y + z * z);
  const double angle = 2 * std::atan2(s, w);
  for (int i = 0; i < 3; ++i) aa[i] = s > 0 ? angle * q[1 + i] / s : 0.0;
  for (int i = 0; i < 3; ++i) C[i] = -(R[i] * t[0] + R[3 + i] * t[1] + R[6 + i] * t[2]);
}

}  // namespace

extern "C" {

int unc_scene_read_colmap(const char *dir, unc_scene **out) {
  if (!dir || !out) return UNC_ERR_BAD_INPUT;
  *out = nullptr;
  const std::string d(dir);
  std::map<long long, ColmapCamera> cameras;
  int rc = read_cameras(d + "/cameras.txt", cameras);
  if (rc) return rc;
  // points3D.txt: POINT3D_ID X Y Z R G B ERROR TRACK[]
  std::map<long long, std::array<double, 3>> pts;
  {
    std::ifstream in(d + "/points3D.txt");
    if (!in.good()) return UNC_ERR_BAD_INPUT;
    std::string line;
    while (std::getline(in, line)) {
      if (!content_line(line)) continue;
      std::istringstream ls(line);
      long long id;
      std::array<double, 3> X;
      if (!(ls >> id >> X[0] >> X[1] >> X[2])) return UNC_ERR_BAD_INPUT;
      pts[id] = X;
    }
  }
  std::map<long long, int> pt_index;
  unc_scene *sc = new unc_scene();
  for (const auto &kv : pts) {
    pt_index[kv.first] = (int)pt_index.size();
    sc->points.insert(sc->points.end(), kv.second.begin(), kv.second.end());
  }
  // images.txt: IMAGE_ID QW QX QY QZ TX TY TZ CAMERA_ID NAME, then X Y POINT3D_ID ...
  struct Img {
    double aa[3], C[3];
    long long cam;
    std::vector<int> obs;  // point indices
  };
  std::map<long long, Img> images;
  {
    std::ifstream in(d + "/images.txt");
    if (!in.good()) {
      delete sc;
      return UNC_ERR_BAD_INPUT;
    }
    std::string line;
    while (std::getline(in, line)) {
      if (!content_line(line)) continue;
      std::istringstream ls(line);
      long long id, cam;
      double q[4], t[3];
      if (!(ls >> id >> q[0] >> q[1] >> q[2] >> q[3] >> t[0] >> t[1] >> t[2] >> cam) ||
          !cameras.count(cam)) {
        delete sc;
        return UNC_ERR_BAD_INPUT;
      }
      Img im;
      quat_to_aa_center(q, t, im.aa, im.C);
      im.cam = cam;
      std::string obs_line;
      std::getline(in, obs_line);
      std::istringstream os(obs_line);
      double x, y;
      long long pid;
      while (os >> x >> y >> pid) {
        if (pid == -1) continue;
        auto it = pt_index.find(pid);
        if (it == pt_index.end()) {
          delete sc;
          return UNC_ERR_BAD_INPUT;
        }
        im.obs.push_back(it->second);
      }
      // point ids and indices share their order, so sorting indices sorts by id
      std::stable_sort(im.obs.begin(), im.obs.end());
      images[id] = std::move(im);
    }
  }
  for (const auto &kv : images) {
    const Img &im = kv.second;
    if (im.obs.empty()) continue;
    const ColmapCamera &c = cameras[im.cam];
    const int ci = (int)(sc->cams.size() / 9);
    const double row[9] = {im.aa[0], im.aa[1], im.aa[2], im.C[0], im.C[1], im.C[2], c.f, c.k1, c.k2};
    sc->cams.insert(sc->cams.end(), row, row + 9);
    for (int p : im.obs) {
      sc->obs_cam.push_back(ci);
      sc->obs_pt.push_back(p);
    }
  }
  *out = sc;
  return UNC_OK;
}

int unc_scene_read_bal(const char *path, unc_scene **out) {
  if (!path || !out) return UNC_ERR_BAD_INPUT;
  *out = nullptr;
  // The whole file at once, then a forward scan with strtol / strtod: BAL
  // problems run to millions of observations and a stream extractor per value
  // is several times slower.
  std::string text;
  {
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return UNC_ERR_BAD_INPUT;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) return UNC_ERR_BAD_INPUT;
  }
  const char *p = text.c_str();
  bool ok = true;
  auto next_long = [&](long long &v) {
    char *e;
    v = std::strtoll(p, &e, 10);
    if (e == p) ok = false;
    p = e;
  };
  auto next_double = [&](double &v) {
    char *e;
    v = std::strtod(p, &e);
    if (e == p) ok = false;
    p = e;
  };
  // header: num_cameras num_points num_observations (bal_problem.cpp:121-123)
  long long nc, np, nobs;
  next_long(nc);
  next_long(np);
  next_long(nobs);
  if (!ok || nc < 1 || np < 1 || nobs < 1 || nc > INT32_MAX / 9 || np > INT32_MAX / 3 ||
      nobs > INT32_MAX / 2)
    return UNC_ERR_BAD_INPUT;
  unc_scene *sc = new unc_scene();
  sc->obs_cam.resize(nobs);
  sc->obs_pt.resize(nobs);
  // observations: camera_index point_index x y (bal_problem.cpp:133-139); the
  // Jacobian does not depend on x, y, which are parsed and dropped
  for (long long i = 0; i < nobs && ok; ++i) {
    long long c, q;
    double x, y;
    next_long(c);
    next_long(q);
    next_double(x);
    next_double(y);
    if (c < 0 || c >= nc || q < 0 || q >= np) ok = false;
    sc->obs_cam[i] = (int)c;
    sc->obs_pt[i] = (int)q;
  }
  // 9 per camera, then 3 per point (bal_problem.cpp:141-143)
  sc->cams.resize(9 * nc);
  sc->points.resize(3 * np);
  for (auto &v : sc->cams) {
    if (!ok) break;
    next_double(v);
  }
  for (auto &v : sc->points) {
    if (!ok) break;
    next_double(v);
  }
  if (!ok) {
    delete sc;
    return UNC_ERR_BAD_INPUT;
  }
  *out = sc;
  return UNC_OK;
}

int unc_scene_counts(const unc_scene *sc, int *num_cams, int *num_points, int *num_obs) {
  if (!sc) return UNC_ERR_BAD_INPUT;
  if (num_cams) *num_cams = (int)(sc->cams.size() / 9);
  if (num_points) *num_points = (int)(sc->points.size() / 3);
  if (num_obs) *num_obs = (int)sc->obs_cam.size();
  return UNC_OK;
}

int unc_scene_arrays(const unc_scene *sc, double *cams, double *points, int *obs_cam,
                     int *obs_pt) {
  if (!sc) return UNC_ERR_BAD_INPUT;
  if (cams) std::copy(sc->cams.begin(), sc->cams.end(), cams);
  if (points) std::copy(sc->points.begin(), sc->points.end(), points);
  if (obs_cam) std::copy(sc->obs_cam.begin(), sc->obs_cam.end(), obs_cam);
  if (obs_pt) std::copy(sc->obs_pt.begin(), sc->obs_pt.end(), obs_pt);
  return UNC_OK;
}

void unc_scene_destroy(unc_scene *sc) { delete sc; }

int unc_find_pts2fix(int num_points, const double *points3D, int out[3]) {
  // Seed of the search (the reference draws one from std::random_device) and
  // its number of accepted triples (JacobianComposer.cpp:70).
  constexpr unsigned kSeed = 5489u;
  constexpr int kTrials = 100000;
  if (num_points < 3 || !points3D || !out) return UNC_ERR_BAD_INPUT;
  auto dist = [points3D](int a, int b) {
    const double dx = points3D[3 * a] - points3D[3 * b], dy = points3D[3 * a + 1] - points3D[3 * b + 1],
                 dz = points3D[3 * a + 2] - points3D[3 * b + 2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
  };
  std::mt19937 gen(kSeed);
  std::uniform_real_distribution<> dis(0, num_points);
  double best = DBL_MIN;
  bool found = false;
  int f[3] = {-1, -1, -1};
  for (int i = 0; i < kTrials;) {
    const int p1 = (int)std::floor(dis(gen)), p2 = (int)std::floor(dis(gen)),
              p3 = (int)std::floor(dis(gen));
    if (p1 == p2 || p2 == p3 || p1 == p3) continue;
    const double d = dist(p1, p2) + dist(p1, p3) + dist(p2, p3);
    if (d > best) {
      best = d;
      f[0] = p1;
      f[1] = p2;
      f[2] = p3;
      found = true;
    }
    ++i;
  }
  if (!found) return UNC_ERR_NO_FIXED_POINTS;  // every triple coincides
  if (f[1] < f[0]) std::swap(f[0], f[1]);     // the reference sorts the first two only
  out[0] = f[0];
  out[1] = f[1];
  out[2] = f[2];
  return UNC_OK;
}

}  // extern "C"
