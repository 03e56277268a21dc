// pipeline.hip -- orchestration of the covariance pipeline on one MI355X and
// the C ABI of include/uncertainty_amd.h.
//
// Stage order mirrors computeCovariances (compute.cpp:458-540, TAYLOR branch):
//   decode + fixPts -> point lists -> J^T J blocks, scales, 3x3 inverses ->
//   Schur complement -> (Z + lambda I)^-1 (+ compensated Newton step) ->
//   Taylor chain -> camera output -> point back-propagation.
// Everything after the input upload runs on one stream (the Z^+ recursion forks
// one product per node onto a second); the one host round trip is at the end,
// reading the error word, the kept-observation count and the Taylor control
// block (TE path) or the conditioning bound (Z^+ path).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"

using namespace unc;

struct unc_context {
  int device = 0;
  hipStream_t stream = nullptr;
  char *ws = nullptr;
  size_t ws_bytes = 0;
  char *io = nullptr;
  size_t io_bytes = 0;
  int max_lds = 65536;
  hipEvent_t stage_ev[16];
  hipEvent_t gemm_ev[2 * 24];
  // the Z^+ recursion's second stream (default priority) and its fork / join
  // events, one pair per recursion depth; fork = false keeps the recursion on
  // the context's own stream (set by a request server whose slots would
  // otherwise drive more streams than the device has hardware queues)
  hipStream_t side = nullptr;
  hipEvent_t fork_ev[40], join_ev[40];
  bool fork = true;
  // page-locked staging of outputs bound for pageable host memory, and the
  // events that mark each staged chunk's arrival
  char *stage = nullptr;
  size_t stage_bytes = 0;
  hipEvent_t copy_ev[64];
  std::mutex mu;
};

namespace {

const char *kStageNames[12] = {"upload",   "decode",   "point_lists", "assemble",
                               "reserved", "schur",    "inverse",     "taylor",
                               "cam_out",  "point_out", "download",   "total"};

struct Sizes {
  int nc, cp, np, npc, nobs, N;
  size_t nn;
  bool te;     // the plain FP64 TE path (its Taylor chain needs the chain slots)
  bool mixed;  // the Z^+ path with CRT products (needs their residue planes)
  int moduli;  // their moduli
};

enum {
  B_OBS_CAM, B_OBS_PT, B_CAM_OFF, B_CAM_LIST, B_SORT_K0, B_SORT_K1, B_SORT_V,
  B_SORT_TMP, B_PT_CNT, B_PT_OFF, B_PT_LIST, B_PT_CURSOR, B_ERR, B_SCAN, B_CAM_PTS, B_PT_CAMS,
  B_GHASH, B_XD, B_US, B_D, B_VI, B_WS, B_PT_HAS, B_CTRL, B_P, B_TP, B_RP, B_NULL, B_DA, B_DB,
  B_DC, B_DD, B_DE, B_SPLIT, B_CRT, B_COUNT
};

struct Layout {
  size_t off[B_COUNT];
  size_t total;
};

constexpr int GJ_BLOCK = 64;
size_t zplus_work_doubles(int n);
size_t zplus_crt_bytes(int n, int nmod);
// Mixed precision: the Cholesky recursion's nodes of at least ZPLUS_CRT_MIN_N
// rows form their products on the int8 matrix cores, by default with 12
// moduli (exact products of operands scaled to 46-bit rows and columns); the
// smaller nodes stay FP64.
constexpr int ZPLUS_CRT_MOD = 12;
constexpr int ZPLUS_CRT_MIN_N = 2048;

// The Taylor chain forms its powers four at a time where one product alone
// would leave the device half idle (from N = 1024; below, products are small
// and launch-bound either way).
constexpr int TE_BATCH_MIN_N = 1024;
bool te_batched(int N) { return N >= TE_BATCH_MIN_N; }

// Taylor chain: the first product's K chunks, and the five power slots of the
// batched chain where it runs (the chunks alias the first three)
size_t make_sizes_split_bytes(const Sizes &z) {
  return z.te ? sizeof(double) * std::max(gemm_sym_split_work_doubles(z.N),
                                          te_batched(z.N) ? 5 * z.nn : (size_t)0)
              : 0;
}

Layout make_layout(const Sizes &z) {
  size_t sz[B_COUNT];
  sz[B_OBS_CAM] = sz[B_OBS_PT] = sz[B_PT_LIST] = sizeof(int) * (size_t)z.nobs;
  sz[B_CAM_OFF] = sizeof(int) * ((size_t)z.nc + 1);
  sz[B_CAM_LIST] = sz[B_SORT_K0] = sz[B_SORT_K1] = sz[B_SORT_V] = sizeof(int) * (size_t)z.nobs;
  sz[B_SORT_TMP] = camera_sort_temp_bytes(z.nobs, z.nc) + 256;
  sz[B_PT_CNT] = sz[B_PT_CURSOR] = sz[B_PT_HAS] = sizeof(int) * (size_t)z.npc;
  sz[B_PT_OFF] = sizeof(int) * ((size_t)z.npc + 1);
  sz[B_ERR] = 64;
  sz[B_SCAN] = sizeof(int) * ((size_t)std::max(z.npc, z.nc) / 1024 + 2);
  sz[B_CAM_PTS] = sz[B_PT_CAMS] = sizeof(int) * (size_t)z.nobs;
  sz[B_GHASH] = 4 * sizeof(int) * (size_t)z.nobs;  // global Schur tables of large cameras
  sz[B_US] = sz[B_XD] = sizeof(double) * (size_t)z.nc * z.cp * z.cp;
  sz[B_D] = sizeof(double) * ((size_t)z.N + 3 * (size_t)z.npc);
  sz[B_VI] = sizeof(double) * 9 * (size_t)z.npc;
  sz[B_WS] = sizeof(double) * 3 * (size_t)z.cp * z.nobs;
  sz[B_CTRL] = sizeof(TeCtrl);
  sz[B_P] = sizeof(double) * 3 * GJ_BLOCK * GJ_BLOCK;
  sz[B_TP] = sz[B_RP] = sizeof(double) * GJ_BLOCK * (size_t)z.N;
  sz[B_DA] = sz[B_DB] = sz[B_DC] = sz[B_DD] = sizeof(double) * z.nn;
  sz[B_DE] = sizeof(double) * z.nn;
  sz[B_SPLIT] = make_sizes_split_bytes(z);
  sz[B_NULL] = sizeof(double) * zplus_work_doubles(z.N);
  sz[B_CRT] = z.mixed ? zplus_crt_bytes(z.N, z.moduli) : 0;
  Layout L;
  size_t o = 0;
  for (int i = 0; i < B_COUNT; ++i) {
    L.off[i] = o;
    o += (sz[i] + 255) & ~(size_t)255;
  }
  L.total = o;
  return L;
}

int map_err(unsigned int e) {
  if (e & (ERRBIT_LAYOUT | ERRBIT_RANGE)) return UNC_ERR_BAD_INPUT;
  if (e & ERRBIT_DUPLICATE) return UNC_ERR_UNSUPPORTED;
  if (e & (ERRBIT_ZERO_COLUMN | ERRBIT_SINGULAR_V | ERRBIT_GJ_PIVOT)) return UNC_ERR_SINGULAR;
  return UNC_ERR_BAD_INPUT;
}

int fact32(int n) {  // compute.cpp:57-62, with the int32 wrap made explicit
  uint32_t f = 1u;
  for (int i = 2; i <= n; ++i) f *= (uint32_t)i;
  return (int)(int32_t)f;
}

int check_options(const unc_options *o, int nrows, int ncols, Sizes *z, int fix[3]) {
  if (!o) return UNC_ERR_BAD_INPUT;
  if (o->algorithm != UNC_ALG_TAYLOR_EXPANSION) return UNC_ERR_UNSUPPORTED;
  if (o->precision != UNC_PREC_FP64 && o->precision != UNC_PREC_MIXED &&
      o->precision != UNC_PREC_ZPLUS_FP64)
    return UNC_ERR_BAD_INPUT;
  if (o->cam_params < 1 || o->cam_params > 9) return UNC_ERR_UNSUPPORTED;
  if (o->num_cams <= 0 || o->num_points < 4) return UNC_ERR_BAD_INPUT;
  if (o->crt_moduli != 0 && (o->crt_moduli < CRT_MIN_MOD || o->crt_moduli > CRT_MAX_MOD))
    return UNC_ERR_BAD_INPUT;
  const long long N = (long long)o->num_cams * o->cam_params;
  if ((long long)ncols != N + 3LL * o->num_points) return UNC_ERR_BAD_INPUT;
  if (nrows <= 0 || (nrows & 1)) return UNC_ERR_BAD_INPUT;
  if ((long long)nrows * (o->cam_params + 3) > 0x7fffffffLL) return UNC_ERR_UNSUPPORTED;
  if (N > 200000) return UNC_ERR_UNSUPPORTED;
  if (o->pts2fix[0] < 0 || o->pts2fix[1] < 0 || o->pts2fix[2] < 0) return UNC_ERR_NO_FIXED_POINTS;
  for (int i = 0; i < 3; ++i)
    if (o->pts2fix[i] >= o->num_points) return UNC_ERR_BAD_INPUT;
  if (o->pts2fix[0] == o->pts2fix[1] || o->pts2fix[0] == o->pts2fix[2] ||
      o->pts2fix[1] == o->pts2fix[2])
    return UNC_ERR_BAD_INPUT;
  fix[0] = o->pts2fix[0];
  fix[1] = o->pts2fix[1];
  fix[2] = o->pts2fix[2];
  if (fix[1] < fix[0]) std::swap(fix[0], fix[1]);  // sort(pts, pts + 2), compute.cpp:285
  z->nc = o->num_cams;
  z->cp = o->cam_params;
  z->np = o->num_points;
  z->npc = o->num_points - 3;
  z->nobs = nrows / 2;
  z->N = (int)N;
  z->nn = (size_t)N * (size_t)N;
  z->te = o->precision == UNC_PREC_FP64;
  z->mixed = o->precision == UNC_PREC_MIXED;
  z->moduli = o->crt_moduli ? o->crt_moduli : ZPLUS_CRT_MOD;
  return UNC_OK;
}

// Grow a cached device buffer.  Stream-ordered (hipMallocAsync/hipFreeAsync on
// the context's stream), so a context that grows does not stall the other
// streams of the device the way hipFree's implicit synchronisation would.
int ensure(char **buf, size_t *have, size_t need, hipStream_t s) {
  if (*have >= need) return UNC_OK;
  if (*buf) (void)hipFreeAsync(*buf, s);
  *buf = nullptr;
  *have = 0;
  if (hipMallocAsync((void **)buf, need, s) != hipSuccess) {
    (void)hipGetLastError();
    *buf = nullptr;
    return UNC_ERR_OOM;
  }
  *have = need;
  return UNC_OK;
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
    (void)hipGetLastError();
    return 0.f;
  }
  return ms;
}

#define HC(expr)                                                                        \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      std::fprintf(stderr, "uncertainty_amd: %s failed (%s:%d): %s\n", #expr, __FILE__, \
                   __LINE__, hipGetErrorString(e_));                                    \
      return UNC_ERR_DEVICE;                                                            \
    }                                                                                   \
  } while (0)

// Blocked in-place Gauss-Jordan inverse of the SPD matrix in W (n x n) with
// 64-column panels K: P = W[K,K]^-1, T = W[:,K] P, R = W[K,:], W -= T R, then
// W[K,:] = P R, W[:,K] = -T, W[K,K] = P.  Two launches per panel: gj_panel_a
// (T, R and a copy of the next diagonal block W[K1,K1]) and gj_update (the
// trailing update outside the panel's rows and columns, those rows and
// columns, and the next panel's P = (W[K1,K1] - T[K1,:] R[:,K1])^-1, which
// needs nothing the trailing update writes).  P3: 3 x 64 x 64 doubles (two P
// buffers and the copy); T: n x 64; R: 64 x n.
int gj_inverse(int n, double *W, double *P3, double *T, double *R, unsigned int *err,
               hipStream_t s) {
  const int panels = (int)cdiv(n, GJ_BLOCK);
  double *Pb[2] = {P3, P3 + GJ_BLOCK * GJ_BLOCK};
  double *Dsrc = P3 + 2 * GJ_BLOCK * GJ_BLOCK;
  HC(spd_inverse64(std::min(GJ_BLOCK, n), W, n, Pb[0], err, s));
  for (int p = 0; p < panels; ++p) {
    const int k0 = p * GJ_BLOCK, kb = std::min(GJ_BLOCK, n - k0);
    const int k1 = k0 + kb, kb1 = p + 1 < panels ? std::min(GJ_BLOCK, n - k1) : 0;
    HC(gj_panel_a(n, k0, kb, W, Pb[p & 1], T, R, k1, kb1, Dsrc, s));
    HC(gj_update(n, k0, kb, W, Pb[p & 1], T, R, kb1, Dsrc, Pb[(p + 1) & 1], err, s));
  }
  return UNC_OK;
}

// Products of the Z^+ recursion: FP64 MFMA, or (mixed precision) the int8 CRT
// product with `nmod` moduli for every node of at least `min_n` rows.  The CRT
// area holds one product's scratch (operands and C residues at the top node's
// size) and, per recursion depth, the residues of that depth's L21, which serve
// three products of the node (A22 -= L21 L21^T as both operands, then L21 X11
// after the recursion into A22).
constexpr int REC_MAX_DEPTH = 40;
struct RecProducts {
  int nmod = 0;          // 0: FP64 throughout
  int min_n = 0;
  char *scratch = nullptr;
  size_t scratch_bytes = 0;
  char *level[REC_MAX_DEPTH] = {};
  // Second stream: a node of at least fork_min_n rows forms L21 X11 there while
  // the main stream descends into A22 (that product needs only L21 and X11); its
  // result goes to the depth's own region wside[depth] of the recursion scratch,
  // its CRT scratch to side_scratch.
  hipStream_t side = nullptr;
  const hipEvent_t *fork = nullptr, *join = nullptr;
  int fork_min_n = 0;
  double *wside[REC_MAX_DEPTH] = {};
  char *side_scratch = nullptr;
  size_t side_scratch_bytes = 0;
};

// Nodes of at least this many rows fork their L21 X11 onto the second stream
// (walk1500, 13500 rows, mixed: 38.8 ms for the recursion with 1024, 37.7 with
// 128; the FP64 path is level within noise).
constexpr int REC_FORK_MIN_N = 128;

// The first half of a recursion node of n rows: the multiple of 64 at or just
// above n / 2 (the second half, n - n1, can be one row larger when n is odd).
inline int rec_split(int n) { return ((n / 2 + 63) / 64) * 64; }

// CRT area of an n x n recursion (0 below ZPLUS_CRT_MIN_N); with `P` also its
// carve-up from `base`.  Walks the distinct node sizes depth by depth (a node's
// halves differ by less than 64, in either order): the scratch covers the
// largest node's products, each depth's L21 area its largest node's L21.
size_t zplus_crt_layout(int n, int nmod, char *base, RecProducts *P) {
  if (n < ZPLUS_CRT_MIN_N) return 0;
  auto r256 = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t scratch = 0, side = 0, lvl[REC_MAX_DEPTH] = {};
  std::vector<int> cur{n};
  int depths = 0;
  for (; depths < REC_MAX_DEPTH && !cur.empty(); ++depths) {
    std::vector<int> next;
    for (int m : cur) {
      if (m < ZPLUS_CRT_MIN_N) continue;
      const int m1 = rec_split(m), m2 = m - m1, h = std::max(m1, m2);
      scratch = std::max(scratch, r256(crt_gemm_work_bytes(h, h, h, nmod)));
      // the forked L21 X11: X11's residues and the product's
      if (m >= REC_FORK_MIN_N)
        side = std::max(side, r256(crt_operand_bytes(m1, m1, nmod)) +
                                  r256(crt_product_bytes(m2, m1, nmod)));
      lvl[depths] = std::max(lvl[depths], r256(crt_operand_bytes(m2, m1, nmod)));
      next.push_back(m1);
      next.push_back(m2);
    }
    std::sort(next.begin(), next.end());
    next.erase(std::unique(next.begin(), next.end()), next.end());
    cur.swap(next);
  }
  size_t o = scratch + side;
  if (P) {
    P->nmod = nmod;
    P->min_n = ZPLUS_CRT_MIN_N;
    P->scratch = base;
    P->scratch_bytes = scratch;
    P->side_scratch = base + scratch;
    P->side_scratch_bytes = side;
  }
  for (int d = 0; d < depths; ++d) {
    if (!lvl[d]) continue;
    if (P) P->level[d] = base + o;
    o += lvl[d];
  }
  return o;
}

// Regions of the recursion scratch W (n x n doubles): the main stream's
// products use W from its start (at most max n2 n1 doubles), and each depth's
// forked L21 X11 its own region after that.  Returns false when they would not
// fit (then nothing forks).
bool rec_fork_layout(int n, double *W, RecProducts *P) {
  size_t main = 0, wsz[REC_MAX_DEPTH] = {};
  std::vector<int> cur{n};
  int depths = 0;
  for (; depths < REC_MAX_DEPTH && !cur.empty(); ++depths) {
    std::vector<int> next;
    for (int m : cur) {
      if (m <= 64) continue;
      const int m1 = rec_split(m), m2 = m - m1;
      main = std::max(main, (size_t)m2 * m1);
      if (m >= REC_FORK_MIN_N) wsz[depths] = std::max(wsz[depths], (size_t)m2 * m1);
      next.push_back(m1);
      next.push_back(m2);
    }
    std::sort(next.begin(), next.end());
    next.erase(std::unique(next.begin(), next.end()), next.end());
    cur.swap(next);
  }
  size_t o = (main + 31) / 32 * 32;
  for (int d = 0; d < depths; ++d) {
    if (!wsz[d]) continue;
    P->wside[d] = W + o;
    o += (wsz[d] + 31) / 32 * 32;
  }
  if (o > (size_t)n * n) {
    for (auto &w : P->wside) w = nullptr;
    return false;
  }
  P->fork_min_n = REC_FORK_MIN_N;
  return true;
}

size_t zplus_crt_bytes(int n, int nmod) { return zplus_crt_layout(n, nmod, nullptr, nullptr); }

// Recursive Cholesky of the SPD block A[off.., off..] (n x n, ld N) with the
// inverse of its factor: X = L^-1 goes to the same block of X (its upper
// triangle untouched).  With A = [A11 .; A21 A22]:
//   (L11, X11) = chol_inv(A11);  L21 = A21 X11^T;  A22 -= L21 L21^T;
//   (L22, X22) = chol_inv(A22);  X21 = -X22 (L21 X11).
// L21 is held in X21's block, which stays free until the last product fills it
// (so A keeps L only in the diagonal blocks of the leaves and 128-row nodes;
// A21 keeps its input).  W is scratch of (n/2)^2 doubles at least.  Leaves are
// 64 x 64; a node of at most 128 rows runs in one launch (chol_inv_node128).
int chol_inv(int N, double *A, double *X, double *W, int off, int n, unsigned int *err,
             const RecProducts &P, int depth, hipStream_t s) {
  if (n <= 64) {
    HC(chol_inv_leaf(n, A + off + (size_t)off * N, N, X + off + (size_t)off * N, N, err, s));
    return UNC_OK;
  }
  if (n <= 128) {  // halves 64 and n - 64: the whole node in one workgroup
    HC(chol_inv_node128(n, A + off + (size_t)off * N, N, X + off + (size_t)off * N, N, err, s));
    return UNC_OK;
  }
  const int n1 = rec_split(n), n2 = n - n1;
  int rc = chol_inv(N, A, X, W, off, n1, err, P, depth + 1, s);
  if (rc) return rc;
  const double *A21 = A + (off + n1) + (size_t)off * N;
  double *A22 = A + (off + n1) + (size_t)(off + n1) * N;
  const double *X11 = X + off + (size_t)off * N;
  const double *X22 = X + (off + n1) + (size_t)(off + n1) * N;
  double *X21 = X + (off + n1) + (size_t)off * N;
  double *L21p = X21;  // L21 until X21 is formed
  const bool crt = P.nmod > 0 && n >= P.min_n && depth < REC_MAX_DEPTH && P.level[depth];
  const bool fork = P.side && n >= P.fork_min_n && depth < REC_MAX_DEPTH && P.wside[depth] &&
                    (!crt || P.side_scratch);
  const int nm = P.nmod;
  const size_t sb = P.scratch_bytes;
  // L21 = A21 X11^T (into X21's block)
  if (crt)
    HC(gemm_crt(n2, n1, n1, 1.0, A21, N, X11, N, true, 0.0, L21p, N, GEMM_BT_LOWER, nm, P.scratch,
                sb, s));
  else
    HC(gemm_f64_ex(n2, n1, n1, 1.0, A21, N, X11, N, true, 0.0, L21p, N, GEMM_BT_LOWER, s));
  CrtOperand L21;
  if (crt) HC(crt_split(n2, n1, L21p, N, true, 0, nm, P.level[depth], &L21, s));
  // W3 = L21 X11, here or on the second stream
  double *W3 = fork ? P.wside[depth] : W;
  auto l21_x11 = [&](hipStream_t st, char *cs, size_t csb) -> hipError_t {
    if (!crt)
      return gemm_f64_ex(n2, n1, n1, 1.0, L21p, N, X11, N, false, 0.0, W3, n2, GEMM_B_LOWER, st);
    CrtOperand Xc;  // the columns of X11 (lower) as the second operand
    const size_t bx = crt_operand_bytes(n1, n1, nm);
    hipError_t e = crt_split(n1, n1, X11, N, false, 2, nm, cs, &Xc, st);
    if (e != hipSuccess) return e;
    return crt_product(1.0, L21, Xc, 0.0, W3, n2, GEMM_B_LOWER, cs + bx, csb - bx, st);
  };
  if (fork) {
    HC(hipEventRecord(P.fork[depth], s));
    HC(hipStreamWaitEvent(P.side, P.fork[depth], 0));
    HC(l21_x11(P.side, P.side_scratch, P.side_scratch_bytes));
    HC(hipEventRecord(P.join[depth], P.side));
  }
  // A22 -= L21 L21^T, then the recursion into A22
  if (crt)
    HC(crt_product(-1.0, L21, L21, 1.0, A22, N, GEMM_C_LOWER, P.scratch, sb, s));
  else
    HC(gemm_f64_ex(n2, n2, n1, -1.0, L21p, N, L21p, N, true, 1.0, A22, N, GEMM_C_LOWER, s));
  rc = chol_inv(N, A, X, W, off + n1, n2, err, P, depth + 1, s);
  if (rc) return rc;
  if (fork)
    HC(hipStreamWaitEvent(s, P.join[depth], 0));
  else
    HC(l21_x11(s, P.scratch, sb));
  // X21 = -X22 W3
  if (crt)
    HC(gemm_crt(n2, n1, n2, -1.0, X22, N, W3, n2, false, 0.0, X21, N, GEMM_A_LOWER, nm, P.scratch,
                sb, s));
  else
    HC(gemm_f64_ex(n2, n1, n2, -1.0, X22, N, W3, n2, false, 0.0, X21, N, GEMM_A_LOWER, s));
  return UNC_OK;
}

// Work area of the Z^+ path inside B_NULL (doubles, n = N):
//   U n x 14 ([N | N0]), P n x 14, MU n x 14, Y n x 7, R n x 7, Q n x 7,
//   skinny partials, gram partials, small matrices.
constexpr int GAUGE = 7;
size_t zplus_work_doubles(int n) {
  return (size_t)(3 * 2 * GAUGE + 3 * GAUGE) * n + skinny_part_doubles(n) + 64 * 256 + 4096;
}

// Gauge-fixed Z^+: Z in dA on entry (kept), dB/dC/dD are N x N scratch.  Leaves
// the cameras' own blocks of Z^+ in Xd (row-major cp x cp per camera).
int zplus(int N, int newton_steps, double *dA, double *dB, double *dC, double *dD, double *work,
          const RecProducts &P, const Structure &S, double *Xd, unsigned int *err,
          hipStream_t s) {
  const size_t nn = (size_t)N * N;
  const int k = GAUGE;
  double *U = work, *Pm = U + 2 * k * (size_t)N, *MU = Pm + 2 * k * (size_t)N;
  double *Y = MU + 2 * k * (size_t)N, *R = Y + k * (size_t)N, *Q = R + k * (size_t)N;
  double *part = Q + k * (size_t)N;
  double *gpart = part + skinny_part_doubles(N);
  double *G = gpart + 64 * 256, *T = G + 256, *Sw = T + 256, *Th = Sw + 256;
  double *Nk = U, *N0 = U + k * (size_t)N;
  // N0: fixed pseudo-random orthonormal block
  HC(fill_hash((size_t)k * N, N0, 0x5eedull, s));
  HC(gram(N, k, N0, k, N0, gpart, G, s));
  HC(small_orth(k, G, T, err, s));
  HC(small_apply(N, k, k, 0.0, nullptr, 1.0, N0, T, N0, s));
  // Z~ = Z + N0 N0^T, Cholesky with the inverse factor X (upper part zero)
  HC(rank_copy(N, k, dA, N0, dB, s));
  HC(hipMemsetAsync(dC, 0, sizeof(double) * nn, s));
  int rc = chol_inv(N, dB, dC, dD, 0, N, err, P, 0, s);
  if (rc) return rc;
  const double *X = dC;
  // N = orth(Z~^-1 N0) = orth(X^T X N0)
  HC(skinny_nn(N, k, X, N, N0, R, part, true, s));
  HC(skinny_tn(N, k, X, N, R, Y, true, s));
  HC(gram(N, k, Y, k, Y, gpart, G, s));
  HC(small_orth(k, G, T, err, s));
  HC(small_apply(N, k, k, 0.0, nullptr, 1.0, Y, T, Nk, s));
  // Newton steps on the invariant subspace: R = Z N - N (N^T Z N), projected;
  // correction P X^T X P R
  for (int it = 0; it < newton_steps; ++it) {
    HC(skinny_tn(N, k, dA, N, Nk, R, false, s));            // Z N (Z symmetric)
    HC(gram(N, k, Nk, k, R, gpart, Th, s));                  // N^T Z N
    HC(small_apply(N, k, k, 1.0, R, -1.0, Nk, Th, R, s));    // R = ZN - N Th (already _|_ N)
    HC(gram(N, k, Nk, k, R, gpart, G, s));
    HC(small_apply(N, k, k, 1.0, R, -1.0, Nk, G, R, s));     // re-project
    HC(skinny_nn(N, k, X, N, R, Q, part, true, s));
    HC(skinny_tn(N, k, X, N, Q, Y, true, s));                // M P R
    HC(gram(N, k, Nk, k, Y, gpart, G, s));
    HC(small_apply(N, k, k, 1.0, Y, -1.0, Nk, G, Y, s));     // P M P R
    HC(small_apply(N, k, k, 1.0, Nk, -1.0, Y, nullptr, Y, s));  // N - correction
    HC(gram(N, k, Y, k, Y, gpart, G, s));
    HC(small_orth(k, G, T, err, s));
    HC(small_apply(N, k, k, 0.0, nullptr, 1.0, Y, T, Nk, s));
  }
  // Woodbury: U = [N N0], P = X U, MU = X^T P, S = (C + P^T P)^-1
  HC(skinny_nn(N, 2 * k, X, N, U, Pm, part, true, s));
  HC(skinny_tn(N, 2 * k, X, N, Pm, MU, true, s));
  HC(gram(N, 2 * k, Pm, 2 * k, Pm, gpart, G, s));
  HC(small_woodbury(2 * k, G, Sw, err, s));
  HC(cam_gram(S, X, MU, Sw, Nk, Xd, s));
  return UNC_OK;
}

// ---- SVD algorithm modes (compute.cpp:483-495, svdInverse :185-249), k_svd.hip
constexpr long long kSvdMaxParams = 8192;          // dense n x n eigen-decomposition
constexpr long long kSvdMaxJtBytes = 8LL << 30;    // J^T held densely

struct SvdLayout {
  size_t jt, a0, a1, v, lam, w, zc, err, flag, total;
};

SvdLayout svd_layout(long long npar, long long nrows, int nc, int cp) {
  auto r256 = [](size_t b) { return ((b + 255) / 256) * 256; };
  SvdLayout L;
  size_t o = 0;
  L.jt = o; o += r256(8 * (size_t)npar * (size_t)nrows);
  L.a0 = o; o += r256(8 * (size_t)npar * (size_t)npar);
  L.a1 = o; o += r256(8 * (size_t)npar * (size_t)npar);
  L.v = o; o += r256(8 * (size_t)npar * (size_t)npar);
  L.lam = o; o += r256(8 * (size_t)npar);
  L.w = o; o += r256(8 * (size_t)npar);
  L.zc = o; o += r256(8 * (size_t)nc * cp * cp);
  L.err = o; o += 256;
  L.flag = o; o += 256;
  L.total = o;
  return L;
}

int svd_run(unc_context *ctx, const unc_options *o, int nrows, int ncols, const int *d_rows,
            const int *d_cols, const double *d_vals, double *d_cam, double *d_pt, unc_stats *st,
            hipStream_t s) {
  const int cp = o->cam_params, nc = o->num_cams, np = o->num_points;
  if (cp < 1 || cp > 9) return UNC_ERR_UNSUPPORTED;
  if (nc <= 0 || np <= 0 || nrows <= 0 || (nrows & 1)) return UNC_ERR_BAD_INPUT;
  const long long N = (long long)nc * cp, npar = N + 3LL * np;
  if ((long long)ncols != npar) return UNC_ERR_BAD_INPUT;
  if (npar > kSvdMaxParams || 8LL * npar * nrows > kSvdMaxJtBytes) return UNC_ERR_UNSUPPORTED;
  const int n = (int)npar, m = n + (n & 1);
  hipEvent_t *E = ctx->stage_ev;
  HC(hipEventRecord(E[1], s));
  const SvdLayout L = svd_layout(npar, nrows, nc, cp);
  int rc = ensure(&ctx->ws, &ctx->ws_bytes, L.total, s);
  if (rc) return rc;
  char *base = ctx->ws;
  double *Jt = (double *)(base + L.jt), *A0 = (double *)(base + L.a0);
  double *A1 = (double *)(base + L.a1), *V = (double *)(base + L.v);
  double *lam = (double *)(base + L.lam), *wt = (double *)(base + L.w);
  double *Zc = (double *)(base + L.zc);
  unsigned int *err = (unsigned int *)(base + L.err);
  int *flag = (int *)(base + L.flag);
  HC(hipMemsetAsync(err, 0, 256, s));
  HC(hipMemsetAsync(Jt, 0, 8 * (size_t)npar * (size_t)nrows, s));
  HC(svd_jt_scatter(nrows, n, d_rows, d_cols, d_vals, Jt, err, s));
  // M = J^T J (unscaled, every camera and point), svdInverse's Z (compute.cpp:186)
  HC(gemm_f64_ex(n, n, nrows, 1.0, Jt, n, Jt, n, true, 0.0, A0, n, 0, s));
  HC(svd_symmetrize(n, A0, s));
  HC(svd_identity(n, V, s));
  HC(hipEventRecord(E[5], s));
  double *Ao = A0, *An = A1;
  int sweeps = 0;
  for (; sweeps < 40; ++sweeps) {
    HC(hipMemsetAsync(flag, 0, sizeof(int), s));
    for (int r = 0; r < m - 1; ++r) {
      HC(svd_jacobi_round(n, m, r, Ao, An, V, flag, s));
      std::swap(Ao, An);
    }
    int any = 0;
    HC(hipMemcpyAsync(&any, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    if (!any) break;
  }
  if (sweeps >= 40) return UNC_ERR_NUMERIC;
  // singular values = |eigenvalues|; the kept ones (svdInverse :222-240) get 1/lambda
  HC(svd_diag(n, Ao, lam, s));
  std::vector<double> hl(n), hw(n, 0.0);
  HC(hipMemcpyAsync(hl.data(), lam, 8 * (size_t)n, hipMemcpyDeviceToHost, s));
  HC(hipStreamSynchronize(s));
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return std::fabs(hl[a]) > std::fabs(hl[b]); });
  if (o->svd_remove_n != -1) {
    const long long keep = std::max(0LL, std::min((long long)n, (long long)n - o->svd_remove_n));
    for (long long i = 0; i < keep; ++i) hw[order[i]] = 1.0 / hl[order[i]];
  } else {
    const double eps = o->epsilon < 0 ? 1e-10 : o->epsilon;
    for (int i = 0; i < n; ++i)
      if (std::fabs(hl[i]) > eps) hw[i] = 1.0 / hl[i];
  }
  HC(hipMemcpyAsync(wt, hw.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s));
  HC(hipEventRecord(E[7], s));
  HC(svd_cam_blocks(nc, cp, n, V, wt, Zc, d_cam, s));
  HC(hipEventRecord(E[8], s));
  HC(svd_points(nrows / 2, cp, nc, np, d_cols, d_vals, Zc, d_pt, err, s));
  HC(hipEventRecord(E[10], s));
  unsigned int herr = 0;
  HC(hipMemcpyAsync(&herr, err, sizeof(herr), hipMemcpyDeviceToHost, s));
  HC(hipStreamSynchronize(s));
  if (herr) return map_err(herr);
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->fixed_pts[0] = st->fixed_pts[1] = st->fixed_pts[2] = -1;
    st->num_obs_kept = nrows / 2;
    st->stage_ms[5] = ev_ms(E[1], E[5]);    // J^T J
    st->stage_ms[6] = ev_ms(E[5], E[7]);    // eigen-decomposition + pseudo-inverse weights
    st->stage_ms[8] = ev_ms(E[7], E[8]);
    st->stage_ms[9] = ev_ms(E[8], E[10]);
    st->stage_ms[11] = ev_ms(E[1], E[10]);
    st->time_create_j = st->stage_ms[5] * 1e-3;
    st->time_inv_z = st->stage_ms[6] * 1e-3;
    st->time_pts_unc = (st->stage_ms[8] + st->stage_ms[9]) * 1e-3;
    st->time_all = st->stage_ms[11] * 1e-3;
  }
  return UNC_OK;
}

// Workspace of a run: the context's cached buffer (grown on demand), or a
// caller-provided region of at least the layout's size (no allocation).
struct Workspace {
  char *base = nullptr;
  size_t bytes = 0;
};

int run(unc_context *ctx, const unc_options *o, int nrows, int ncols, const int *d_rows,
        const int *d_cols, const double *d_vals, double *d_cam, double *d_pt, unc_stats *st,
        unc_debug *dbg, hipStream_t s, bool host_api_upload_timed, const Workspace *fixed_ws,
        const int *d_ids = nullptr) {
  if (o && (o->algorithm == UNC_ALG_SVD_QR_ITERATION ||
            o->algorithm == UNC_ALG_SVD_DEVIDE_AND_CONQUER)) {
    if (fixed_ws) return UNC_ERR_UNSUPPORTED;
    return svd_run(ctx, o, nrows, ncols, d_rows, d_cols, d_vals, d_cam, d_pt, st, s);
  }
  Sizes z;
  int fix[3];
  int rc = check_options(o, nrows, ncols, &z, fix);
  if (rc) return rc;
  const bool refine = o->refine != 0;
  const Layout L = make_layout(z);
  char *w = nullptr;
  if (fixed_ws) {
    if (fixed_ws->bytes < L.total) return UNC_ERR_OOM;
    w = fixed_ws->base;
  } else {
    rc = ensure(&ctx->ws, &ctx->ws_bytes, L.total, s);
    if (rc) return rc;
    w = ctx->ws;
  }
  auto ptr = [&](int b) { return (void *)(w + L.off[b]); };

  Structure S;
  S.cp = z.cp; S.nc = z.nc; S.np = z.np; S.npc = z.npc; S.nobs = z.nobs; S.N = z.N;
  S.fix[0] = fix[0]; S.fix[1] = fix[1]; S.fix[2] = fix[2];
  S.rows = d_rows; S.cols = d_cols; S.vals = d_vals; S.ids = d_ids;
  S.obs_cam = (int *)ptr(B_OBS_CAM); S.obs_pt = (int *)ptr(B_OBS_PT);
  S.cam_off = (int *)ptr(B_CAM_OFF);
  S.cam_list = (int *)ptr(B_CAM_LIST);
  S.pt_cnt = (int *)ptr(B_PT_CNT); S.pt_off = (int *)ptr(B_PT_OFF);
  S.pt_list = (int *)ptr(B_PT_LIST); S.pt_cursor = (int *)ptr(B_PT_CURSOR);
  S.err = (unsigned int *)ptr(B_ERR);
  Blocks Bk;
  Bk.Us = (double *)ptr(B_US); Bk.d = (double *)ptr(B_D); Bk.Vi = (double *)ptr(B_VI);
  Bk.Ws = (double *)ptr(B_WS); Bk.pt_has = (int *)ptr(B_PT_HAS);
  TeCtrl *ctrl = (TeCtrl *)ptr(B_CTRL);
  double *dA = (double *)ptr(B_DA), *dB = (double *)ptr(B_DB), *dC = (double *)ptr(B_DC);
  double *dD = (double *)ptr(B_DD), *dE = (double *)ptr(B_DE);
  const int N = z.N;
  const size_t nn = z.nn;
  hipEvent_t *E = ctx->stage_ev;
  if (!host_api_upload_timed) HC(hipEventRecord(E[0], s));
  HC(hipEventRecord(E[1], s));

  // ---- decode + fixPts.  Invalid observations are flagged in the error word
  // and treated as dropped by every later kernel; the word is read once, with
  // the results, so no stage waits for the host.
  HC(launch_decode(S, s));
  HC(hipEventRecord(E[2], s));
  HC(launch_point_lists(S, (int *)ptr(B_SCAN), s));
  int *cam_pts = (int *)ptr(B_CAM_PTS), *pt_cams = (int *)ptr(B_PT_CAMS);
  HC(sort_obs_by_camera(S, pt_cams, cam_pts, (int *)ptr(B_SORT_K0), (int *)ptr(B_SORT_K1),
                        (int *)ptr(B_SORT_V), ptr(B_SORT_TMP),
                        camera_sort_temp_bytes(z.nobs, z.nc) + 256, s));
  HC(hipEventRecord(E[3], s));
  // ---- (a) + (b): camera blocks, point blocks + 3x3 inverses, W_s
  HC(launch_cam_blocks(S, Bk, s));
  HC(hipEventRecord(E[11], s));
  HC(launch_point_blocks(S, Bk, s));
  HC(hipEventRecord(E[12], s));
  HC(launch_obs_blocks(S, Bk, s));
  HC(hipEventRecord(E[4], s));
  HC(hipEventRecord(E[5], s));
  // ---- (c) Schur complement into dA; blocks of cameras without a shared
  // point stay zero
  HC(fill_zero(dA, nn, s));
  HC(launch_schur(S, Bk, dA, cam_pts, pt_cams, (int *)ptr(B_GHASH), s));
  HC(hipEventRecord(E[6], s));
  if (dbg && dbg->Z) HC(hipMemcpyAsync(dbg->Z, dA, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
  unsigned int herr = 0;
  int kept = 0;
  auto fill_stats = [&](double lambda, int nterms, const double *chg) {
    std::memset(st, 0, sizeof(*st));
    st->lambda = lambda;
    st->fixed_pts[0] = fix[0]; st->fixed_pts[1] = fix[1]; st->fixed_pts[2] = fix[2];
    st->num_terms = nterms;
    for (int i = 0; i < nterms && i < 20; ++i) st->cycle_change[i] = chg[i];
    st->num_obs_kept = kept;
    st->stage_ms[1] = ev_ms(E[1], E[2]);
    st->stage_ms[2] = ev_ms(E[2], E[3]);
    st->stage_ms[3] = ev_ms(E[3], E[4]);
    st->stage_ms[5] = ev_ms(E[5], E[6]);
    st->stage_ms[6] = ev_ms(E[6], E[7]);
    st->stage_ms[7] = ev_ms(E[7], E[8]);
    st->stage_ms[8] = ev_ms(E[8], E[9]);
    st->stage_ms[9] = ev_ms(E[9], E[10]);
    st->stage_ms[11] = ev_ms(E[host_api_upload_timed ? 0 : 1], E[10]);
    if (host_api_upload_timed) st->stage_ms[0] = ev_ms(E[0], E[1]);
    st->time_create_j = st->stage_ms[0] * 1e-3;
    st->time_fix_j = (st->stage_ms[1] + st->stage_ms[2]) * 1e-3;
    // column norms are the diagonals of the assembled blocks and the blocks are
    // written split: J^T J = camera blocks + W_s, V^-1 = point blocks
    st->time_norm_j = 0.0;
    st->time_multiply_jj = (ev_ms(E[3], E[11]) + ev_ms(E[12], E[4])) * 1e-3;
    st->time_split_jj = 0.0;
    st->time_inv_v = ev_ms(E[11], E[12]) * 1e-3;
    st->time_compose_z = st->stage_ms[5] * 1e-3;
    st->time_inv_z = st->stage_ms[6] * 1e-3;
    st->time_te = st->stage_ms[7] * 1e-3;
    st->time_pts_unc = (st->stage_ms[8] + st->stage_ms[9]) * 1e-3;
    st->time_all = st->stage_ms[11] * 1e-3;
  };
  auto copy_debug_blocks = [&]() -> int {
    if (!dbg) return UNC_OK;
    if (dbg->U_s) HC(hipMemcpyAsync(dbg->U_s, Bk.Us, sizeof(double) * z.nc * z.cp * z.cp, hipMemcpyDeviceToHost, s));
    if (dbg->Vinv_s) HC(hipMemcpyAsync(dbg->Vinv_s, Bk.Vi, sizeof(double) * 9 * (size_t)z.npc, hipMemcpyDeviceToHost, s));
    if (dbg->d) HC(hipMemcpyAsync(dbg->d, Bk.d, sizeof(double) * ((size_t)N + 3 * (size_t)z.npc), hipMemcpyDeviceToHost, s));
    return UNC_OK;
  };
  if (o->precision != UNC_PREC_FP64) {
    // gauge-fixed Z^+ path
    const int steps = o->refine >= 0 ? o->refine : 1;
    double *Xd = (double *)ptr(B_XD);
    RecProducts P;
    if (z.mixed) zplus_crt_layout(N, z.moduli, (char *)ptr(B_CRT), &P);
    if (ctx->fork && rec_fork_layout(N, dD, &P)) {
      P.side = ctx->side;
      P.fork = ctx->fork_ev;
      P.join = ctx->join_ev;
    }
    rc = zplus(N, steps, dA, dB, dC, dD, (double *)ptr(B_NULL), P, S, Xd, S.err, s);
    if (rc) return rc;
    double *d_condlb = (double *)((char *)ptr(B_ERR) + 8);
    HC(diag_max_sq(N, dC, N, d_condlb, s));
    HC(hipEventRecord(E[7], s));
    HC(hipEventRecord(E[8], s));
    HC(cam_output_cov(S, Xd, Bk.d, d_cam, s));
    HC(hipEventRecord(E[9], s));
    HC(hipMemsetAsync(d_pt, 0, sizeof(double) * 6 * (size_t)z.np, s));
    HC(launch_point_output(S, Bk, Xd, d_pt, z.np, s));
    HC(hipEventRecord(E[10], s));
    rc = copy_debug_blocks();
    if (rc) return rc;
    // X: the inverse Cholesky factor L^-1 of Z + N0 N0^T (lower triangle)
    if (dbg && dbg->X) HC(hipMemcpyAsync(dbg->X, dC, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
    HC(hipMemcpyAsync(&herr, S.err, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
    HC(hipMemcpyAsync(&kept, S.pt_off + z.npc, sizeof(int), hipMemcpyDeviceToHost, s));
    double condlb = 0.0;
    HC(hipMemcpyAsync(&condlb, d_condlb, sizeof(double), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    if (herr) return map_err(herr);
    if (st) {
      fill_stats(0.0, 0, nullptr);
      st->zplus_moduli = P.nmod;
      st->zplus_cond_lb = condlb;  // X = L^-1: 1 / min L_ii^2 = max X_ii^2
    }
    return UNC_OK;
  }
  // ---- (d) A = Z + lambda I ; X0 = A^-1
  const double lambda = (o->lambda == -1) ? pow(10, -1.2653 * log10((double)N) - 2.9415) : o->lambda;
  HC(add_diag(N, dA, N, lambda, s));
  HC(hipMemcpyAsync(dB, dA, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
  rc = gj_inverse(N, dB, (double *)ptr(B_P), (double *)ptr(B_TP), (double *)ptr(B_RP), S.err, s);
  if (rc) return rc;
  hipEvent_t *GE = ctx->gemm_ev;
  int gemm_pairs = 0;
  double *X0 = dB;
  // Taylor-chain K chunks (the symmetric products' scratch)
  double *sym_work = (double *)ptr(B_SPLIT);
  const size_t sym_work_doubles = gemm_sym_split_work_doubles(N);
  if (refine) {
    // X0 from Gauss-Jordan is symmetric only to about cond * eps; symmetrised
    // first, X0 R = X0 - X0 A X0 is symmetric too (up to rounding), so the
    // lower-tile product serves.  (Left unsymmetrised, X0 R carries an
    // antisymmetric part of the size of X0's asymmetry, which mirroring would
    // fold into the result.)
    HC(svd_symmetrize(N, dB, s));
    if (te_batched(N)) {
      // on the int8 matrix cores; its residue planes take D, E and the chain
      // slots, all free until X1 is formed
      const size_t span = L.off[B_SPLIT] + make_sizes_split_bytes(z) - L.off[B_DD];
      HC(residual_crt(N, dA, dB, dC, ptr(B_DD), span, s));
    } else {
      HC(residual_ex(N, dA, dB, dC, dE, s));  // dE is free until X1 is formed
    }
    HC(hipEventRecord(GE[2 * gemm_pairs], s));
    HC(gemm_f64_sym(N, 1.0, dB, dC, dD, sym_work, sym_work_doubles, nullptr, nullptr, s));
    HC(hipEventRecord(GE[2 * gemm_pairs + 1], s));
    ++gemm_pairs;
    // X0 was symmetrised and the lower-tile product mirrors D: the plain sum
    // is then already symmetric
    if (sym_work_doubles > 0) HC(add_symmetric(N, dB, dD, dE, s));
    else HC(sym_update(N, dB, dD, dE, s));
    X0 = dE;
  }
  HC(hipEventRecord(E[7], s));
  if (dbg && dbg->X0) HC(hipMemcpyAsync(dbg->X0, X0, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
  // ---- Taylor chain (compute.cpp:142-160); X in dA (A is dead)
  double *X = dA;
  HC(hipMemcpyAsync(X, X0, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
  HC(te_init(ctrl, s));
  // chain launches: events and the first term each produces (a launch runs iff
  // every earlier term was accepted), products per launch
  int lfirst[24], lcount[24];
  const int first_chain_pair = gemm_pairs;
  auto chain_launch_begin = [&](int first_term, int products) {
    lfirst[gemm_pairs] = first_term;
    lcount[gemm_pairs] = products;
    return hipEventRecord(GE[2 * gemm_pairs], s);
  };
  auto chain_launch_end = [&]() {
    const hipError_t e = hipEventRecord(GE[2 * gemm_pairs + 1], s);
    ++gemm_pairs;
    return e;
  };
  auto coef = [&](int i) { return pow(lambda, i) / (double)fact32(i - 1); };
  // T_1 = X0^2; X0^(i+1) = T_i is symmetric: only lower tiles are computed
  double *T1 = dC;
  HC(chain_launch_begin(1, 1));
  HC(gemm_f64_sym(N, 1.0, X0, X0, T1, sym_work, sym_work_doubles, &ctrl->maxbits[1], nullptr, s));
  HC(chain_launch_end());
  const bool batched = te_batched(N);
  if (!batched) {
    double *Tc = T1, *Tn = dD;
    for (int i = 1; i < 20; ++i) {
      HC(te_step(N, ctrl, i, coef(i), X, Tc, s));
      HC(chain_launch_begin(i + 1, 1));
      HC(gemm_f64_sym(N, 1.0, Tc, X0, Tn, sym_work, sym_work_doubles, &ctrl->maxbits[i + 1],
                      &ctrl->active, s));
      HC(chain_launch_end());
      std::swap(Tc, Tn);
    }
  } else {
    // powers X0^1..4 stay resident; X0^(4j+1..4j+4) = X0^(4j) X0^(1..4) fill
    // five rotating slots (the base X0^(4j) of a batch is the last product of
    // the one before, so four fresh slots are always free)
    double *pw[4] = {X0, T1, dD, X0 == dB ? dE : dB};
    double *slot = sym_work;
    HC(te_step(N, ctrl, 1, coef(1), X, T1, s));
    {
      const double *Bq[4] = {pw[0], pw[1], pw[0], pw[0]};
      double *Cq[4] = {pw[2], pw[3], pw[2], pw[2]};
      HC(chain_launch_begin(2, 2));
      HC(gemm_f64_sym_batch(N, 2, T1, Bq, Cq, &ctrl->maxbits[2], &ctrl->active, s));
      HC(chain_launch_end());
      const double kq[4] = {coef(2), coef(3), 0.0, 0.0};
      const double *Tq[4] = {pw[2], pw[3], pw[2], pw[2]};
      HC(te_steps(N, ctrl, 2, 2, kq, X, Tq, s));
    }
    const double *base = pw[3];
    for (int j = 1; 4 * j < 20; ++j) {
      const int i0 = 4 * j, m = std::min(4, 20 - i0);
      double *Cq[4];
      const double *Tq[4];
      double kq[4];
      for (int q = 0; q < 4; ++q) {
        Cq[q] = slot + (size_t)((4 * (j - 1) + std::min(q, m - 1)) % 5) * nn;
        Tq[q] = Cq[q];
        kq[q] = q < m ? coef(i0 + q) : 0.0;
      }
      const double *Bq[4] = {pw[0], pw[1], pw[2], pw[3]};
      HC(chain_launch_begin(i0, m));
      HC(gemm_f64_sym_batch(N, m, base, Bq, Cq, &ctrl->maxbits[i0], &ctrl->active, s));
      HC(chain_launch_end());
      HC(te_steps(N, ctrl, i0, m, kq, X, Tq, s));
      base = Cq[m - 1];
    }
  }
  HC(hipEventRecord(E[8], s));
  if (dbg && dbg->X) HC(hipMemcpyAsync(dbg->X, X, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
  // ---- outputs
  HC(launch_cam_output(S, Bk, X, d_cam, s));
  HC(hipEventRecord(E[9], s));
  HC(hipMemsetAsync(d_pt, 0, sizeof(double) * 6 * (size_t)z.np, s));
  HC(launch_extract_diag(S, X, (double *)ptr(B_XD), s));
  HC(launch_point_output(S, Bk, (double *)ptr(B_XD), d_pt, z.np, s));
  HC(hipEventRecord(E[10], s));
  rc = copy_debug_blocks();
  if (rc) return rc;
  TeCtrl hc;
  HC(hipMemcpyAsync(&hc, ctrl, sizeof(hc), hipMemcpyDeviceToHost, s));
  HC(hipMemcpyAsync(&herr, S.err, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
  HC(hipMemcpyAsync(&kept, S.pt_off + z.npc, sizeof(int), hipMemcpyDeviceToHost, s));
  HC(hipStreamSynchronize(s));
  if (herr) return map_err(herr);

  if (st) {
    fill_stats(lambda, hc.nterms, hc.change);
    double gms = 0.0, gflops = 0.0, bms = 0.0, bflops = 0.0;
    int gcount = 0, bcount = 0;
    // a symmetric product computed on its lower tiles needs N (N + 1) / 2
    // entries of 2 N flops each; the first product falls back to the full
    // 2 N^3 where its split form does not apply
    const double nd = (double)N, sym_flops = nd * nd * (nd + 1.0);
    for (int p = first_chain_pair; p < gemm_pairs; ++p) {
      if (hc.nterms < lfirst[p] - 1) continue;  // exited at once: the series had stopped
      const double ms = ev_ms(GE[2 * p], GE[2 * p + 1]);
      const bool first = p == first_chain_pair;
      const double fl = (first || !batched) ? (sym_work_doubles > 0 ? sym_flops : 2.0 * nd * nd * nd)
                                            : lcount[p] * sym_flops;
      gms += ms;
      gflops += fl;
      ++gcount;
      if (lcount[p] == 4) {
        bms += ms;
        bflops += fl;
        ++bcount;
      }
    }
    st->gemm_ms = gms;
    st->gemm_count = gcount;
    st->gemm_flops = gflops;
    st->batch_ms = bms;
    st->batch_flops = bflops;
    st->batch_count = bcount;
  }
  return UNC_OK;
}

// ---- fixed scratch of the no-allocation entry point.  One array per device
// (each GPU the library runs on holds it until the process ends).  Host code
// never uses the array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device when its first context is created
// (the library's set-up call), and is kept per device.
constexpr int kMaxDevices = 64;
static __device__ __align__(256) char scratch[UNC_NOALLOC_SCRATCH_BYTES];
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {nullptr};

int scratch_setup(int device) {  // with `device` current
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return UNC_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    std::fprintf(stderr, "uncertainty_amd: hipGetSymbolAddress(scratch) failed on device %d: %s\n",
                 device, hipGetErrorString(e));
    (void)hipGetLastError();
    return UNC_ERR_DEVICE;
  }
  g_scratch_ptr[device] = (char *)p;
  return UNC_OK;
}

char *scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

int context_init(unc_context *c) {
  HC(hipSetDevice(c->device));
  HC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) ==
          hipSuccess &&
      lds > 0)
    c->max_lds = lds;
  for (auto &e : c->stage_ev) HC(hipEventCreate(&e));
  for (auto &e : c->gemm_ev) HC(hipEventCreate(&e));
  HC(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
  for (auto &e : c->fork_ev) HC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (auto &e : c->join_ev) HC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (auto &e : c->copy_ev) HC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return scratch_setup(c->device);
}

std::mutex g_default_mu;
unc_context *g_default[kMaxDevices] = {nullptr};

unc_context *default_context(int device, int *rc) {
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (device < 0 || device >= kMaxDevices) {
    *rc = UNC_ERR_BAD_INPUT;
    return nullptr;
  }
  if (!g_default[device]) {
    unc_context *c = nullptr;
    *rc = unc_context_create(device, &c);
    if (*rc) return nullptr;
    g_default[device] = c;
  }
  *rc = UNC_OK;
  return g_default[device];
}

}  // namespace

// ---- internal helpers for the request server (server.cpp)
size_t unc_ctx_held_bytes(unc_context *ctx) { return ctx ? ctx->ws_bytes + ctx->io_bytes : 0; }
size_t unc_ctx_ws_bytes(unc_context *ctx) { return ctx ? ctx->ws_bytes : 0; }
size_t unc_ctx_io_bytes(unc_context *ctx) { return ctx ? ctx->io_bytes : 0; }

void unc_ctx_set_fork(unc_context *ctx, bool on) {
  if (!ctx) return;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->fork = on;
}

void unc_ctx_release(unc_context *ctx) {
  if (!ctx) return;
  std::lock_guard<std::mutex> lk(ctx->mu);
  (void)hipSetDevice(ctx->device);
  if (ctx->ws) (void)hipFreeAsync(ctx->ws, ctx->stream);
  if (ctx->io) (void)hipFreeAsync(ctx->io, ctx->stream);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->ws = ctx->io = nullptr;
  ctx->ws_bytes = ctx->io_bytes = 0;
}

size_t unc_io_bytes_for(const unc_options *o, int num_rows, size_t nnz) {
  const int cp = o->cam_params;
  const size_t per = (size_t)cp * (cp + 1) / 2;
  auto r256 = [](size_t b) { return ((b + 255) / 256) * 256; };
  return r256(sizeof(int) * ((size_t)num_rows + 1)) + r256(sizeof(int) * nnz) +
         r256(sizeof(double) * nnz) + r256(sizeof(double) * (size_t)o->num_cams * per) +
         r256(sizeof(double) * (size_t)o->num_points * 6) + 256;
}

namespace {

bool page_locked(const void *p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

// Device -> host copies of the outputs, ordered after the work on s.  Page-locked
// destinations take one DMA each.  Pageable ones go through the context's
// page-locked staging buffer in chunks: every chunk's DMA is queued with an
// event behind it, and host threads copy chunks out as they land, so the host
// copy overlaps the transfer (the runtime's own path for pageable memory moves
// small pieces through a blit kernel, one host copy at a time).
struct Seg {
  void *dst;
  const void *src;
  size_t bytes;
};
constexpr size_t STAGE_MIN_CHUNK = 256u << 10, STAGE_MAX_CHUNK = 8u << 20;
constexpr size_t STAGE_THREAD_BYTES = 2u << 20;  // host threads only beyond this
constexpr int STAGE_THREADS = 4;

int download(unc_context *ctx, const Seg *segs, int nseg, hipStream_t s) {
  size_t pageable = 0;
  int npage = 0;
  bool locked[4] = {};
  if (nseg > 4) return UNC_ERR_BAD_INPUT;
  for (int i = 0; i < nseg; ++i) {
    if (segs[i].bytes == 0) continue;
    locked[i] = page_locked(segs[i].dst);
    if (locked[i])
      HC(hipMemcpyAsync(segs[i].dst, segs[i].src, segs[i].bytes, hipMemcpyDeviceToHost, s));
    else {
      pageable += segs[i].bytes;
      ++npage;
    }
  }
  if (pageable == 0) {
    HC(hipStreamSynchronize(s));
    return UNC_OK;
  }
  if (ctx->stage_bytes < pageable) {
    HC(hipStreamSynchronize(s));  // the old buffer may still be read by queued copies
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    ctx->stage = nullptr;
    ctx->stage_bytes = 0;
    if (hipHostMalloc((void **)&ctx->stage, pageable, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      ctx->stage = nullptr;
      return UNC_ERR_OOM;
    }
    ctx->stage_bytes = pageable;
  }
  const int max_chunks = (int)(sizeof(ctx->copy_ev) / sizeof(ctx->copy_ev[0])) - npage;
  size_t chunk = (pageable + max_chunks - 1) / max_chunks;
  chunk = std::max(chunk, std::min(STAGE_MAX_CHUNK, std::max(STAGE_MIN_CHUNK, pageable / 16)));
  chunk = (chunk + 4095) / 4096 * 4096;
  struct Piece {
    char *dst;
    const char *from;
    size_t bytes;
  };
  std::vector<Piece> pieces;
  size_t off = 0;
  for (int i = 0; i < nseg; ++i) {
    if (segs[i].bytes == 0 || locked[i]) continue;
    for (size_t o = 0; o < segs[i].bytes; o += chunk) {
      const size_t len = std::min(chunk, segs[i].bytes - o);
      HC(hipMemcpyAsync(ctx->stage + off, (const char *)segs[i].src + o, len,
                        hipMemcpyDeviceToHost, s));
      HC(hipEventRecord(ctx->copy_ev[pieces.size()], s));
      pieces.push_back({(char *)segs[i].dst + o, ctx->stage + off, len});
      off += len;
    }
  }
  std::atomic<int> next{0}, failed{0};
  auto worker = [&]() {
    for (int k; (k = next.fetch_add(1)) < (int)pieces.size();) {
      if (hipEventSynchronize(ctx->copy_ev[k]) != hipSuccess) {
        failed = 1;
        continue;
      }
      std::memcpy(pieces[k].dst, pieces[k].from, pieces[k].bytes);
    }
  };
  std::vector<std::thread> helpers;
  if (pageable > STAGE_THREAD_BYTES) {
    try {
      for (int t = 1; t < STAGE_THREADS && t < (int)pieces.size(); ++t)
        helpers.emplace_back(worker);
    } catch (...) {  // no helper threads to be had: this thread copies the rest
    }
  }
  worker();
  for (auto &h : helpers) h.join();
  HC(hipStreamSynchronize(s));
  return failed ? UNC_ERR_DEVICE : UNC_OK;
}

// The CRS index arrays of the standard layout (every observation two rows of
// cp camera columns then 3 point columns, rows equally long) reduced on the host
// to (first camera column, first point column - n_cam) per observation, in
// parallel chunks; false at the first row that departs from the layout (the
// device decode then reads the full arrays and reports it).
constexpr int IDS_THREADS = 8;
constexpr int IDS_THREAD_OBS = 1 << 16;  // observations per thread at least

bool host_ids(const int *rows, const int *cols, int nobs, int cp, int ncam_cols, int *ids) {
  const int rl = cp + 3;
  std::atomic<bool> ok{true};
  auto part = [&](int o0, int o1) {
    for (int o = o0; o < o1; ++o) {
      if ((o & 1023) == 0 && !ok.load(std::memory_order_relaxed)) return;
      const int r0 = 2 * o;
      const int *c0 = cols + (size_t)r0 * rl, *c1 = c0 + rl;
      bool good = rows[r0 + 1] == (r0 + 1) * rl && rows[r0 + 2] == (r0 + 2) * rl;
      const int cam0 = c0[0], pc = c0[cp] - ncam_cols;
      for (int j = 0; j < cp; ++j) good &= c0[j] == cam0 + j && c1[j] == cam0 + j;
      for (int j = 0; j < 3; ++j)
        good &= c0[cp + j] == ncam_cols + pc + j && c1[cp + j] == ncam_cols + pc + j;
      if (!good) {
        ok = false;
        return;
      }
      ids[2 * o] = cam0;
      ids[2 * o + 1] = pc;
    }
  };
  const int nt = std::max(1, std::min(IDS_THREADS, nobs / IDS_THREAD_OBS));
  const int per = (nobs + nt - 1) / nt;
  std::vector<std::thread> helpers;
  int t0 = 1;
  try {
    for (; t0 < nt; ++t0) helpers.emplace_back(part, t0 * per, std::min(nobs, (t0 + 1) * per));
  } catch (...) {  // fewer threads: this one takes the remaining chunks
  }
  part(0, std::min(nobs, per));
  for (int t = t0; t < nt; ++t) part(t * per, std::min(nobs, (t + 1) * per));
  for (auto &h : helpers) h.join();
  return ok.load();
}

int compute_host(unc_context *ctx, const unc_options *opts, int num_rows, int num_cols,
                 const int *rows, const int *cols, const double *values, double *cam_unc,
                 double *pt_unc, unc_stats *stats, unc_debug *dbg) {
  if (!opts || !rows || !cols || !values || !cam_unc || !pt_unc) return UNC_ERR_BAD_INPUT;
  int rc = UNC_OK;
  if (!ctx) ctx = default_context(opts->device, &rc);
  if (!ctx) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HC(hipSetDevice(ctx->device));
  if (num_rows <= 0 || (num_rows & 1)) return UNC_ERR_BAD_INPUT;
  const int cp = opts->cam_params;
  if (cp < 1 || cp > 9) return UNC_ERR_UNSUPPORTED;
  if (rows[0] != 0 || rows[num_rows] != num_rows * (cp + 3)) return UNC_ERR_BAD_INPUT;
  const size_t nnz = (size_t)rows[num_rows];
  const size_t per = (size_t)cp * (cp + 1) / 2;
  const size_t cam_n = (size_t)opts->num_cams * per, pt_n = (size_t)opts->num_points * 6;
  const size_t b_rows = ((sizeof(int) * (num_rows + 1) + 255) / 256) * 256;
  const size_t b_cols = ((sizeof(int) * nnz + 255) / 256) * 256;
  const size_t b_vals = ((sizeof(double) * nnz + 255) / 256) * 256;
  const size_t b_cam = ((sizeof(double) * cam_n + 255) / 256) * 256;
  const size_t b_pt = ((sizeof(double) * pt_n + 255) / 256) * 256;
  rc = ensure(&ctx->io, &ctx->io_bytes, b_rows + b_cols + b_vals + b_cam + b_pt + 256,
              ctx->stream);
  if (rc) return rc;
  int *d_rows = (int *)ctx->io;
  int *d_cols = (int *)(ctx->io + b_rows);
  double *d_vals = (double *)(ctx->io + b_rows + b_cols);
  double *d_cam = (double *)(ctx->io + b_rows + b_cols + b_vals);
  double *d_pt = (double *)(ctx->io + b_rows + b_cols + b_vals + b_cam);
  hipStream_t s = ctx->stream;
  HC(hipEventRecord(ctx->stage_ev[0], s));
  // the values first: the largest transfer is under way while the host checks
  // the index arrays' layout; where it holds, 8 bytes of ids per observation
  // cross the link instead of the index arrays (2 (cp + 3) + 2 ints)
  HC(hipMemcpyAsync(d_vals, values, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
  const int nobs = num_rows / 2;
  const int *d_ids = nullptr;
  const bool svd = opts->algorithm == UNC_ALG_SVD_QR_ITERATION ||
                   opts->algorithm == UNC_ALG_SVD_DEVIDE_AND_CONQUER;
  if (!svd && nobs > 0) {
    const size_t ids_bytes = sizeof(int) * 2 * (size_t)nobs;
    if (ctx->stage_bytes < ids_bytes) {
      HC(hipStreamSynchronize(s));  // a previous call's copies may still read the old buffer
      if (ctx->stage) (void)hipHostFree(ctx->stage);
      ctx->stage = nullptr;
      ctx->stage_bytes = 0;
      if (hipHostMalloc((void **)&ctx->stage, ids_bytes, hipHostMallocDefault) == hipSuccess)
        ctx->stage_bytes = ids_bytes;
      else
        (void)hipGetLastError();
    }
    if (ctx->stage && host_ids(rows, cols, nobs, cp, opts->num_cams * cp, (int *)ctx->stage)) {
      HC(hipMemcpyAsync(d_cols, ctx->stage, ids_bytes, hipMemcpyHostToDevice, s));
      d_ids = d_cols;
    }
  }
  if (!d_ids) {
    HC(hipMemcpyAsync(d_rows, rows, sizeof(int) * (num_rows + 1), hipMemcpyHostToDevice, s));
    HC(hipMemcpyAsync(d_cols, cols, sizeof(int) * nnz, hipMemcpyHostToDevice, s));
  }
  rc = run(ctx, opts, num_rows, num_cols, d_rows, d_cols, d_vals, d_cam, d_pt, stats, dbg, s, true,
           nullptr, d_ids);
  if (rc) return rc;
  // non-finite outputs are flagged on the device (one pass over both arrays)
  unsigned int *d_flag = (unsigned int *)(ctx->io + b_rows + b_cols + b_vals + b_cam + b_pt);
  HC(hipMemsetAsync(d_flag, 0, sizeof(unsigned int), s));
  HC(flag_nonfinite(cam_n, d_cam, d_flag, s));
  HC(flag_nonfinite(pt_n, d_pt, d_flag, s));
  unsigned int hflag = 0;
  const Seg segs[3] = {{cam_unc, d_cam, sizeof(double) * cam_n},
                       {pt_unc, d_pt, sizeof(double) * pt_n},
                       {&hflag, d_flag, sizeof(unsigned int)}};
  rc = download(ctx, segs, 3, s);
  if (rc) return rc;
  return hflag ? UNC_ERR_NUMERIC : UNC_OK;
}

}  // namespace

extern "C" {

void unc_default_options(unc_options *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->algorithm = UNC_ALG_TAYLOR_EXPANSION;
  o->epsilon = 1e-10;
  o->lambda = -1;
  o->cam_params = 9;
  o->svd_remove_n = -1;
  o->max_iter_te = -1;
  o->pts2fix[0] = o->pts2fix[1] = o->pts2fix[2] = -1;
  o->precision = UNC_PREC_FP64;
  o->refine = -1;
  o->device = 0;
}

int unc_context_create(int device, unc_context **out) {
  if (!out) return UNC_ERR_BAD_INPUT;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return UNC_ERR_DEVICE;
  }
  if (device < 0 || device >= ndev || device >= kMaxDevices) return UNC_ERR_BAD_INPUT;
  unc_context *c = new unc_context();
  c->device = device;
  for (auto &e : c->stage_ev) e = nullptr;
  for (auto &e : c->gemm_ev) e = nullptr;
  for (auto &e : c->fork_ev) e = nullptr;
  for (auto &e : c->join_ev) e = nullptr;
  const int rc = context_init(c);
  if (rc) {
    unc_context_destroy(c);
    return rc;
  }
  *out = c;
  return UNC_OK;
}

void unc_context_destroy(unc_context *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &e : c->stage_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->gemm_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->side) (void)hipStreamSynchronize(c->side);
  for (auto &e : c->fork_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->join_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->side) (void)hipStreamDestroy(c->side);
  for (auto &e : c->copy_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stage) (void)hipHostFree(c->stage);
  if (c->ws) (void)hipFreeAsync(c->ws, c->stream);
  if (c->io) (void)hipFreeAsync(c->io, c->stream);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int unc_compute(unc_context *ctx, const unc_options *opts, int num_rows, int num_cols,
                const int *rows, const int *cols, const double *values, double *cam_unc,
                double *pt_unc, unc_stats *stats) {
  return compute_host(ctx, opts, num_rows, num_cols, rows, cols, values, cam_unc, pt_unc, stats,
                      nullptr);
}

int unc_compute_debug(unc_context *ctx, const unc_options *opts, int num_rows, int num_cols,
                      const int *rows, const int *cols, const double *values, double *cam_unc,
                      double *pt_unc, unc_stats *stats, unc_debug *dbg) {
  return compute_host(ctx, opts, num_rows, num_cols, rows, cols, values, cam_unc, pt_unc, stats,
                      dbg);
}

int unc_compute_device(unc_context *ctx, const unc_options *opts, int num_rows, int num_cols,
                       const int *d_rows, const int *d_cols, const double *d_values,
                       double *d_cam_unc, double *d_pt_unc, unc_stats *stats, void *stream) {
  if (!opts || !d_rows || !d_cols || !d_values || !d_cam_unc || !d_pt_unc) return UNC_ERR_BAD_INPUT;
  int rc = UNC_OK;
  if (!ctx) ctx = default_context(opts->device, &rc);
  if (!ctx) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HC(hipSetDevice(ctx->device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  return run(ctx, opts, num_rows, num_cols, d_rows, d_cols, d_values, d_cam_unc, d_pt_unc, stats,
             nullptr, s, false, nullptr);
}

int unc_compute_device_noalloc(unc_context *ctx, const unc_options *opts, int num_rows,
                               int num_cols, const int *d_rows, const int *d_cols,
                               const double *d_values, double *d_cam_unc, double *d_pt_unc,
                               unc_stats *stats, void *stream) {
  if (!ctx || !opts || !d_rows || !d_cols || !d_values || !d_cam_unc || !d_pt_unc)
    return UNC_ERR_BAD_INPUT;
  if (opts->algorithm != UNC_ALG_TAYLOR_EXPANSION) return UNC_ERR_UNSUPPORTED;
  const unsigned long long need = unc_workspace_bytes(opts, num_rows);
  if (need == 0) return UNC_ERR_BAD_INPUT;
  if (need > UNC_NOALLOC_SCRATCH_BYTES) {
    std::fprintf(stderr,
                 "uncertainty_amd: unc_compute_device_noalloc: this solve needs %llu bytes of "
                 "device workspace, the fixed scratch holds %llu\n",
                 need, (unsigned long long)UNC_NOALLOC_SCRATCH_BYTES);
    return UNC_ERR_TOO_LARGE;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  HC(hipSetDevice(ctx->device));
  char *base = scratch_of(ctx->device);
  if (!base) return UNC_ERR_DEVICE;
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  Workspace ws;
  ws.base = base;
  ws.bytes = UNC_NOALLOC_SCRATCH_BYTES;
  return run(ctx, opts, num_rows, num_cols, d_rows, d_cols, d_values, d_cam_unc, d_pt_unc, stats,
             nullptr, s, false, &ws);
}

int unc_eig_cameras_device(unc_context *ctx, int num_cams, int cam_params,
                           const double *d_cam_unc, double *d_eigvals, double *d_center_vecs,
                           void *stream) {
  if (num_cams < 0 || cam_params < 1 || cam_params > 9) return UNC_ERR_BAD_INPUT;
  if (d_center_vecs && cam_params < 6) return UNC_ERR_BAD_INPUT;  // block(3, 3, 3, 3)
  if (num_cams == 0 || (!d_eigvals && !d_center_vecs)) return UNC_OK;
  if (!d_cam_unc) return UNC_ERR_BAD_INPUT;
  int rc = UNC_OK;
  if (!ctx) ctx = default_context(0, &rc);
  if (!ctx) return rc;
  HC(hipSetDevice(ctx->device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  HC(unc::launch_eig_cameras(d_cam_unc, num_cams, cam_params, d_eigvals, d_center_vecs, s));
  return UNC_OK;
}

int unc_eig_points_device(unc_context *ctx, int num_points, const double *d_pt_unc,
                          double *d_eigvals, double *d_eigvecs, void *stream) {
  if (num_points < 0) return UNC_ERR_BAD_INPUT;
  if (num_points == 0 || (!d_eigvals && !d_eigvecs)) return UNC_OK;
  if (!d_pt_unc) return UNC_ERR_BAD_INPUT;
  int rc = UNC_OK;
  if (!ctx) ctx = default_context(0, &rc);
  if (!ctx) return rc;
  HC(hipSetDevice(ctx->device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  HC(unc::launch_eig_points(d_pt_unc, num_points, d_eigvals, d_eigvecs, s));
  return UNC_OK;
}

int unc_eig_cameras(unc_context *ctx, int num_cams, int cam_params, const double *cam_unc,
                    double *eigvals, double *center_vecs) {
  if (num_cams < 0 || cam_params < 1 || cam_params > 9) return UNC_ERR_BAD_INPUT;
  if (center_vecs && cam_params < 6) return UNC_ERR_BAD_INPUT;
  if (num_cams == 0 || (!eigvals && !center_vecs)) return UNC_OK;
  if (!cam_unc) return UNC_ERR_BAD_INPUT;
  int rc = UNC_OK;
  if (!ctx) ctx = default_context(0, &rc);
  if (!ctx) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HC(hipSetDevice(ctx->device));
  auto r256 = [](size_t b) { return ((b + 255) / 256) * 256; };
  const size_t n_in = (size_t)num_cams * cam_params * (cam_params + 1) / 2;
  const size_t n_val = (size_t)num_cams * cam_params, n_vec = (size_t)num_cams * 9;
  const size_t b_in = r256(8 * n_in), b_val = eigvals ? r256(8 * n_val) : 0;
  const size_t b_vec = center_vecs ? r256(8 * n_vec) : 0;
  hipStream_t s = ctx->stream;
  rc = ensure(&ctx->io, &ctx->io_bytes, b_in + b_val + b_vec, s);
  if (rc) return rc;
  double *d_in = (double *)ctx->io;
  double *d_val = eigvals ? (double *)(ctx->io + b_in) : nullptr;
  double *d_vec = center_vecs ? (double *)(ctx->io + b_in + b_val) : nullptr;
  HC(hipMemcpyAsync(d_in, cam_unc, 8 * n_in, hipMemcpyHostToDevice, s));
  HC(unc::launch_eig_cameras(d_in, num_cams, cam_params, d_val, d_vec, s));
  if (eigvals) HC(hipMemcpyAsync(eigvals, d_val, 8 * n_val, hipMemcpyDeviceToHost, s));
  if (center_vecs) HC(hipMemcpyAsync(center_vecs, d_vec, 8 * n_vec, hipMemcpyDeviceToHost, s));
  HC(hipStreamSynchronize(s));
  return UNC_OK;
}

int unc_eig_points(unc_context *ctx, int num_points, const double *pt_unc, double *eigvals,
                   double *eigvecs) {
  if (num_points < 0) return UNC_ERR_BAD_INPUT;
  if (num_points == 0 || (!eigvals && !eigvecs)) return UNC_OK;
  if (!pt_unc) return UNC_ERR_BAD_INPUT;
  int rc = UNC_OK;
  if (!ctx) ctx = default_context(0, &rc);
  if (!ctx) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HC(hipSetDevice(ctx->device));
  auto r256 = [](size_t b) { return ((b + 255) / 256) * 256; };
  const size_t np = (size_t)num_points;
  const size_t b_in = r256(48 * np), b_val = eigvals ? r256(24 * np) : 0;
  const size_t b_vec = eigvecs ? r256(72 * np) : 0;
  hipStream_t s = ctx->stream;
  rc = ensure(&ctx->io, &ctx->io_bytes, b_in + b_val + b_vec, s);
  if (rc) return rc;
  double *d_in = (double *)ctx->io;
  double *d_val = eigvals ? (double *)(ctx->io + b_in) : nullptr;
  double *d_vec = eigvecs ? (double *)(ctx->io + b_in + b_val) : nullptr;
  HC(hipMemcpyAsync(d_in, pt_unc, 48 * np, hipMemcpyHostToDevice, s));
  HC(unc::launch_eig_points(d_in, num_points, d_val, d_vec, s));
  if (eigvals) HC(hipMemcpyAsync(eigvals, d_val, 24 * np, hipMemcpyDeviceToHost, s));
  if (eigvecs) HC(hipMemcpyAsync(eigvecs, d_vec, 72 * np, hipMemcpyDeviceToHost, s));
  HC(hipStreamSynchronize(s));
  return UNC_OK;
}

namespace {
int scene_jacobian_on(unc_context *ctx, int model, int nc, int np, int nobs, const double *cams,
                      const double *pts, const int *obs_cam, const int *obs_pt, int *rows,
                      int *cols, double *vals, hipStream_t s) {
  // the error word lives at the front of the context's upload buffer
  int rc = ensure(&ctx->io, &ctx->io_bytes, 256, s);
  if (rc) return rc;
  unsigned int *err = (unsigned int *)ctx->io;
  HC(hipMemsetAsync(err, 0, sizeof(unsigned int), s));
  HC(unc::launch_scene_jacobian(model, nc, np, nobs, cams, pts, obs_cam, obs_pt, rows, cols, vals,
                                err, s));
  unsigned int herr = 0;
  HC(hipMemcpyAsync(&herr, err, sizeof(herr), hipMemcpyDeviceToHost, s));
  HC(hipStreamSynchronize(s));
  if (herr & unc::ERRBIT_RANGE) return UNC_ERR_BAD_INPUT;
  if (herr & unc::ERRBIT_PROJECTION) return UNC_ERR_SINGULAR;
  return UNC_OK;
}
}  // namespace

int unc_scene_jacobian_device(unc_context *ctx, int cam_model, int num_cams, int num_points,
                              int num_obs, const double *d_cams, const double *d_points,
                              const int *d_obs_cam, const int *d_obs_pt, int *d_rows,
                              int *d_cols, double *d_vals, void *stream) {
  if (unc::scene_cam_params(cam_model) < 0) return UNC_ERR_UNSUPPORTED;
  if (num_cams <= 0 || num_points <= 0 || num_obs <= 0) return UNC_ERR_BAD_INPUT;
  if (!d_cams || !d_points || !d_obs_cam || !d_obs_pt || !d_rows || !d_cols || !d_vals)
    return UNC_ERR_BAD_INPUT;
  int rc = UNC_OK;
  if (!ctx) ctx = default_context(0, &rc);
  if (!ctx) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HC(hipSetDevice(ctx->device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  return scene_jacobian_on(ctx, cam_model, num_cams, num_points, num_obs, d_cams, d_points,
                           d_obs_cam, d_obs_pt, d_rows, d_cols, d_vals, s);
}

int unc_scene_jacobian(unc_context *ctx, int cam_model, int num_cams, int num_points,
                       int num_obs, const double *cams, const double *points,
                       const int *obs_cam, const int *obs_pt, int *rows, int *cols,
                       double *vals) {
  const int cp = unc::scene_cam_params(cam_model);
  if (cp < 0) return UNC_ERR_UNSUPPORTED;
  if (num_cams <= 0 || num_points <= 0 || num_obs <= 0) return UNC_ERR_BAD_INPUT;
  if (!cams || !points || !obs_cam || !obs_pt || !rows || !cols || !vals) return UNC_ERR_BAD_INPUT;
  if ((long long)2 * num_obs * (cp + 3) > 0x7fffffffLL) return UNC_ERR_UNSUPPORTED;
  int rc = UNC_OK;
  if (!ctx) ctx = default_context(0, &rc);
  if (!ctx) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HC(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  auto r256 = [](size_t b) { return ((b + 255) / 256) * 256; };
  const size_t nnz = (size_t)2 * num_obs * (cp + 3);
  const size_t b_err = 256, b_cam = r256(72 * (size_t)num_cams), b_pt = r256(24 * (size_t)num_points);
  const size_t b_obs = r256(4 * (size_t)num_obs), b_rows = r256(4 * (2 * (size_t)num_obs + 1));
  const size_t b_cols = r256(4 * nnz), b_vals = r256(8 * nnz);
  rc = ensure(&ctx->io, &ctx->io_bytes, b_err + b_cam + b_pt + 2 * b_obs + b_rows + b_cols + b_vals, s);
  if (rc) return rc;
  char *p = ctx->io + b_err;
  double *d_cam = (double *)p;
  double *d_pt = (double *)(p += b_cam);
  int *d_oc = (int *)(p += b_pt);
  int *d_op = (int *)(p += b_obs);
  int *d_rows = (int *)(p += b_obs);
  int *d_cols = (int *)(p += b_rows);
  double *d_vals = (double *)(p += b_cols);
  HC(hipMemcpyAsync(d_cam, cams, 72 * (size_t)num_cams, hipMemcpyHostToDevice, s));
  HC(hipMemcpyAsync(d_pt, points, 24 * (size_t)num_points, hipMemcpyHostToDevice, s));
  HC(hipMemcpyAsync(d_oc, obs_cam, 4 * (size_t)num_obs, hipMemcpyHostToDevice, s));
  HC(hipMemcpyAsync(d_op, obs_pt, 4 * (size_t)num_obs, hipMemcpyHostToDevice, s));
  rc = scene_jacobian_on(ctx, cam_model, num_cams, num_points, num_obs, d_cam, d_pt, d_oc, d_op,
                         d_rows, d_cols, d_vals, s);
  if (rc) return rc;
  HC(hipMemcpyAsync(rows, d_rows, 4 * (2 * (size_t)num_obs + 1), hipMemcpyDeviceToHost, s));
  HC(hipMemcpyAsync(cols, d_cols, 4 * nnz, hipMemcpyDeviceToHost, s));
  HC(hipMemcpyAsync(vals, d_vals, 8 * nnz, hipMemcpyDeviceToHost, s));
  HC(hipStreamSynchronize(s));
  return UNC_OK;
}

unsigned long long unc_workspace_bytes(const unc_options *o, int num_rows) {
  if (!o || o->num_cams <= 0 || o->cam_params <= 0 || o->num_points < 4 || num_rows <= 0) return 0;
  if (o->algorithm == UNC_ALG_SVD_QR_ITERATION || o->algorithm == UNC_ALG_SVD_DEVIDE_AND_CONQUER)
    return (unsigned long long)svd_layout((long long)o->num_cams * o->cam_params +
                                              3LL * o->num_points,
                                          num_rows, o->num_cams, o->cam_params)
        .total;
  Sizes z;
  z.nc = o->num_cams; z.cp = o->cam_params; z.np = o->num_points; z.npc = z.np - 3;
  z.nobs = num_rows / 2; z.N = z.nc * z.cp; z.nn = (size_t)z.N * z.N;
  z.te = o->precision == UNC_PREC_FP64;
  z.mixed = o->precision == UNC_PREC_MIXED;
  z.moduli = o->crt_moduli ? o->crt_moduli : ZPLUS_CRT_MOD;
  return (unsigned long long)make_layout(z).total;
}

const char *unc_error_string(int code) {
  switch (code) {
    case UNC_OK: return "ok";
    case UNC_ERR_BAD_INPUT: return "bad input (Jacobian layout, sizes or ids)";
    case UNC_ERR_NO_FIXED_POINTS: return "no fixed points given (the reference crashes here)";
    case UNC_ERR_SINGULAR: return "singular input (zero Jacobian column, singular 3x3 point block, or a null space beyond the 7-dim gauge)";
    case UNC_ERR_OOM: return "out of device memory";
    case UNC_ERR_DEVICE: return "HIP runtime error";
    case UNC_ERR_UNSUPPORTED: return "unsupported option or layout";
    case UNC_ERR_NUMERIC: return "non-finite result";
    case UNC_ERR_TOO_LARGE: return "the solve needs more device workspace than the fixed scratch of the no-allocation entry point";
    default: return "unknown error";
  }
}

const char *unc_stage_name(int stage) {
  return (stage >= 0 && stage < 12) ? kStageNames[stage] : "";
}

int unc_abi_version(void) { return UNC_ABI_VERSION; }

int unc_host_alloc(size_t bytes, void **ptr) {
  if (!ptr || bytes == 0) return UNC_ERR_BAD_INPUT;
  *ptr = nullptr;
  if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return UNC_ERR_OOM;
  }
  return UNC_OK;
}

void unc_host_free(void *ptr) {
  if (ptr) (void)hipHostFree(ptr);
}

int unc_device_lds_bytes(int device) {
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  return lds;
}

}  // extern "C"

// ---- entry points of the internal algorithms for the diagnostics library
// (uncertaintyte_amd/csrc/diag.hip), which tests kernels one at a time
namespace unc {
int dense_inverse_gj(int n, double *W, double *P3, double *T, double *R, unsigned int *err,
                     hipStream_t s) {
  return gj_inverse(n, W, P3, T, R, err, s);
}
int cholesky_inverse(int n, double *A, double *X, double *W, unsigned int *err, hipStream_t s) {
  return chol_inv(n, A, X, W, 0, n, err, RecProducts(), 0, s);
}
int error_code(unsigned int device_error_bits) { return map_err(device_error_bits); }
}  // namespace unc
