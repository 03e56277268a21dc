// common.h -- shared definitions of the gfx950 covariance pipeline.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/uncertainty_amd.h"

namespace unc {

// Device-side error word bits (OR-ed by validation kernels).
enum : unsigned int {
  ERRBIT_LAYOUT = 1u,        // row pointers / column pattern not the .jacob layout
  ERRBIT_RANGE = 2u,         // camera or point id out of range
  ERRBIT_ZERO_COLUMN = 4u,   // a Jacobian column is all zero (d = 1/0)
  ERRBIT_SINGULAR_V = 8u,    // a 3x3 point block has no finite inverse
  ERRBIT_GJ_PIVOT = 32u,     // non-positive pivot in the dense inverse
  ERRBIT_PROJECTION = 128u,     // a point lies in a camera's focal plane (depth 0)
  ERRBIT_DUPLICATE = 256u,      // a camera observes the same point twice
};

// Control block of the Taylor chain, resident in device memory so the whole
// chain is enqueued without host round trips (compute.cpp:142-160).
struct TeCtrl {
  unsigned long long maxbits[21];  // max|T_i| as IEEE bits (non-negative doubles order as integers)
  double old_change;
  double coef;                     // k of the term accepted by the last control step
  double change[20];
  int active;                      // 1 while terms are being accepted
  int nterms;
  int pad[2];
};

#define UNC_HIP_CHECK(expr)                                                       \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      std::fprintf(stderr, "uncertainty_amd: %s failed at %s:%d: %s\n", #expr,  \
                   __FILE__, __LINE__, hipGetErrorString(e_));                   \
      return UNC_ERR_DEVICE;                                                      \
    }                                                                             \
  } while (0)

static inline unsigned int cdiv(long long a, long long b) { return (unsigned int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// kernel launchers (defined in the .hip files)
// ---------------------------------------------------------------------------
struct Structure {
  int cp, nc, np, npc, nobs, N;     // N = nc*cp
  int fix[3];                       // partially sorted triple
  const int *rows, *cols;           // input CRS (device)
  const double *vals;
  // per observation (first camera column, first point column - N) when the host
  // has checked the CRS layout; rows and cols are then not read
  const int *ids = nullptr;
  // outputs of decode
  int *obs_cam;                     // camera per observation, -1 when dropped
  int *obs_pt;                      // compact point per observation, -1 when dropped
  int *cam_off, *cam_list;          // camera-grouped observations (point order)
  int *pt_cnt, *pt_off, *pt_list;   // point-grouped observations (row order)
  int *pt_cursor;
  unsigned int *err;                // device error word
};

struct Blocks {
  double *Us;     // nc*cp*cp
  double *d;      // N + 3*npc
  double *Vi;     // npc*9
  double *Ws;     // nobs*cp*3
  int *pt_has;    // npc
};

hipError_t launch_decode(const Structure &s, hipStream_t st);
hipError_t launch_point_lists(const Structure &s, int *scan_tmp, hipStream_t st);
size_t camera_sort_temp_bytes(int nobs, int nc);
// cam_list / cam_off: the kept observations grouped by camera, each camera's in
// point order (a stable sort by camera of the point-grouped list).  Also
// pt_cams[u] = camera of the u-th point-grouped observation and cam_pts[t] =
// point of the t-th camera-grouped one.
hipError_t sort_obs_by_camera(const Structure &s, int *pt_cams, int *cam_pts, int *keys_in,
                              int *keys_out, int *vals_in, void *temp, size_t temp_bytes,
                              hipStream_t st);
// assembly in three launches: camera blocks U_s and d_c; point blocks V_s,
// d_p and V_s^-1; per-observation W_s
hipError_t launch_cam_blocks(const Structure &s, const Blocks &b, hipStream_t st);
hipError_t launch_point_blocks(const Structure &s, const Blocks &b, hipStream_t st);
hipError_t launch_obs_blocks(const Structure &s, const Blocks &b, hipStream_t st);
// Schur complement (k_schur.hip): cam_pts / pt_cams from sort_obs_by_camera;
// ghash holds 4 ints per observation (the global tables of cameras with more
// than 4096 observations)
hipError_t launch_schur(const Structure &s, const Blocks &b, double *Z, const int *cam_pts,
                        const int *pt_cams, int *ghash, hipStream_t st);
size_t scan_temp_bytes(int n);
hipError_t exclusive_scan_int(const int *in, int *out, int n, void *temp, size_t temp_bytes,
                              hipStream_t st);
hipError_t launch_cam_output(const Structure &s, const Blocks &b, const double *X,
                             double *cam_out, hipStream_t st);
hipError_t launch_extract_diag(const Structure &s, const double *X, double *Xd, hipStream_t st);
hipError_t launch_point_output(const Structure &s, const Blocks &b, const double *Xd,
                               double *pt_out, int npts_out, hipStream_t st);
// *flag |= 1 when any of a[0 .. n) is not finite (flag cleared by the caller)
hipError_t flag_nonfinite(size_t n, const double *a, unsigned int *flag, hipStream_t st);

// dense
hipError_t gemm_f64(int M, int N, int K, double alpha, const double *A, int lda, const double *B,
                    int ldb, double beta, double *C, int ldc, unsigned long long *absmax,
                    const int *active, hipStream_t st);
// C = alpha A B with the K range split over two workgroups per tile where that
// fills idle CUs (gemm_split_applies); work: M x N doubles of scratch, or
// nullptr for the unsplit kernel.  Same absmax / active contract as gemm_f64.
bool gemm_split_applies(int M, int N, int K);
// C = alpha A B for n x n operands (ld n) whose product is symmetric (powers
// of one symmetric matrix): only the 128 x 128 tiles on and below the
// diagonal are computed, over three or four K chunks where gemm_split_applies (work:
// gemm_sym_split_work_doubles(n) doubles), and the lower triangle is mirrored;
// otherwise the plain product.  Same absmax / active contract as gemm_f64.
// parts: K chunks (1..4), 0 = by the launch-size rule.
size_t gemm_sym_split_work_doubles(int n);
hipError_t gemm_f64_sym(int n, double alpha, const double *A, const double *B, double *C,
                        double *work, size_t work_doubles, unsigned long long *absmax,
                        const int *active, hipStream_t st, int parts = 0);
// nb <= 4 symmetric products C_b = A B_b in one launch (A and every B_b
// symmetric, every product symmetric: powers of one matrix), on the 128 x 128
// tiles on and below the diagonal, each workgroup forming one tile over the
// whole K range, lower triangle mirrored in the epilogue; max|C_b| folded into
// absmax[b].  The grid (nb x lower tiles) is dealt to the XCDs product by
// product.  Exits at once when *active == 0.
hipError_t gemm_f64_sym_batch(int n, int nb, const double *A, const double *const B[4],
                              double *const C[4], unsigned long long *absmax, const int *active,
                              hipStream_t st);
hipError_t gemm_f64_split(int M, int N, int K, double alpha, const double *A, int lda,
                          const double *B, int ldb, double *C, int ldc, double *work,
                          unsigned long long *absmax, const int *active, hipStream_t st);
// triangular-operand flags of gemm_f64_ex
constexpr int GEMM_A_LOWER = 1;   // A lower triangular: K range ends at the tile's last row
constexpr int GEMM_B_LOWER = 2;   // B lower triangular: K range starts at the tile's first column
constexpr int GEMM_BT_LOWER = 4;  // op(B) = S^T with S lower: K range ends at the tile's last column
constexpr int GEMM_C_LOWER = 8;   // only tiles on or below the diagonal of C are written
// with GEMM_C_LOWER on a square C: the grid lists only the lower tiles (times
// the K parts), in 4 x 4 tile blocks, each XCD taking a contiguous range
constexpr int GEMM_LOWER_LIST = 16;
// all tiles (times the K parts) in 4 x 4 tile blocks, each XCD taking a
// contiguous range of them
constexpr int GEMM_BLOCKED = 32;
// with GEMM_LOWER_LIST: the tile index is the lower-list position itself (the
// caller has already spread its grid over the XCDs); K is not split
constexpr int GEMM_TILE_DIRECT = 64;
// C is symmetric: of a lower tile only the elements on or below the diagonal
// are formed, each written at (row, col) and (col, row); max|C| folds over them
constexpr int GEMM_SYM_OUT = 128;
hipError_t gemm_f64_ex(int M, int N, int K, double alpha, const double *A, int lda, const double *B,
                       int ldb, bool trans_b, double beta, double *C, int ldc, int tri,
                       hipStream_t st);
// C = alpha op(A) op(B) + beta C from the exact integer product of op(A) and
// op(B) scaled to q-bit rows/columns, formed modulo nmod coprime moduli on the
// int8 matrix cores and rebuilt by the Chinese remainder theorem
// (k_crt_gemm.hip); same operand flags as gemm_f64_ex.  work: 256-byte aligned,
// crt_gemm_work_bytes(M, N, K, nmod) bytes.
constexpr int CRT_MIN_MOD = 6, CRT_MAX_MOD = 20, CRT_MAX_K = 131071;
int crt_bits(int nmod);  // q of nmod moduli (0 when out of range)
size_t crt_gemm_work_bytes(int M, int N, int K, int nmod);
hipError_t gemm_crt(int M, int N, int K, double alpha, const double *A, int lda, const double *B,
                    int ldb, bool trans_b, double beta, double *C, int ldc, int tri, int nmod,
                    void *work, size_t work_bytes, hipStream_t st);
// The same in two parts, so that one operand's residues can serve several
// products: crt_split cuts the R rows of an operand (element (r, k) at
// P[r + k ld] when gather, else at P[k + r ld]; mask 1 keeps k <= r, 2 keeps
// k >= r) into residue planes in `area` (crt_operand_bytes, 256-byte aligned);
// crt_product forms C = alpha A B^T + beta C of two such operands (rows of A
// are C's rows, rows of B its columns; tri as for gemm_f64_ex, consistent with
// the operands' masks) with its residue planes in `cplanes` (crt_product_bytes);
// doff >= 0 subtracts the identity exactly, C = alpha (A B^T - D) + beta C with
// D_ij = 1 where i + doff == j (I - A B^T for alpha = -1).
struct CrtOperand {
  signed char *planes = nullptr;
  int *expo = nullptr;
  int R = 0, K = 0, Rp = 0, Kp = 0, nmod = 0;
};
size_t crt_operand_bytes(int R, int K, int nmod);
size_t crt_product_bytes(int M, int N, int nmod);
hipError_t crt_split(int R, int K, const double *P, int ld, bool gather, int mask, int nmod,
                     void *area, CrtOperand *out, hipStream_t st);
hipError_t crt_product(double alpha, const CrtOperand &A, const CrtOperand &B, double beta,
                       double *C, int ldc, int tri, void *cplanes, size_t cplanes_size,
                       hipStream_t st, int doff = -1);
// R = I - A X for symmetric A (the Newton step's residual) on the int8 matrix
// cores: CRT_RESID_MOD moduli (q = 77 bits per cut row / column), A in row
// blocks sized to `work` (256-byte aligned; residual_crt_work_bytes(n, blocks)
// for `blocks` row blocks); invalid argument when even 64 blocks do not fit.
constexpr int CRT_RESID_MOD = 20;
size_t residual_crt_work_bytes(int n, int blocks);
hipError_t residual_crt(int n, const double *A, const double *X, double *R, void *work,
                        size_t work_bytes, hipStream_t st);
// gauge-fixed Z^+ path (k_zplus.hip)
// a recursion node of 64 < n <= 128 rows (halves 64 and n - 64) in one launch
hipError_t chol_inv_node128(int n, double *A, int lda, double *X, int ldx, unsigned int *err,
                            hipStream_t st);
hipError_t chol_inv_leaf(int b, double *A, int lda, double *X, int ldx, unsigned int *err,
                         hipStream_t st);
// P = W^-1 of a b x b SPD block (b <= 64) by Cholesky, L^-1 and X^T X (k_zplus.hip)
hipError_t spd_inverse64(int b, const double *W, int ldw, double *P, unsigned int *err,
                         hipStream_t st);
// blocked Gauss-Jordan panel kernels (k_zplus.hip): the first half of a panel
// (T, R and the copy of the next diagonal block), then the trailing update,
// the panel's rows and columns and the next panel's P in one launch
hipError_t gj_panel_a(int n, int k0, int kb, const double *W, const double *P, double *T,
                      double *R, int k1, int kb1, double *Dsrc, hipStream_t st);
hipError_t gj_update(int n, int k0, int kb, double *W, const double *P, const double *T,
                     const double *R, int kb1, const double *Dsrc, double *Pn, unsigned int *err,
                     hipStream_t st);
hipError_t skinny_tn(int n, int kc, const double *A, int lda, const double *B, double *C,
                     bool lower, hipStream_t st);
hipError_t skinny_nn(int n, int kc, const double *A, int lda, const double *B, double *C,
                     double *part, bool lower, hipStream_t st);
size_t skinny_part_doubles(int n);
hipError_t gram(int n, int k1, const double *V, int k2, const double *W, double *partial,
                double *G, hipStream_t st);
hipError_t small_orth(int k, const double *G, double *T, unsigned int *err, hipStream_t st);
hipError_t small_woodbury(int k, const double *G, double *S, unsigned int *err, hipStream_t st);
hipError_t small_apply(int n, int k1, int k2, double alpha, const double *A, double beta,
                       const double *B, const double *T, double *out, hipStream_t st);
hipError_t fill_hash(size_t cnt, double *V, unsigned long long seed, hipStream_t st);
hipError_t rank_copy(int n, int k, const double *A, const double *V, double *out, hipStream_t st);
// *out = max_i X_ii^2 (one workgroup; X n x n, leading dimension ldx)
hipError_t diag_max_sq(int n, const double *X, int ldx, double *out, hipStream_t st);
hipError_t cam_gram(const Structure &s, const double *X, const double *MU, const double *S,
                    const double *Nb, double *Xd, hipStream_t st);
hipError_t cam_output_cov(const Structure &s, const double *Xd, const double *d, double *out,
                          hipStream_t st);
// R = I - A X accurate to about eps |R| by error-free extraction.  scratch:
// 2 n doubles for the row/column maxima.  tile_cols 16 or 32 forces 64 x 16 or
// 64 x 32 tiles; anything else chooses by the grid's fill of the device.
hipError_t residual_ex(int n, const double *A, const double *X, double *R, double *scratch,
                       hipStream_t st, int tile_cols = 0);
hipError_t copy_block(int m, int n, const double *src, int lds, double *dst, int ldd, double scale,
                      hipStream_t st);
hipError_t add_diag(int n, double *A, int lda, double v, hipStream_t st);
hipError_t sym_update(int n, const double *X0, const double *D, double *out, hipStream_t st);
// out = X0 + D when both are exactly symmetric (the same value as sym_update)
hipError_t add_symmetric(int n, const double *X0, const double *D, double *out, hipStream_t st);
// mx[j] = max_i |A_ij| (n x n, column-major) as IEEE bits
hipError_t col_absmax(int n, const double *A, unsigned long long *mx, hipStream_t st);
hipError_t te_init(TeCtrl *ctrl, hipStream_t st);
// Terms i0 .. i0+m-1 (m <= 4, coefficients k[]) of the chain at once: the
// accept / stop decisions of te_step taken in term order on one thread, then
// X += k_j T_j over the accepted terms, in term order.
hipError_t te_steps(int n, TeCtrl *ctrl, int i0, int m, const double k[4], double *X,
                    const double *const T[4], hipStream_t st);
hipError_t te_step(int n, TeCtrl *ctrl, int i, double k, double *X, const double *T,
                   hipStream_t st);
// A[0, n) = 0 (A 16-byte aligned)
hipError_t fill_zero(double *A, size_t n, hipStream_t st);
hipError_t absmax_f64(size_t n, const double *A, unsigned long long *out, hipStream_t st);
// SVD algorithm modes (k_svd.hip)
hipError_t svd_jt_scatter(int nrows, int npar, const int *rows, const int *cols,
                          const double *vals, double *Jt, unsigned int *err, hipStream_t st);
hipError_t svd_symmetrize(int n, double *M, hipStream_t st);
hipError_t svd_identity(int n, double *V, hipStream_t st);
hipError_t svd_jacobi_round(int n, int m, int r, const double *Aold, double *Anew, double *V,
                            int *any_rot, hipStream_t st);
hipError_t svd_diag(int n, const double *A, double *d, hipStream_t st);
hipError_t svd_cam_blocks(int nc, int cp, int n, const double *V, const double *w, double *Zc,
                          double *cam_out, hipStream_t st);
hipError_t svd_points(int nobs, int cp, int nc, int np, const int *cols, const double *vals,
                      const double *Zc, double *pt_out, unsigned int *err, hipStream_t st);
// scene Jacobian (k_scene.hip)
int scene_cam_params(int model);  // cp of a UNC_CAM_* model, -1 if unknown
hipError_t launch_scene_jacobian(int model, int nc, int np, int nobs, const double *cams,
                                 const double *pts, const int *obs_cam, const int *obs_pt,
                                 int *rows, int *cols, double *vals, unsigned int *err,
                                 hipStream_t st);
// eigen accessors (k_eig.hip): one lane per block, descending eigenvalues
hipError_t launch_eig_cameras(const double *cam, int nc, int cp, double *vals, double *cvecs,
                              hipStream_t st);
hipError_t launch_eig_points(const double *pts, long long np, double *vals, double *vecs,
                             hipStream_t st);

// pipeline algorithms (pipeline.hip) for the diagnostics library: blocked
// Gauss-Jordan inverse of SPD W in place (P3: 3 x 64 x 64, T: n x 64, R: 64 x n
// doubles of scratch); recursive Cholesky A = L L^T with X = L^-1 (W: scratch
// of (n/2)^2 doubles); the C ABI code of a device error word.  They return
// UNC_* codes.
int dense_inverse_gj(int n, double *W, double *P3, double *T, double *R, unsigned int *err,
                     hipStream_t s);
int cholesky_inverse(int n, double *A, double *X, double *W, unsigned int *err, hipStream_t s);
int error_code(unsigned int device_error_bits);

// 1/sqrt(x) and 1/x for normal, finite x: the hardware estimates refined by two
// Newton steps each (quadratic convergence from the estimate's >= 20 bits),
// without the special-value handling of the library versions; callers use the
// result only where the operand is never zero, denormal, infinite or NaN.
__device__ __forceinline__ double rsqrt_normal(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = fma(y, fma(-hx * y, y, 0.5), y);
  y = fma(y, fma(-hx * y, y, 0.5), y);
  return y;
}

__device__ __forceinline__ double rcp_normal(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  y = fma(y, fma(-x, y, 1.0), y);
  return y;
}

}  // namespace unc
