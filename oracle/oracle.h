/*
 * oracle.h -- CPU restatement of uncertaintyTE's TAYLOR_EXPANSION covariance
 * pipeline (src/uncertaintyTE/compute.cpp:458-540 and the containers it calls).
 *
 * TEST INFRASTRUCTURE ONLY.  This library gives the reference results the MI355X
 * path is compared with:
 * it may be linked or loaded by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, and by nothing else.  The product library
 * (uncertaintyte_amd/) never calls into it.
 *
 * Arithmetic contract.  Stages that the GPU path reproduces bit for bit
 * (fixPts, column scaling, camera/point block assembly, 3x3 inverses, W/Y
 * blocks, the Schur complement, camera output packing and the point
 * back-propagation given the same X) are written with an explicit operation
 * order: every multiply-add that is fused is an explicit fma(), everything
 * else is compiled with -ffp-contract=off.  The dense inverse and the Taylor
 * GEMM chain are tolerance stages (their summation order is free).
 *
 * Pinning (see tests/test_oracle_*.py):
 *   pipeline            -> test/EigValues/output.cov (Daliborka, 6 digits)
 *   generic J^T J       -> test/SparseMatrix/multipl_{in,out}_01.txt
 *   dense inverses      -> test/SparseMatrix/inv_{in,out}_01.txt,
 *                          test/DenseMatrix/inv_{in,out}_01.txt
 *   3x3 inverse, Schur, -> closed forms / brute force in numpy on small cases
 *   point step             (A^+ B Sigma B^T A^+^T with an SVD pseudo-inverse)
 *   fixPts              -> indirectly by output.cov (row-dropping semantics)
 */
#ifndef UNC_ORACLE_H
#define UNC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_OK = 0,
  ORC_ERR_BAD_INPUT = 1,
  ORC_ERR_NO_FIXED_POINTS = 2,
  ORC_ERR_SINGULAR = 3,
  ORC_ERR_OOM = 4,
  ORC_ERR_UNSUPPORTED = 6   /* a camera observes a point twice (Schur order needs one) */
};

/* Inverse used for X0 = (Z_s + lambda I)^-1 before the Taylor chain. */
enum {
  ORC_INV_REFINED = 0,   /* LU + one Newton step with a compensated (Dot2) residual */
  ORC_INV_REFERENCE = 1  /* row/column-equilibrated LU inverse, plain FP64
                            (ScaledDenseMatrix.cpp:232-279) */
};

typedef struct {
  int inverse_mode;   /* ORC_INV_* */
  int nthreads;       /* <=0: OpenMP default */
  double lambda;      /* -1: automatic, compute.cpp:123-125 */
} orc_opts;

typedef struct {
  double lambda;
  int n_terms;               /* accepted Taylor terms (size of cycle_change) */
  double cycle_change[20];
  int fixed[3];              /* triple after the partial sort of compute.cpp:285 */
  int n_obs_kept;            /* observations left after fixPts */
  int n_pts_compact;         /* numPoints - 3 */
} orc_stats;

/* Optional intermediate outputs (caller-allocated, any may be NULL).
 * N = ncams*cp, npc = npts-3.  Dense matrices are column-major. */
typedef struct {
  double *U_s;     /* ncams*cp*cp  scaled camera blocks, row-major per block */
  double *Vinv_s;  /* npc*9        inverse of the scaled point blocks          */
  double *d;       /* N + 3*npc    column scales 1/||J_:,j||                   */
  double *Z;       /* N*N          scaled Schur complement (without lambda)    */
  double *X0;      /* N*N          (Z + lambda I)^-1 as used by the chain      */
  double *X;       /* N*N          Taylor result                                */
} orc_debug;

/* fixPts (compute.cpp:283-368).  pts is partially sorted in place
 * (sort(pts, pts+2), compute.cpp:285).  Output buffers must hold
 * nrows+1 row pointers and nnz entries.  Returns ORC_ERR_* . */
int orc_fixpts(int nrows, int ncols, const int *rows, const int *cols,
               const double *vals, int cp, int ncams, int pts[3],
               int *out_nrows, int *out_ncols, int *out_rows, int *out_cols,
               double *out_vals);

/* Whole TAYLOR_EXPANSION pipeline on the ORIGINAL Jacobian (before fixPts). */
int orc_pipeline(const orc_opts *opts, int ncams, int cp, int npts,
                 const int fix_in[3], int nrows, int ncols, const int *rows,
                 const int *cols, const double *vals, double *cam_out,
                 double *pt_out, orc_stats *stats, orc_debug *dbg);

/* Front half of the pipeline only (fixPts, assembly, Schur): the scaled Schur
 * complement Z (column-major N*N) and the column scales d (N + 3*npc). */
int orc_front(const orc_opts *opts, int ncams, int cp, int npts, const int fix_in[3],
              int nrows, int ncols, const int *rows, const int *cols, const double *vals,
              double *Z_out, double *d_out);

/* Taylor chain alone on a given scaled Z (column-major N*N); X out. */
int orc_taylor(const orc_opts *opts, int N, const double *Z, double lambda,
               double *X0_out, double *X_out, orc_stats *stats);

/* Point back-propagation alone, for stage tests: given the fixed Jacobian
 * (2 rows per observation, 12 nnz per row) and a scaled X (N*N col-major). */
int orc_points_from_X(int ncams, int cp, int npts, const int fix[3],
                      int nrows_fixed, const int *cols_fixed,
                      const double *vals_fixed, const double *X, double *pt_out);
/* The same point step from compact camera blocks Xd (ncams x cp x cp,
 * column-major per block), only for the points all of whose observing cameras
 * have cam_ok set (pt_done[q] = 1 at their output id q); other points 0. */
int orc_points_from_blocks(int ncams, int cp, int npts, const int fix[3],
                           int nrows_fixed, const int *cols_fixed, const double *vals_fixed,
                           const double *Xd, const char *cam_ok, double *pt_out,
                           char *pt_done);

/* Generic CSR J^T J (sequential, dense column-major output ncols*ncols). */
int orc_jtj_dense(int nrows, int ncols, const int *rows, const int *cols,
                  const double *vals, double *out);

/* Dense inverses, column-major n*n. */
int orc_inv_lu(int n, const double *A, double *X);            /* partial pivoting */
int orc_inv_equilibrated(int n, const double *A, double *X);  /* reference form   */
/* R = I - A X with a compensated (Dot2) dot product, rounded once. */
void orc_residual_dd(int n, const double *A, const double *X, double *R);
/* C = A B, column-major n*n (blocked, OpenMP). */
void orc_gemm(int n, const double *A, const double *B, double *C);

/* The int32 factorial of compute.cpp:57-62 with its two's-complement wrap. */
int orc_factorial32(int n);
/* lambda of compute.cpp:125 for N camera parameters. */
double orc_auto_lambda(int N);

#ifdef __cplusplus
}
#endif
#endif
