/*
 * uncertainty_amd.h -- C ABI of the MI355X (gfx950) camera + point covariance
 * pipeline.  Plain pointers and sizes; no C++, HIP or torch types cross it.
 *
 * It replaces the reference's computation entry point
 *     void computeCovariances(cov::Options&, cov::Statistic&, ceres::CRSMatrix&,
 *                             double* camUnc, double* ptUnc)
 *     (src/uncertaintyTE/compute.h:92-97, implemented at compute.cpp:458-540)
 * which the reference's public getCovariances (include/uncertainty.h:48-61) and
 * its CLI (src/software/main_uncertainty.cpp:86) call.  The C++ wrapper with the
 * reference's own signature is declared in include/uncertainty.h.
 *
 * Semantics follow the reference TAYLOR_EXPANSION path:
 *   - the Jacobian is a CRS matrix with 2 rows per observation and cp+3
 *     nonzeros per row (cp camera columns, then 3 point columns);
 *   - three fixed points remove the point columns and every observation of
 *     those points (compute.cpp:283-368);
 *   - output: cp(cp+1)/2 upper-triangle values per camera, row-major per block
 *     (compute.cpp:168-183, including its (X[j,k] + X[k,j]/2) form), and 6
 *     values (xx, xy, xz, yy, yz, zz) per point (compute.cpp:443-455), fixed
 *     points left at zero.
 * Differences from the reference: errors are returned as codes instead of
 * exit(1)/segfaults, stage times are true seconds (the reference prints ten
 * times the seconds, compute.cpp:33), and the call is reentrant per context.
 */
#ifndef UNCERTAINTY_AMD_H
#define UNCERTAINTY_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UNC_ABI_VERSION 3  /* 2: unc_stats gained batch_ms, batch_flops, batch_count;
                              3: unc_options.crt_moduli, unc_stats.zplus_moduli and
                                 zplus_cond_lb */

/* error codes */
#define UNC_OK 0
#define UNC_ERR_BAD_INPUT 1        /* layout/size/ids inconsistent with the options      */
#define UNC_ERR_NO_FIXED_POINTS 2  /* reference crashes here (compute.cpp:374-378, 446)  */
#define UNC_ERR_SINGULAR 3         /* zero Jacobian column or singular 3x3 point block   */
#define UNC_ERR_OOM 4              /* device memory (or admission) refused               */
#define UNC_ERR_DEVICE 5           /* HIP runtime error                                  */
#define UNC_ERR_UNSUPPORTED 6      /* option combination not implemented on the device   */
#define UNC_ERR_NUMERIC 7          /* NaN/Inf in the result                              */
#define UNC_ERR_TOO_LARGE 8        /* workspace beyond UNC_NOALLOC_SCRATCH_BYTES           */

/* cov::Options::_algorithm values (include/uncertainty.h:11-13) */
#define UNC_ALG_SVD_QR_ITERATION 0
#define UNC_ALG_SVD_DEVIDE_AND_CONQUER 1
#define UNC_ALG_TAYLOR_EXPANSION 2

/* arithmetic of the Z inverse (the only stage whose precision is selectable) */
#define UNC_PREC_FP64 0        /* plain FP64 TE path: reference output, <=1e-9 per block  */
#define UNC_PREC_MIXED 1       /* gauge-fixed Z^+ (contract: <=1e-5 per block vs the FP64
                                  pseudo-inverse): the large products of the Cholesky
                                  recursion on the int8 matrix cores (exact integer
                                  products modulo crt_moduli coprime moduli, rebuilt
                                  in FP64), the rest as ZPLUS_FP64                     */
#define UNC_PREC_ZPLUS_FP64 2  /* gauge-fixed Z^+ with FP64 inverses (the FP64 Z^+)     */
/* Output semantics: UNC_PREC_FP64 reproduces the reference TAYLOR output
 * (including its (X[j,k] + X[k,j]/2) camera packing).  The two Z^+ modes return
 * the covariance D Z^+ D of the gauge-fixed problem (camera upper triangles, no
 * 1.5 factor) and the points propagated from it; they are compared with an FP64
 * pseudo-inverse that drops the 7 gauge eigenvalues, not with the reference. */

typedef struct unc_options {
  int algorithm;      /* UNC_ALG_*: TAYLOR_EXPANSION, or the two SVD modes (dense
                         pseudo-inverse of J^T J, at most 8192 parameters)          */
  double epsilon;     /* cov::Options::_epsilon (SVD modes only)                      */
  double lambda;      /* -1: 10^(-1.2653 log10(N) - 2.9415), compute.cpp:123-125    */
  int num_cams;
  int cam_params;     /* cp, 1..9                                                     */
  int num_points;
  int num_obs;        /* informational (header of the .jacob)                         */
  int svd_remove_n;   /* SVD modes: drop this many smallest singular values; -1 =
                         drop those <= epsilon instead (compute.cpp:222-240)        */
  int max_iter_te;    /* unused by the reference; kept for layout parity              */
  int pts2fix[3];     /* fixed point ids, -1 -1 -1 = none                             */
  int precision;      /* UNC_PREC_*                                                   */
  int refine;         /* Newton steps on the inverse; -1 = automatic (default):
                         FP64 path 1 step with an error-free residual on
                         (Z + lambda I)^-1; Z^+ modes 1 step on the null-space
                         basis N; 0 disables refinement                           */
  int device;         /* HIP device ordinal                                           */
  int flags;          /* reserved, 0                                                  */
  int crt_moduli;     /* UNC_PREC_MIXED: moduli of the int8 CRT products, 6..20 (more
                         moduli = more bits kept: q = 23, 27, 31, 35, 39, 43, 46, 50,
                         54, 58, 62, 65, 69, 73, 77); 0 = automatic (12)             */
} unc_options;

typedef struct unc_stats {
  /* cov::Statistic times (include/uncertainty.h:36), in true seconds */
  double time_create_j, time_fix_j, time_norm_j, time_multiply_jj, time_split_jj,
      time_inv_v, time_compose_z, time_inv_z, time_te, time_pts_unc, time_all;
  double lambda;
  int fixed_pts[3];        /* after the reference's partial sort (compute.cpp:285) */
  int num_terms;           /* accepted Taylor terms = cycle_change size            */
  double cycle_change[20];
  /* device-side measurements */
  double gemm_ms;          /* summed duration of the Taylor chain's product launches    */
                           /* that did work (HIP events; one launch forms 1-4 products) */
  double gemm_flops;       /* their algorithmic flops: 2 N^3 per product, N^2 (N + 1)   */
                           /* for the symmetric form that computes the lower tiles only */
  int gemm_count;          /* launches                                                */
  int num_obs_kept;        /* observations left after fixPts                          */
                           /* time_norm_j and time_split_jj are 0: the column norms are
                              the assembled blocks' diagonals and the blocks are
                              written split; time_multiply_jj = camera blocks + W_s,
                              time_inv_v = point blocks with their 3x3 inverses     */
  double stage_ms[12];     /* per device stage, see unc_stage_name()                  */
  /* the chain's launches of four products each (where the chain is batched,
     N >= 1024): their summed duration, algorithmic flops and count; included
     in gemm_ms / gemm_flops / gemm_count, which count launches */
  double batch_ms;
  double batch_flops;
  int batch_count;
  /* Z^+ paths: moduli of the CRT products (0 = FP64 products throughout), and
     1 / min_i L_ii^2 for the Cholesky factor L of Z~ = Z_s + N0 N0^T, a lower
     bound of ||Z~^-1||_2 (each pivot L_ii^2 is at least lambda_min(Z~)); Z_s has
     a unit-scale diagonal, so this is a lower estimate of cond(Z~) */
  int zplus_moduli;
  double zplus_cond_lb;
} unc_stats;

/* Optional host copies of intermediates (any pointer may be NULL).
 * N = num_cams*cam_params, npc = num_points-3, dense matrices column-major. */
typedef struct unc_debug {
  double *U_s;     /* num_cams*cp*cp scaled camera blocks (row-major per block) */
  double *Vinv_s;  /* npc*9 inverses of the scaled point blocks                  */
  double *d;       /* N + 3*npc column scales                                     */
  double *Z;       /* N*N scaled Schur complement, before + lambda I              */
  double *X0;      /* N*N (Z + lambda I)^-1 as used by the chain                  */
  double *X;       /* N*N Taylor result (FP64 path); on the Z^+ paths L^-1 of the
                      Cholesky factor of Z + N0 N0^T (lower triangle)            */
} unc_debug;

typedef struct unc_context unc_context;

/* Fill defaults: TAYLOR_EXPANSION, lambda -1, FP64, refine -1 (auto), device 0. */
void unc_default_options(unc_options *opts);

/* Per-device context: stream, reusable workspace, timing events.  Reentrant
 * across contexts; one context must not be used by two threads at once. */
int unc_context_create(int device, unc_context **ctx);
void unc_context_destroy(unc_context *ctx);

/* Host-memory drop-in for computeCovariances (compute.cpp:458).  rows has
 * num_rows+1 entries, cols/values rows[num_rows] entries.  cam_unc holds
 * num_cams*cp(cp+1)/2 doubles, pt_unc num_points*6.  ctx may be NULL (a
 * per-device default context is used, serialised by a lock). */
int unc_compute(unc_context *ctx, const unc_options *opts, int num_rows,
                int num_cols, const int *rows, const int *cols,
                const double *values, double *cam_unc, double *pt_unc,
                unc_stats *stats);

/* Same with every array already resident in device memory (HBM) on ctx's
 * device; outputs are device pointers.  stream: hipStream_t or NULL for the
 * context's stream.  Returns after the results are complete. */
int unc_compute_device(unc_context *ctx, const unc_options *opts, int num_rows,
                       int num_cols, const int *d_rows, const int *d_cols,
                       const double *d_values, double *d_cam_unc,
                       double *d_pt_unc, unc_stats *stats, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): unc_compute_device with every device buffer of the solve
 * carved from a fixed scratch area of UNC_NOALLOC_SCRATCH_BYTES that the
 * library holds on each device it runs on; the call allocates nothing.
 * ctx must be a context from unc_context_create (whose creation takes the
 * scratch's device address).  A solve whose unc_workspace_bytes() exceeds the
 * scratch returns UNC_ERR_TOO_LARGE (a message names both sizes).  All calls
 * on one device share its scratch and find the last call's contents there
 * (the solve initialises everything it reads): calls that overlap on one GPU
 * race and must be serialised by the caller.  TAYLOR_EXPANSION only. */
#define UNC_NOALLOC_SCRATCH_BYTES (64ull << 20)
int unc_compute_device_noalloc(unc_context *ctx, const unc_options *opts, int num_rows,
                               int num_cols, const int *d_rows, const int *d_cols,
                               const double *d_values, double *d_cam_unc,
                               double *d_pt_unc, unc_stats *stats, void *stream);

/* unc_compute plus host copies of intermediates (parity tests). */
int unc_compute_debug(unc_context *ctx, const unc_options *opts, int num_rows,
                      int num_cols, const int *rows, const int *cols,
                      const double *values, double *cam_unc, double *pt_unc,
                      unc_stats *stats, unc_debug *dbg);

/* Device bytes a solve of this size needs (admission control). */
unsigned long long unc_workspace_bytes(const unc_options *opts, int num_rows);

/* ---- eigen accessors of cov::Uncertainty on the device
 * (src/uncertaintyTE/uncertainty.cpp:74-115 and :146-185, declared at
 * src/uncertaintyTE/uncertainty.h:110-142).  Inputs are the packed outputs
 * above: cp(cp+1)/2 values per camera, 6 per point.  Any output may be NULL.
 *   eigvals      cp per camera / 3 per point, descending, as
 *                getCamerasUncEigenValues / getPointsUncEigenValues return them
 *                (the unpacked blocks are symmetric, so the real parts the
 *                reference keeps are the eigenvalues)
 *   center_vecs  9 per camera: eigenvectors of the camera-centre block, rows
 *                and columns 3..5 (getCamerasCentersUncEigVec); needs cp >= 6
 *   eigvecs      9 per point (getPointsUncEigVectors)
 * Vector k of a block sits at [3k, 3k+3) and has unit length.  Order and sign
 * differ from the reference, whose order and signs are whatever Eigen's
 * EigenSolver produces: here the vectors follow the descending eigenvalues and
 * each has its largest-magnitude component positive.
 * The host variants copy through ctx's device (ctx NULL: device 0) and return
 * with the results in place.  The _device variants take device pointers,
 * enqueue on stream (NULL: ctx's stream) and return without waiting. */
int unc_eig_cameras(unc_context *ctx, int num_cams, int cam_params, const double *cam_unc,
                    double *eigvals, double *center_vecs);
int unc_eig_points(unc_context *ctx, int num_points, const double *pt_unc, double *eigvals,
                   double *eigvecs);
int unc_eig_cameras_device(unc_context *ctx, int num_cams, int cam_params,
                           const double *d_cam_unc, double *d_eigvals, double *d_center_vecs,
                           void *stream);
int unc_eig_points_device(unc_context *ctx, int num_points, const double *d_pt_unc,
                          double *d_eigvals, double *d_eigvecs, void *stream);

/* ---- reconstruction front-end: COLMAP text model -> CRS Jacobian
 * Replaces the reference's ColmapIO::read (ColmapIO.cpp:24-146) and
 * JacobianComposer::scene2Jacobian (JacobianComposer.cpp:84-194, Ceres
 * autodiff of snavely_reprojection_error.h:148-302) with a C++ reader and an
 * analytic Jacobian computed on the device.  Camera models (-cam of the
 * reference CLI) and the camera columns they free: */
#define UNC_CAM_SIMPLE_PINHOLE 0  /* cp 6: aa, C          (f, k1, k2 held)  */
#define UNC_CAM_PINHOLE 1         /* cp 7: aa, C, f       (k1, k2 held)     */
#define UNC_CAM_SIMPLE_RADIAL 2   /* cp 8: aa, C, f, k1   (no k2 term)      */
#define UNC_CAM_RADIAL 3          /* cp 9: aa, C, f, k1, k2                 */
typedef struct unc_scene unc_scene;
/* cameras.txt, images.txt and points3D.txt of the COLMAP text model in dir.
 * Order (that of the reference's parameter blocks, JacobianComposer.cpp:
 * 123-157): images by id, those without observations dropped; points by id;
 * observations image by image, each image's by point id. */
int unc_scene_read_colmap(const char *dir, unc_scene **out);
/* A BAL problem file (replaces BALProblem::BALProblem(filename),
 * bal_problem.cpp:111-172): "nc np nobs", nobs lines "cam pt x y", 9 nc camera
 * numbers, 3 np point numbers.  File order throughout.  The camera numbers mean
 * what the reference's BAL functor makes of them (SnavelyReprojectionError,
 * snavely_reprojection_error.h:309-347, bundle type 0 with unit weights):
 * aa[3], the camera centre C[3], f, k1, k2 -- the UNC_CAM_RADIAL model.
 * A truncated or malformed file or an index out of range: UNC_ERR_BAD_INPUT. */
int unc_scene_read_bal(const char *path, unc_scene **out);
int unc_scene_counts(const unc_scene *sc, int *num_cams, int *num_points, int *num_obs);
/* cams: num_cams x 9 (aa[3], C[3], f, k1, k2); points: num_points x 3;
 * obs_cam / obs_pt: num_obs indices.  Any pointer may be NULL. */
int unc_scene_arrays(const unc_scene *sc, double *cams, double *points, int *obs_cam,
                     int *obs_pt);
void unc_scene_destroy(unc_scene *sc);
/* JacobianComposer::findPts2Fix (JacobianComposer.cpp:63-83): the most distant
 * of 100 000 random distinct triples (triangle perimeter), first two ids sorted;
 * the seed is fixed (the reference seeds from std::random_device). */
int unc_find_pts2fix(int num_points, const double *points3D, int out[3]);
/* The CRS Jacobian in the .jacob layout: rows[2 num_obs + 1], cols and vals
 * [2 num_obs (cp + 3)], cp from cam_model.  Host arrays (ctx NULL: device 0);
 * the _device variant takes device pointers, enqueues on stream (NULL: ctx's
 * stream) and waits for the kernel, since it reports index errors. */
int unc_scene_jacobian(unc_context *ctx, int cam_model, int num_cams, int num_points,
                       int num_obs, const double *cams, const double *points,
                       const int *obs_cam, const int *obs_pt, int *rows, int *cols,
                       double *vals);
int unc_scene_jacobian_device(unc_context *ctx, int cam_model, int num_cams, int num_points,
                              int num_obs, const double *d_cams, const double *d_points,
                              const int *d_obs_cam, const int *d_obs_pt, int *d_rows,
                              int *d_cols, double *d_vals, void *stream);

const char *unc_error_string(int code);
const char *unc_stage_name(int stage);
int unc_abi_version(void);
/* Largest LDS allocation of one workgroup on the device (bytes), -1 on error. */
int unc_device_lds_bytes(int device);

/* Page-locked host memory (hipHostMalloc) for Jacobians and outputs that travel
 * over PCIe often, e.g. a server's request payloads: copies from it run at the
 * DMA engines' rate instead of being staged through pageable buffers.  No
 * counterpart in the reference.  UNC_ERR_OOM when the allocation fails. */
int unc_host_alloc(size_t bytes, void **ptr);
void unc_host_free(void *ptr);

/* ---- request server: many reconstructions of different sizes -----------------
 * No counterpart in the reference (which handles one problem per call).
 * unc_server_submit queues a request and returns a ticket at once; the host
 * buffers must stay valid until unc_server_wait returns for that ticket.  Slots
 * (each a context with its own HIP streams) run admitted requests concurrently; a
 * request is admitted when a slot is idle and the device bytes all slots then
 * hold stay within hbm_budget (see unc_workspace_bytes).  FIFO, with a later
 * request allowed to pass a head that does not fit yet for at most bypass_ms.
 * Results are those of unc_compute on the same inputs. */
typedef struct unc_server unc_server;
typedef struct unc_server_config {
  int device;          /* HIP device ordinal                                          */
  int slots;           /* requests in flight together, 1..64                          */
  size_t hbm_budget;   /* device bytes all slots may hold together                    */
  double bypass_ms;    /* how long later requests may pass a head that does not fit   */
  int flags;           /* reserved, 0                                                 */
} unc_server_config;
typedef struct unc_server_stats {
  long long queued, completed, admitted_out_of_order;
  size_t held_bytes, peak_held_bytes;
  int max_concurrent;
} unc_server_stats;
/* defaults: device 0, 2 slots, 64 GiB, 50 ms.  With more than 2 slots each slot
 * keeps the Z+ recursion on its own single stream (no second-stream fork), so
 * that the slots' streams stay within the device's four hardware queues. */
void unc_server_default_config(unc_server_config *cfg);
int unc_server_create(const unc_server_config *cfg, unc_server **out);
/* UNC_ERR_OOM when the request alone exceeds hbm_budget */
int unc_server_submit(unc_server *srv, const unc_options *opts, int num_rows, int num_cols,
                      const int *rows, const int *cols, const double *values,
                      double *cam_unc, double *pt_unc, unc_stats *stats,
                      long long *ticket);
/* blocks until the ticket's request finished; status = its unc_compute code,
 * latency = submit to finish, queue = submit to start (milliseconds).  Each
 * ticket is claimed by one wait: a second wait on it (also while the first
 * is still waiting) returns UNC_ERR_BAD_INPUT.  unc_server_destroy must not
 * run while a wait is in progress. */
int unc_server_wait(unc_server *srv, long long ticket, int *status, double *latency_ms,
                    double *queue_ms);
int unc_server_query(unc_server *srv, unc_server_stats *out);
/* finishes every queued request, then frees the slots */
void unc_server_destroy(unc_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* UNCERTAINTY_AMD_H */
