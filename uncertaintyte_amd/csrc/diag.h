/*
 * diag.h -- kernel-level diagnostics of the covariance library (test hooks).
 *
 * Built into a separate shared library (libuncertainty_amd_diag.so, the same
 * kernel objects plus diag.hip); the product library does not export these.
 * Every entry point works on device 0 with host pointers (column-major
 * matrices), allocates and frees its own device buffers, and keeps no state
 * between calls: kernel choices are arguments, never process-wide settings.
 * None has a counterpart in the reference.
 */
#ifndef UNCERTAINTY_AMD_DIAG_H
#define UNCERTAINTY_AMD_DIAG_H

#ifdef __cplusplus
extern "C" {
#endif

/* C = alpha A B + beta C with the FP64 MFMA GEMM; *absmax_out = max|C| from its
 * fused epilogue */
int unc_diag_gemm_f64(int M, int N, int K, double alpha, const double *A, int lda,
                      const double *B, int ldb, double beta, double *C, int ldc,
                      double *absmax_out);
/* C = alpha A B by the split-K form (two workgroups per tile over the halves of
 * K, then an epilogue adding them and folding max|C|); UNC_ERR_UNSUPPORTED where
 * the library would not split this shape (96-256 tiles of 128 x 128, K >= 1024) */
int unc_diag_gemm_f64_split(int M, int N, int K, double alpha, const double *A, int lda,
                            const double *B, int ldb, double *C, int ldc, double *absmax_out);
/* C = alpha A B for n x n A, B whose product is symmetric: the lower 128 x 128
 * tiles only, K in `parts` chunks (1..4; 0 = the library's rule), lower triangle
 * mirrored; UNC_ERR_UNSUPPORTED for shapes the library would not take this way */
int unc_diag_gemm_f64_sym(int n, double alpha, const double *A, const double *B, double *C,
                          double *absmax_out, int parts);
/* C_b = A B_b (b < nb <= 4; A and every B_b n x n symmetric, products symmetric)
 * by the batched lower-tile kernel of the Taylor chain in one launch.  B and C
 * hold nb matrices one after the other; absmax_out[b] = max|C_b| from the
 * fused epilogue.  avg_ms (may be NULL): average of `reps` further launches. */
int unc_diag_gemm_f64_sym_batch(int n, int nb, const double *A, const double *B, double *C,
                                double *absmax_out, int reps, double *avg_ms);
/* C = alpha A op(B) + beta C (op(B) = B^T when trans_b, B then N x K) with the
 * operand flags of the library's general product (1 A lower, 2 B lower, 4 B^T
 * lower, 8 lower C only): nmod = 0 by the FP64 MFMA product, 6..20 by the int8
 * CRT product with that many moduli.  avg_ms (may be NULL): average of `reps` further calls on
 * the device-resident operands. */
int unc_diag_gemm_ex(int M, int N, int K, double alpha, const double *A, int lda,
                     const double *B, int ldb, int trans_b, double beta, double *C, int ldc,
                     int tri, int nmod, int reps, double *avg_ms);
/* X = A^-1 (SPD A) by the blocked Gauss-Jordan kernels, plus the Newton step
 * with the error-free residual when refine != 0 */
int unc_diag_inverse(int n, const double *A, double *X, int refine);
/* R = I - A X by the refinement's residual kernel; tile_cols 16 or 32 forces the
 * 64 x 16 or 64 x 32 tiles, 0 takes the library's rule */
int unc_diag_residual(int n, const double *A, const double *X, double *R, int tile_cols);
/* R = I - A X for symmetric A (X any) by the int8 CRT residual of the Newton
 * step (residual_crt, 20 moduli), A cut in `blocks` row blocks (1..64);
 * avg_ms (may be NULL): average of `reps` more calls */
int unc_diag_residual_crt(int n, const double *A, const double *X, double *R, int blocks,
                          int reps, double *avg_ms);
/* X = L^-1 (upper triangle 0) for the Cholesky factor A = L L^T by the recursive
 * kernels of the Z^+ path; L holds the factor in the diagonal blocks of the
 * recursion's leaves and 128-row nodes only (elsewhere the values A had then) */
int unc_diag_chol_inverse(int n, const double *A, double *L, double *X);
/* average ms of one n x n x n product over reps launches on device-resident
 * operands: form 0 plain, 1 split-K, 2 symmetric lower-tile form with `parts`
 * K chunks (0 = rule) */
int unc_diag_gemm_bench(int n, int reps, int form, int parts, double *avg_ms);
/* average ms of the n x n residual (tile_cols as for unc_diag_residual) */
int unc_diag_residual_bench(int n, int reps, int tile_cols, double *avg_ms);
/* average us of one 64 x 64 leaf: 0 = Cholesky + L^-1 of the Z^+ recursion
 * (plus a 32 KB copy of its input), 1 = the Gauss-Jordan panel's SPD inverse */
int unc_diag_leaf_bench(int reps, int variant, double *avg_us);

#ifdef __cplusplus
}
#endif
#endif /* UNCERTAINTY_AMD_DIAG_H */
