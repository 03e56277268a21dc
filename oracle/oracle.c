/*
 * oracle.c -- CPU restatement of uncertaintyTE's TAYLOR_EXPANSION pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -ffp-contract=off:
 * every fused multiply-add below is an explicit fma(), so the GPU kernels can
 * reproduce the bit-exact stages operation for operation.
 *
 * Reference correspondence (paths relative to the reference checkout):
 *   orc_fixpts         compute.cpp:283-368            (fixPts)
 *   assemble_*         compute.cpp:76-100, ScaledSparseMatrix.cpp:334-358,
 *                      :384-388, :459-465, :238-299   (scaling, J^T J, split)
 *   inv3x3_ref         ScaledSparseMatrix.cpp:210-230 (inv3x3blockSymmDiag)
 *   schur              compute.cpp:107-112, ScaledSparseMatrix.cpp:554-641
 *   orc_taylor         compute.cpp:57-71, 121-166     (factorial, iZupdate, teInverse)
 *   cam_output         compute.cpp:168-183            (removeScaleJ4Z, 1.5x quirk)
 *   point_output       compute.cpp:370-455, 516-527   (point back-propagation)
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* small helpers                                                             */
/* ------------------------------------------------------------------------ */

int orc_factorial32(int n) {
  /* compute.cpp:57-62: int factorial(int n) { n>1 ? n*factorial(n-1) : 1 }.
   * Signed overflow is undefined in C, so the two's-complement wrap the
   * reference binary exhibits is written out with unsigned arithmetic. */
  uint32_t f = 1u;
  for (int i = 2; i <= n; ++i) f = f * (uint32_t)i;
  return (int)(int32_t)f;
}

double orc_auto_lambda(int N) {
  /* compute.cpp:125 */
  return pow(10, -1.2653 * log10((double)N) - 2.9415);
}

static void set_threads(const orc_opts *o) {
#ifdef _OPENMP
  if (o && o->nthreads > 0) omp_set_num_threads(o->nthreads);
#else
  (void)o;
#endif
}

/* ------------------------------------------------------------------------ */
/* fixPts  (compute.cpp:283-368)                                             */
/* ------------------------------------------------------------------------ */

int orc_fixpts(int nrows, int ncols, const int *rows, const int *cols,
               const double *vals, int cp, int ncams, int pts[3],
               int *out_nrows, int *out_ncols, int *out_rows, int *out_cols,
               double *out_vals) {
  if (pts[0] < 0 || pts[1] < 0 || pts[2] < 0) return ORC_ERR_NO_FIXED_POINTS;
  /* sort(pts, pts + 2): only the first two are ordered (compute.cpp:285) */
  if (pts[1] < pts[0]) { int t = pts[0]; pts[0] = pts[1]; pts[1] = t; }
  int camOffset = ncams * cp;
  int remCols[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) remCols[i * 3 + j] = camOffset + pts[i] * 3 + j;
  for (int i = 0; i < 9; ++i)
    if (remCols[i] < 0 || remCols[i] >= ncols) return ORC_ERR_BAD_INPUT;

  int *newIds = (int *)calloc((size_t)ncols, sizeof(int));
  if (!newIds) return ORC_ERR_OOM;
  for (int i = 0; i < 9; ++i) newIds[remCols[i]] = -1;
  int off = 0;
  for (int i = 0; i < ncols; ++i) { off += newIds[i]; newIds[i] = i + off; }

  int nnz = rows[nrows];
  int *rws = (int *)malloc(sizeof(int) * (size_t)(nrows + 1));
  int *cl = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  double *vl = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  char *use = (char *)malloc((size_t)(nnz > 0 ? nnz : 1));
  if (!rws || !cl || !vl || !use) { free(newIds); free(rws); free(cl); free(vl); free(use); return ORC_ERR_OOM; }

  /* decompose and compose jacobian (compute.cpp:322-335) */
  int cnt = 0;
  rws[0] = 0;
  for (int i = 0; i < nrows; ++i) {
    for (int j = rows[i]; j < rows[i + 1]; ++j) {
      int keep = 1;
      for (int k = 0; k < 9; ++k) if (cols[j] == remCols[k]) keep = 0;
      if (keep) { cl[cnt] = newIds[cols[j]]; vl[cnt] = vals[j]; ++cnt; }
    }
    rws[i + 1] = cnt;
  }
  /* remove the rows that lost their point columns (compute.cpp:337-352) */
  int err = 0, nu = 0;
  for (int i = 1; i <= nrows; ++i) {
    int u = 1;
    if (rws[i] != i * (cp + 3) - err) { err += 3; u = 0; }
    else { use[nu++] = 1; use[nu++] = 1; use[nu++] = 1; }
    for (int j = 0; j < cp; ++j) use[nu++] = (char)u;
  }
  if (nu != cnt) { free(newIds); free(rws); free(cl); free(vl); free(use); return ORC_ERR_BAD_INPUT; }
  int n2 = 0;
  for (int i = 0; i < cnt; ++i)
    if (use[i]) { out_cols[n2] = cl[i]; out_vals[n2] = vl[i]; ++n2; }
  int nr = n2 / (cp + 3) + 1; /* compute.cpp:360-363 */
  for (int i = 0; i < nr; ++i) out_rows[i] = i * (cp + 3);
  *out_nrows = nr - 1;
  *out_ncols = ncols - 9;
  free(newIds); free(rws); free(cl); free(vl); free(use);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* structure of the fixed Jacobian                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
  int cp, nc, npc, nobs, N;
  const double *vals;   /* 2*(cp+3) values per observation */
  int *obs_cam, *obs_pt;
  int *cam_off, *cam_list, *pt_off, *pt_list;
} jstruct;

static void jstruct_free(jstruct *s) {
  free(s->obs_cam); free(s->obs_pt); free(s->cam_off); free(s->cam_list);
  free(s->pt_off); free(s->pt_list);
  memset(s, 0, sizeof(*s));
}

/* Validate the uniform 2-rows-per-observation layout that compute.cpp:341,
 * :373-375 and :393-414 rely on, and group observations by point (row order
 * inside each point) and by camera (point order inside each camera: the
 * point-grouped list distributed stably by camera).  A camera that observes a
 * point twice is refused: the Schur order below has one observation per
 * camera and point. */
static int jstruct_build(jstruct *s, int cp, int nc, int npc, int nrows,
                         const int *cols, const double *vals) {
  memset(s, 0, sizeof(*s));
  int rl = cp + 3;
  if (nrows % 2) return ORC_ERR_BAD_INPUT;
  s->cp = cp; s->nc = nc; s->npc = npc; s->nobs = nrows / 2; s->N = nc * cp;
  s->vals = vals;
  int no = s->nobs;
  s->obs_cam = (int *)malloc(sizeof(int) * (size_t)(no + 1));
  s->obs_pt = (int *)malloc(sizeof(int) * (size_t)(no + 1));
  s->cam_off = (int *)calloc((size_t)nc + 1, sizeof(int));
  s->pt_off = (int *)calloc((size_t)npc + 1, sizeof(int));
  s->cam_list = (int *)malloc(sizeof(int) * (size_t)(no + 1));
  s->pt_list = (int *)malloc(sizeof(int) * (size_t)(no + 1));
  if (!s->obs_cam || !s->obs_pt || !s->cam_off || !s->pt_off || !s->cam_list || !s->pt_list) {
    jstruct_free(s); return ORC_ERR_OOM;
  }
  int N = s->N;
  for (int o = 0; o < no; ++o) {
    const int *c0 = cols + (size_t)(2 * o) * rl;
    const int *c1 = c0 + rl;
    int a = c0[0] / cp;
    if (c0[0] < 0 || c0[0] % cp || a >= nc) { jstruct_free(s); return ORC_ERR_BAD_INPUT; }
    int pc = c0[cp] - N;
    if (pc < 0 || pc % 3 || pc / 3 >= npc) { jstruct_free(s); return ORC_ERR_BAD_INPUT; }
    for (int j = 0; j < cp; ++j)
      if (c0[j] != a * cp + j || c1[j] != c0[j]) { jstruct_free(s); return ORC_ERR_BAD_INPUT; }
    for (int j = 0; j < 3; ++j)
      if (c0[cp + j] != N + pc + j || c1[cp + j] != c0[cp + j]) { jstruct_free(s); return ORC_ERR_BAD_INPUT; }
    s->obs_cam[o] = a;
    s->obs_pt[o] = pc / 3;
  }
  for (int o = 0; o < no; ++o) { s->cam_off[s->obs_cam[o] + 1]++; s->pt_off[s->obs_pt[o] + 1]++; }
  for (int a = 0; a < nc; ++a) s->cam_off[a + 1] += s->cam_off[a];
  for (int p = 0; p < npc; ++p) s->pt_off[p + 1] += s->pt_off[p];
  int *cc = (int *)malloc(sizeof(int) * (size_t)(nc + 1));
  int *pc = (int *)malloc(sizeof(int) * (size_t)(npc + 1));
  if (!cc || !pc) { free(cc); free(pc); jstruct_free(s); return ORC_ERR_OOM; }
  memcpy(cc, s->cam_off, sizeof(int) * (size_t)(nc + 1));
  memcpy(pc, s->pt_off, sizeof(int) * (size_t)(npc + 1));
  for (int o = 0; o < no; ++o) s->pt_list[pc[s->obs_pt[o]]++] = o;
  for (int t = 0; t < no; ++t) {
    int o = s->pt_list[t];
    s->cam_list[cc[s->obs_cam[o]]++] = o;
  }
  free(cc); free(pc);
  for (int a = 0; a < nc; ++a)
    for (int t = s->cam_off[a] + 1; t < s->cam_off[a + 1]; ++t)
      if (s->obs_pt[s->cam_list[t]] == s->obs_pt[s->cam_list[t - 1]]) {
        jstruct_free(s); return ORC_ERR_UNSUPPORTED;
      }
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* (a) assembly + column scaling + (b) 3x3 inverses                           */
/* ------------------------------------------------------------------------ */

/* inv3x3blockSymmDiag (ScaledSparseMatrix.cpp:210-230): cofactor inverse
 * reading only the upper entries, with the reference's evaluation order. */
static void inv3x3_ref(const double v[9], double o[9]) {
  double div = v[0] * v[4] * v[8] - v[0] * v[5] * v[5] - v[1] * v[1] * v[8] +
               2 * v[1] * v[2] * v[5] - v[2] * v[2] * v[4];
  o[0] = (v[4] * v[8] - v[5] * v[5]) / div;
  o[1] = (-v[1] * v[8] + v[2] * v[5]) / div;
  o[2] = (v[1] * v[5] - v[2] * v[4]) / div;
  o[3] = o[1];
  o[4] = (v[0] * v[8] - v[2] * v[2]) / div;
  o[5] = (-v[0] * v[5] + v[1] * v[2]) / div;
  o[6] = o[2];
  o[7] = o[5];
  o[8] = (v[0] * v[4] - v[1] * v[1]) / div;
}

typedef struct {
  double *Us;    /* nc*cp*cp */
  double *d;     /* N + 3*npc */
  double *Vi;    /* npc*9 */
  double *Ws;    /* nobs*cp*3 (row-major [i][k]) */
  double *Y;     /* nobs*cp*3 */
  char *pt_has;  /* npc */
} blocks;

static void blocks_free(blocks *b) {
  free(b->Us); free(b->d); free(b->Vi); free(b->Ws); free(b->Y); free(b->pt_has);
  memset(b, 0, sizeof(*b));
}

static int assemble(const jstruct *s, blocks *b) {
  int cp = s->cp, rl = cp + 3, N = s->N, nc = s->nc, npc = s->npc, no = s->nobs;
  memset(b, 0, sizeof(*b));
  b->Us = (double *)malloc(sizeof(double) * (size_t)nc * cp * cp);
  b->d = (double *)malloc(sizeof(double) * (size_t)(N + 3 * npc));
  b->Vi = (double *)calloc((size_t)npc * 9, sizeof(double));
  b->Ws = (double *)malloc(sizeof(double) * (size_t)no * cp * 3 + 8);
  b->Y = (double *)malloc(sizeof(double) * (size_t)no * cp * 3 + 8);
  b->pt_has = (char *)calloc((size_t)npc + 1, 1);
  if (!b->Us || !b->d || !b->Vi || !b->Ws || !b->Y || !b->pt_has) { blocks_free(b); return ORC_ERR_OOM; }
  int bad = 0;

  /* camera blocks: U_a[i][j] = sum over the camera's observations in point
   * order of fma(J0_i, J0_j, .), fma(J1_i, J1_j, .); d = 1/sqrt(U_ii) */
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
  for (int a = 0; a < nc; ++a) {
    double U[81];
    for (int i = 0; i < cp; ++i)
      for (int j = i; j < cp; ++j) {
        double acc = 0.0;
        for (int t = s->cam_off[a]; t < s->cam_off[a + 1]; ++t) {
          const double *r0 = s->vals + (size_t)s->cam_list[t] * 2 * rl;
          const double *r1 = r0 + rl;
          acc = fma(r0[i], r0[j], acc);
          acc = fma(r1[i], r1[j], acc);
        }
        U[i * cp + j] = acc;
        U[j * cp + i] = acc;
      }
    for (int i = 0; i < cp; ++i) {
      if (!(U[i * cp + i] > 0.0)) bad |= 1;
      b->d[a * cp + i] = 1.0 / sqrt(U[i * cp + i]);
    }
    for (int i = 0; i < cp; ++i)
      for (int j = 0; j < cp; ++j)
        b->Us[((size_t)a * cp + i) * cp + j] = (U[i * cp + j] * b->d[a * cp + i]) * b->d[a * cp + j];
  }
  if (bad) { blocks_free(b); return ORC_ERR_SINGULAR; }

  /* point blocks, their inverses, and the per-observation W_s / Y blocks */
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
  for (int p = 0; p < npc; ++p) {
    int t0 = s->pt_off[p], t1 = s->pt_off[p + 1];
    double *dp = b->d + N + 3 * p;
    if (t0 == t1) { dp[0] = dp[1] = dp[2] = 0.0; continue; }
    b->pt_has[p] = 1;
    double V[9];
    for (int k = 0; k < 3; ++k)
      for (int l = k; l < 3; ++l) {
        double acc = 0.0;
        for (int t = t0; t < t1; ++t) {
          const double *r0 = s->vals + (size_t)s->pt_list[t] * 2 * rl + cp;
          const double *r1 = r0 + rl;
          acc = fma(r0[k], r0[l], acc);
          acc = fma(r1[k], r1[l], acc);
        }
        V[k * 3 + l] = acc;
        V[l * 3 + k] = acc;
      }
    for (int k = 0; k < 3; ++k) {
      if (!(V[k * 3 + k] > 0.0)) bad |= 1;
      dp[k] = 1.0 / sqrt(V[k * 3 + k]);
    }
    double Vs[9];
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) Vs[k * 3 + l] = (V[k * 3 + l] * dp[k]) * dp[l];
    double *Vi = b->Vi + (size_t)p * 9;
    inv3x3_ref(Vs, Vi);
    for (int k = 0; k < 9; ++k)
      if (!isfinite(Vi[k])) bad |= 1;
    for (int t = t0; t < t1; ++t) {
      int o = s->pt_list[t];
      int a = s->obs_cam[o];
      const double *r0 = s->vals + (size_t)o * 2 * rl;
      const double *r1 = r0 + rl;
      const double *dc = b->d + a * cp;
      double *Ws = b->Ws + (size_t)o * cp * 3;
      double *Y = b->Y + (size_t)o * cp * 3;
      for (int i = 0; i < cp; ++i)
        for (int k = 0; k < 3; ++k) {
          double w = fma(r1[i], r1[cp + k], r0[i] * r0[cp + k]);
          Ws[i * 3 + k] = (w * dc[i]) * dp[k];
        }
      for (int i = 0; i < cp; ++i)
        for (int m = 0; m < 3; ++m) {
          double y = Ws[i * 3 + 0] * Vi[0 * 3 + m];
          y = fma(Ws[i * 3 + 1], Vi[1 * 3 + m], y);
          y = fma(Ws[i * 3 + 2], Vi[2 * 3 + m], y);
          Y[i * 3 + m] = y;
        }
    }
  }
  if (bad) { blocks_free(b); return ORC_ERR_SINGULAR; }
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* (c) Schur complement  Z_s = U_s - W_s V_s^-1 W_s^T  (dense, column-major)  */
/* ------------------------------------------------------------------------ */

/* Block (a, b >= a): the points p_0 < p_1 < ... seen by both cameras (the
 * merge of the two point-ordered camera lists) give S_ab = sum_k Y_a,p_k
 * W_b,p_k^T.  Match k goes to stream k mod SCHUR_STREAMS; each element of a
 * stream is one sequential chain of fma(Y[i][t], W[j][t], .) over t = 0, 1, 2
 * of its matches; then S = ((s_0 + s_1) + ...) + s_6.  Z = U_s - S on
 * diagonal blocks (upper triangle, mirrored), -S elsewhere, both (a, b) and
 * (b, a).  Cameras that share no point leave a zero block. */
#define SCHUR_STREAMS 7

static int schur(const jstruct *s, const blocks *bl, double *Z) {
  int cp = s->cp, N = s->N, nc = s->nc;
  memset(Z, 0, sizeof(double) * (size_t)N * N);
  int bad = 0;
#pragma omp parallel reduction(| : bad)
  {
    char *mark = (char *)calloc((size_t)nc + 1, 1);
    double *acc = (double *)malloc(sizeof(double) * SCHUR_STREAMS * 81);
    if (!mark || !acc) bad = 1;
#pragma omp for schedule(dynamic, 1)
    for (int a = 0; a < nc; ++a) {
      if (bad) continue;
      /* neighbours b >= a: cameras of the points of a */
      for (int t = s->cam_off[a]; t < s->cam_off[a + 1]; ++t) {
        int p = s->obs_pt[s->cam_list[t]];
        for (int u = s->pt_off[p]; u < s->pt_off[p + 1]; ++u) {
          int b = s->obs_cam[s->pt_list[u]];
          if (b >= a) mark[b] = 1;
        }
      }
      for (int b = a; b < nc; ++b) {
        if (!mark[b]) continue;
        mark[b] = 0;
        memset(acc, 0, sizeof(double) * SCHUR_STREAMS * 81);
        int ta = s->cam_off[a], tb = s->cam_off[b], k = 0;
        while (ta < s->cam_off[a + 1] && tb < s->cam_off[b + 1]) {
          int o1 = s->cam_list[ta], o2 = s->cam_list[tb];
          int p1 = s->obs_pt[o1], p2 = s->obs_pt[o2];
          if (p1 < p2) { ++ta; continue; }
          if (p2 < p1) { ++tb; continue; }
          const double *Y = bl->Y + (size_t)o1 * cp * 3;
          const double *W = bl->Ws + (size_t)o2 * cp * 3;
          double *sg = acc + (k % SCHUR_STREAMS) * 81;
          for (int i = 0; i < cp; ++i)
            for (int j = 0; j < cp; ++j) {
              double v = sg[i * 9 + j];
              v = fma(Y[i * 3 + 0], W[j * 3 + 0], v);
              v = fma(Y[i * 3 + 1], W[j * 3 + 1], v);
              v = fma(Y[i * 3 + 2], W[j * 3 + 2], v);
              sg[i * 9 + j] = v;
            }
          ++k; ++ta; ++tb;
        }
        for (int i = 0; i < cp; ++i)
          for (int j = (a == b ? i : 0); j < cp; ++j) {
            double sv = acc[i * 9 + j];
            for (int g = 1; g < SCHUR_STREAMS; ++g) sv = sv + acc[g * 81 + i * 9 + j];
            double zv = (a == b) ? bl->Us[((size_t)a * cp + i) * cp + j] - sv : -sv;
            size_t r = (size_t)a * cp + i, c = (size_t)b * cp + j;
            Z[r + c * N] = zv;
            Z[c + r * N] = zv;
          }
      }
    }
    free(mark);
    free(acc);
  }
  return bad ? ORC_ERR_OOM : ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* (d) lambda-regularised inverse + Taylor chain (compute.cpp:121-166)        */
/* ------------------------------------------------------------------------ */

static double absmax(size_t n, const double *A) {
  double m = DBL_MIN; /* ScaledDenseMatrix.cpp:208 */
  for (size_t i = 0; i < n; ++i) {
    double v = fabs(A[i]);
    if (m < v) m = v;
  }
  return m;
}

int orc_taylor(const orc_opts *opts, int N, const double *Z, double lambda,
               double *X0_out, double *X_out, orc_stats *st) {
  set_threads(opts);
  size_t nn = (size_t)N * N;
  double *A = (double *)malloc(sizeof(double) * nn);
  double *X0 = (double *)malloc(sizeof(double) * nn);
  double *T = (double *)malloc(sizeof(double) * nn);
  double *T2 = (double *)malloc(sizeof(double) * nn);
  if (!A || !X0 || !T || !T2) { free(A); free(X0); free(T); free(T2); return ORC_ERR_OOM; }
  memcpy(A, Z, sizeof(double) * nn);
  for (int i = 0; i < N; ++i) A[i + (size_t)i * N] = A[i + (size_t)i * N] + lambda;

  int rc;
  if (opts && opts->inverse_mode == ORC_INV_REFERENCE) {
    rc = orc_inv_equilibrated(N, A, X0);
  } else {
    rc = orc_inv_lu(N, A, X0);
    if (rc == ORC_OK) {
      /* one Newton step X1 = X0 + X0 (I - A X0) with the residual formed in
       * twice the working precision, then exact symmetrisation */
      orc_residual_dd(N, A, X0, T);
      orc_gemm(N, X0, T, T2);
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < nn; ++i) T2[i] = X0[i] + T2[i];
#pragma omp parallel for schedule(static)
      for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r)
          X0[r + (size_t)c * N] = 0.5 * (T2[r + (size_t)c * N] + T2[c + (size_t)r * N]);
    }
  }
  if (rc != ORC_OK) { free(A); free(X0); free(T); free(T2); return rc; }
  if (X0_out) memcpy(X0_out, X0, sizeof(double) * nn);

  /* TE: X = X0; T = X0*X0; for i=1..19 (compute.cpp:142-160) */
  memcpy(X_out, X0, sizeof(double) * nn);
  orc_gemm(N, X0, X0, T);
  double old_change = DBL_MAX;
  int nterms = 0;
  for (int i = 1; i < 20; ++i) {
    double k = pow(lambda, i) / (double)orc_factorial32(i - 1);
    double change = fabs(k * absmax(nn, T));
    if (change < 1e-5 || change > old_change) break;
    old_change = change;
    if (st) st->cycle_change[nterms] = change;
    ++nterms;
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < nn; ++e) X_out[e] = X_out[e] + k * T[e];
    orc_gemm(N, T, X0, T2);
    double *tmp = T; T = T2; T2 = tmp;
  }
  if (st) { st->n_terms = nterms; st->lambda = lambda; }
  free(A); free(X0); free(T); free(T2);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* outputs                                                                   */
/* ------------------------------------------------------------------------ */

/* removeScaleJ4Z (compute.cpp:168-183): (X[j,k] + X[k,j]/2) d_j d_k */
static void cam_output(int nc, int cp, int N, const double *X, const double *d, double *out) {
  int per = cp * (cp + 1) / 2;
#pragma omp parallel for schedule(static)
  for (int a = 0; a < nc; ++a) {
    int l = a * per;
    int st = a * cp;
    for (int j = st; j < st + cp; ++j)
      for (int k = j; k < st + cp; ++k)
        out[l++] = ((X[j + (size_t)k * N] + X[k + (size_t)j * N] / 2) * d[j]) * d[k];
  }
}

/* Point covariance, the V^-1 form of compute.cpp:516-527:
 *   Sigma_p = D_p [Vi + Vi (sum_o W_s,o^T X_cc W_s,o) Vi] D_p
 * equal to pinv(A) B Sigma B^T pinv(A)^T with A = J_p (full column rank),
 * B = [J_c, -I2] per observation and Sigma = blockdiag(X_cc unscaled, I2);
 * only each camera's own block enters (compute.cpp:422-441). */
/* Camera a's own block of X: element (i, j) at X[a * cam_stride + i + j * ld]
 * (the dense N x N column-major X: cam_stride = cp (N + 1), ld = N; compact
 * cp x cp column-major blocks: cam_stride = cp * cp, ld = cp). */
static void point_one(int cp, int N, const jstruct *s, const blocks *bl,
                      const double *X, size_t cam_stride, size_t ld, int p, double out6[6]) {
  double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = s->pt_off[p]; t < s->pt_off[p + 1]; ++t) {
    int o = s->pt_list[t];
    int a = s->obs_cam[o];
    const double *Ws = bl->Ws + (size_t)o * cp * 3;
    const double *Xa = X + (size_t)a * cam_stride;
    double H[27];
    for (int i = 0; i < cp; ++i)
      for (int k = 0; k < 3; ++k) {
        double h = Xa[i + 0 * ld] * Ws[0 * 3 + k];
        for (int j = 1; j < cp; ++j) h = fma(Xa[i + (size_t)j * ld], Ws[j * 3 + k], h);
        H[i * 3 + k] = h;
      }
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) {
        double m = Ws[0 * 3 + k] * H[0 * 3 + l];
        for (int i = 1; i < cp; ++i) m = fma(Ws[i * 3 + k], H[i * 3 + l], m);
        G[k * 3 + l] = G[k * 3 + l] + m;
      }
  }
  const double *Vi = bl->Vi + (size_t)p * 9;
  double R[9], Q[9];
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l) {
      double r = Vi[k * 3 + 0] * G[0 * 3 + l];
      r = fma(Vi[k * 3 + 1], G[1 * 3 + l], r);
      r = fma(Vi[k * 3 + 2], G[2 * 3 + l], r);
      R[k * 3 + l] = r;
    }
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l) {
      double q = R[k * 3 + 0] * Vi[0 * 3 + l];
      q = fma(R[k * 3 + 1], Vi[1 * 3 + l], q);
      q = fma(R[k * 3 + 2], Vi[2 * 3 + l], q);
      Q[k * 3 + l] = Vi[k * 3 + l] + q;
    }
  const double *dp = bl->d + N + 3 * p;
  out6[0] = (Q[0] * dp[0]) * dp[0];
  out6[1] = (Q[1] * dp[0]) * dp[1];
  out6[2] = (Q[2] * dp[0]) * dp[2];
  out6[3] = (Q[4] * dp[1]) * dp[1];
  out6[4] = (Q[5] * dp[1]) * dp[2];
  out6[5] = (Q[8] * dp[2]) * dp[2];
}

/* fillPtUnc2OutArray (compute.cpp:443-455): compact id -> output id using the
 * partially sorted triple, one increment per fixed id passed. */
static int out_point_id(int p, const int fix[3]) {
  for (int i = 0; i < 3; ++i)
    if (p >= fix[i]) ++p;
  return p;
}

static void point_output(const jstruct *s, const blocks *bl, const double *X,
                         const int fix[3], int npts, double *pt_out) {
  memset(pt_out, 0, sizeof(double) * 6 * (size_t)npts);
#pragma omp parallel for schedule(dynamic, 256)
  for (int p = 0; p < s->npc; ++p) {
    if (!bl->pt_has[p]) continue;
    double o6[6];
    point_one(s->cp, s->N, s, bl, X, (size_t)s->cp * ((size_t)s->N + 1), (size_t)s->N, p, o6);
    int q = out_point_id(p, fix);
    if (q >= 0 && q < npts) memcpy(pt_out + (size_t)q * 6, o6, sizeof(o6));
  }
}

/* ------------------------------------------------------------------------ */
/* whole pipeline                                                            */
/* ------------------------------------------------------------------------ */

static int validate_original(int ncams, int cp, int npts, int nrows, int ncols,
                             const int *rows) {
  if (cp <= 0 || cp > 9 || ncams <= 0 || npts < 4) return ORC_ERR_BAD_INPUT;
  if (ncols != ncams * cp + 3 * npts) return ORC_ERR_BAD_INPUT;
  if (nrows <= 0 || nrows % 2) return ORC_ERR_BAD_INPUT;
  for (int i = 0; i <= nrows; ++i)
    if (rows[i] != i * (cp + 3)) return ORC_ERR_BAD_INPUT;
  return ORC_OK;
}

int orc_pipeline(const orc_opts *opts, int ncams, int cp, int npts,
                 const int fix_in[3], int nrows, int ncols, const int *rows,
                 const int *cols, const double *vals, double *cam_out,
                 double *pt_out, orc_stats *st, orc_debug *dbg) {
  set_threads(opts);
  int rc = validate_original(ncams, cp, npts, nrows, ncols, rows);
  if (rc) return rc;
  if (!fix_in || fix_in[0] < 0 || fix_in[1] < 0 || fix_in[2] < 0) return ORC_ERR_NO_FIXED_POINTS;
  int fix[3] = {fix_in[0], fix_in[1], fix_in[2]};
  for (int i = 0; i < 3; ++i)
    if (fix[i] >= npts) return ORC_ERR_BAD_INPUT;
  if (fix[0] == fix[1] || fix[1] == fix[2] || fix[0] == fix[2]) return ORC_ERR_BAD_INPUT;

  int nnz = rows[nrows];
  int *r2 = (int *)malloc(sizeof(int) * (size_t)(nrows + 1));
  int *c2 = (int *)malloc(sizeof(int) * (size_t)nnz);
  double *v2 = (double *)malloc(sizeof(double) * (size_t)nnz);
  if (!r2 || !c2 || !v2) { free(r2); free(c2); free(v2); return ORC_ERR_OOM; }
  int nr2, nc2;
  rc = orc_fixpts(nrows, ncols, rows, cols, vals, cp, ncams, fix, &nr2, &nc2, r2, c2, v2);
  if (rc) { free(r2); free(c2); free(v2); return rc; }

  jstruct s;
  rc = jstruct_build(&s, cp, ncams, npts - 3, nr2, c2, v2);
  if (rc) { free(r2); free(c2); free(v2); return rc; }
  blocks bl;
  rc = assemble(&s, &bl);
  if (rc) { jstruct_free(&s); free(r2); free(c2); free(v2); return rc; }

  int N = s.N;
  size_t nn = (size_t)N * N;
  double *Z = (double *)malloc(sizeof(double) * nn);
  double *X = (double *)malloc(sizeof(double) * nn);
  if (!Z || !X) { free(Z); free(X); blocks_free(&bl); jstruct_free(&s); free(r2); free(c2); free(v2); return ORC_ERR_OOM; }
  rc = schur(&s, &bl, Z);
  if (rc) { free(Z); free(X); blocks_free(&bl); jstruct_free(&s); free(r2); free(c2); free(v2); return rc; }
  double lambda = (opts && opts->lambda != -1) ? opts->lambda : orc_auto_lambda(N);
  orc_stats local;
  memset(&local, 0, sizeof(local));
  rc = orc_taylor(opts, N, Z, lambda, dbg ? dbg->X0 : NULL, X, &local);
  if (rc == ORC_OK) {
    cam_output(ncams, cp, N, X, bl.d, cam_out);
    point_output(&s, &bl, X, fix, npts, pt_out);
    if (dbg) {
      if (dbg->U_s) memcpy(dbg->U_s, bl.Us, sizeof(double) * (size_t)ncams * cp * cp);
      if (dbg->Vinv_s) memcpy(dbg->Vinv_s, bl.Vi, sizeof(double) * (size_t)s.npc * 9);
      if (dbg->d) memcpy(dbg->d, bl.d, sizeof(double) * (size_t)(N + 3 * s.npc));
      if (dbg->Z) memcpy(dbg->Z, Z, sizeof(double) * nn);
      if (dbg->X) memcpy(dbg->X, X, sizeof(double) * nn);
    }
    if (st) {
      *st = local;
      st->fixed[0] = fix[0]; st->fixed[1] = fix[1]; st->fixed[2] = fix[2];
      st->n_obs_kept = s.nobs;
      st->n_pts_compact = s.npc;
    }
  }
  free(Z); free(X); blocks_free(&bl); jstruct_free(&s); free(r2); free(c2); free(v2);
  return rc;
}

int orc_front(const orc_opts *opts, int ncams, int cp, int npts, const int fix_in[3],
              int nrows, int ncols, const int *rows, const int *cols, const double *vals,
              double *Z_out, double *d_out) {
  set_threads(opts);
  int rc = validate_original(ncams, cp, npts, nrows, ncols, rows);
  if (rc) return rc;
  if (!fix_in || fix_in[0] < 0 || fix_in[1] < 0 || fix_in[2] < 0) return ORC_ERR_NO_FIXED_POINTS;
  int fix[3] = {fix_in[0], fix_in[1], fix_in[2]};
  for (int i = 0; i < 3; ++i)
    if (fix[i] >= npts) return ORC_ERR_BAD_INPUT;
  if (fix[0] == fix[1] || fix[1] == fix[2] || fix[0] == fix[2]) return ORC_ERR_BAD_INPUT;
  int nnz = rows[nrows];
  int *r2 = (int *)malloc(sizeof(int) * (size_t)(nrows + 1));
  int *c2 = (int *)malloc(sizeof(int) * (size_t)nnz);
  double *v2 = (double *)malloc(sizeof(double) * (size_t)nnz);
  if (!r2 || !c2 || !v2) { free(r2); free(c2); free(v2); return ORC_ERR_OOM; }
  int nr2, nc2;
  rc = orc_fixpts(nrows, ncols, rows, cols, vals, cp, ncams, fix, &nr2, &nc2, r2, c2, v2);
  if (rc) { free(r2); free(c2); free(v2); return rc; }
  jstruct s;
  rc = jstruct_build(&s, cp, ncams, npts - 3, nr2, c2, v2);
  if (rc) { free(r2); free(c2); free(v2); return rc; }
  blocks bl;
  rc = assemble(&s, &bl);
  if (rc) { jstruct_free(&s); free(r2); free(c2); free(v2); return rc; }
  rc = schur(&s, &bl, Z_out);
  if (rc) { blocks_free(&bl); jstruct_free(&s); free(r2); free(c2); free(v2); return rc; }
  if (d_out) memcpy(d_out, bl.d, sizeof(double) * (size_t)(s.N + 3 * s.npc));
  blocks_free(&bl); jstruct_free(&s); free(r2); free(c2); free(v2);
  return ORC_OK;
}

int orc_points_from_X(int ncams, int cp, int npts, const int fix_in[3],
                      int nrows_fixed, const int *cols_fixed,
                      const double *vals_fixed, const double *X, double *pt_out) {
  int fix[3] = {fix_in[0], fix_in[1], fix_in[2]};
  if (fix[1] < fix[0]) { int t = fix[0]; fix[0] = fix[1]; fix[1] = t; }
  jstruct s;
  int rc = jstruct_build(&s, cp, ncams, npts - 3, nrows_fixed, cols_fixed, vals_fixed);
  if (rc) return rc;
  blocks bl;
  rc = assemble(&s, &bl);
  if (rc) { jstruct_free(&s); return rc; }
  point_output(&s, &bl, X, fix, npts, pt_out);
  blocks_free(&bl);
  jstruct_free(&s);
  return ORC_OK;
}

int orc_points_from_blocks(int ncams, int cp, int npts, const int fix_in[3],
                           int nrows_fixed, const int *cols_fixed, const double *vals_fixed,
                           const double *Xd, const char *cam_ok, double *pt_out,
                           char *pt_done) {
  int fix[3] = {fix_in[0], fix_in[1], fix_in[2]};
  if (fix[1] < fix[0]) { int t = fix[0]; fix[0] = fix[1]; fix[1] = t; }
  jstruct s;
  int rc = jstruct_build(&s, cp, ncams, npts - 3, nrows_fixed, cols_fixed, vals_fixed);
  if (rc) return rc;
  blocks bl;
  rc = assemble(&s, &bl);
  if (rc) { jstruct_free(&s); return rc; }
  memset(pt_out, 0, sizeof(double) * 6 * (size_t)npts);
  memset(pt_done, 0, (size_t)npts);
#pragma omp parallel for schedule(dynamic, 256)
  for (int p = 0; p < s.npc; ++p) {
    if (!bl.pt_has[p]) continue;
    int ok = 1;
    for (int t = s.pt_off[p]; t < s.pt_off[p + 1] && ok; ++t) ok = cam_ok[s.obs_cam[s.pt_list[t]]] != 0;
    if (!ok) continue;
    double o6[6];
    point_one(cp, s.N, &s, &bl, Xd, (size_t)cp * cp, (size_t)cp, p, o6);
    int q = out_point_id(p, fix);
    if (q >= 0 && q < npts) {
      memcpy(pt_out + (size_t)q * 6, o6, sizeof(o6));
      pt_done[q] = 1;
    }
  }
  blocks_free(&bl);
  jstruct_free(&s);
  return ORC_OK;
}

int orc_jtj_dense(int nrows, int ncols, const int *rows, const int *cols,
                  const double *vals, double *out) {
  memset(out, 0, sizeof(double) * (size_t)ncols * ncols);
  for (int r = 0; r < nrows; ++r)
    for (int a = rows[r]; a < rows[r + 1]; ++a)
      for (int b = rows[r]; b < rows[r + 1]; ++b) {
        size_t e = (size_t)cols[a] + (size_t)cols[b] * ncols;
        out[e] = fma(vals[a], vals[b], out[e]);
      }
  return ORC_OK;
}
