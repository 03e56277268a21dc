// k_ozaki.hip -- the Newton step's residual R = I - A X on the int8 matrix
// cores (v_mfma_i32_16x16x64_i8), for symmetric A (A = Z + lambda I) and any X.
// Built into the diagnostics library only: at N = 1800 the per-level products
// re-read 5.9 GB of digit panels and run in 1.6 ms, slower than the product's
// error-free FP64 residual (1.3 ms); the declarations are local to diag.hip.
//
// Ozaki splitting (Ozaki, Ogita, Oishi, Rump, "Error-free transformations of
// matrix multiplication by using fast routines of matrix multiplication and
// its applications", Numer. Algorithms 59 (2012)), with int8 slices: every
// column c of M in {A, X} is scaled by 2^-e_c (max_k |M_kc| < 2^e_c) and cut
// into OZ_S signed digits,
//   M_kc = 2^e_c sum_s q_s(k, c) 2^-(7s - 1),   |q_s| <= 64,
// the first digit carrying 6 bits and every further one 7 (each cut is exact
// in FP64: a power-of-two scaling and the subtraction of the nearest integer).
// A is symmetric, so its rows are its columns and both operands are sliced the
// same way, k contiguous.  Then
//   (A X)_ij = 2^(e_i + f_j) sum_d 2^-(7d - 2) T_d(i, j),
//   T_d = sum_{s + t = d} A_s X_t^T   (int32, exact: |T_d| <= 10 n 64^2 < 2^31),
// and R = I - A X is accumulated in FP64 as ((I - T_2') - T_3') - ... with
// T_d' = 2^(e_i + f_j - 7d + 2) T_d exact: the partial sums shrink with d and
// stay on a grid of at most 53 bits, so the chain is exact unless R itself is
// large (then it rounds by eps |R_ij| at most).  Dropped are the digits past
// OZ_S (at most 2^-7 OZ_S per entry of either scaled operand) and the pairs
// with s + t > OZ_S + 1 (each term below 2^(14 - 7(s + t)), nine pairs at the
// first level): with A symmetric (X need not be),
//   |R - (I - A X)|_ij <= 12 n 2^(e_i + f_j - 7 OZ_S) + 2 eps |R_ij|.
// OZ_S = 10 (55 int8 products): on rings200 (N = 1800, cond ~1e7) the refined
// inverse agrees with the error-free residual's to 5e-13 (9 digits would leave
// 4e-11).
#include "common.h"
#include "gemm_tile.h"

namespace unc {
constexpr int RESIDUAL_OZ_MAX_N = 50000;
size_t residual_oz_work_bytes(int n);
hipError_t residual_oz(int n, const double *A, const double *X, double *R, void *work,
                       hipStream_t st);
}  // namespace unc

namespace unc {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int OZ_S = 10;          // digits per operand
constexpr int OZ_DMAX = OZ_S + 1;  // pairs with s + t <= OZ_DMAX
constexpr int OZ_TILE = 128;      // output tile (rows and columns)
constexpr int OZ_KB = 64;         // bytes of k per slab (one MFMA K)
constexpr int OZ_SUB = 2;         // slabs per step (one LDS buffer)
// LDS holds every 16-row (or column) block of a 64-byte slab as one 1 KB run
// in the order the MFMA reads it (slot l = lane l: row l % 16, bytes 16 (l / 16)),
// so a fragment read and a wave's store each touch 1 KB of consecutive bytes
constexpr int OZ_BLK = 1024;
constexpr int OZ_BUF = OZ_SUB * (OZ_TILE / 16) * OZ_BLK;  // one operand's step

size_t oz_np(int n) { return ((size_t)n + OZ_TILE - 1) / OZ_TILE * OZ_TILE; }
size_t oz_kp(int n) { return ((size_t)n + OZ_KB - 1) / OZ_KB * OZ_KB; }

// Digits of the columns of M (n x n, column-major), zero past n, with e[c] the
// column's exponent; cmax[c] = max_k |M_kc| as IEEE bits.  Digit s of column c
// and k sits where the product kernel's LDS wants it, in 1 KB runs: run
// (s, c / 16, k / 64) holds 16 columns x 64 k, byte 16 (16 (k % 64) / 16 + c % 16)
// + k % 16 -- a 16-column block's slab is then one contiguous 1 KB load.  One
// thread per 8 consecutive k of a column.
__global__ __launch_bounds__(256) void k_oz_slice(int n, int np, int kp,
                                                  const double *__restrict__ M,
                                                  const unsigned long long *__restrict__ cmax,
                                                  signed char *__restrict__ out, int *e_out) {
  const int c = blockIdx.y;
  const int k = 8 * (blockIdx.x * 256 + threadIdx.x);
  int e = 0;
  if (c < n) (void)frexp(__longlong_as_double((long long)cmax[c]), &e);
  if (blockIdx.x == 0 && threadIdx.x == 0) e_out[c] = e;
  if (k >= kp) return;
  unsigned long long w[OZ_S];
#pragma unroll
  for (int s = 0; s < OZ_S; ++s) w[s] = 0ull;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int kk = k + i;
    const double v = (c < n && kk < n) ? ldexp(M[(size_t)kk + (size_t)c * n], -e) : 0.0;
    double r = v * 64.0;
#pragma unroll
    for (int s = 0; s < OZ_S; ++s) {
      const double q = rint(r);
      r = (r - q) * 128.0;
      w[s] |= (unsigned long long)(unsigned char)(signed char)(int)q << (8 * i);
    }
  }
  const size_t runs_per_block = (size_t)kp / OZ_KB;
  const size_t run = (size_t)(c >> 4) * runs_per_block + (size_t)(k / OZ_KB);
  const int kk = k % OZ_KB;
  const int at = (((kk >> 4) * 16 + (c & 15)) << 4) + (kk & 15);
#pragma unroll
  for (int s = 0; s < OZ_S; ++s)
    *(unsigned long long *)(out + ((size_t)s * (np >> 4) * runs_per_block + run) * OZ_BLK + at) =
        w[s];
}

// T_d for one d and one 128 x 128 tile per workgroup (the grid holds every
// (d, tile), largest d first: d - 1 products each), 4 waves of 64 x 64 (4 x 4
// blocks of 16 x 16).  Steps (s, 128-byte k chunk) run through an LDS double
// buffer with the next step prefetched into registers; T_d leaves as int32,
// column-major with leading dimension np.
__global__ __launch_bounds__(256, 2) void k_oz_levels(int np, int kp,
                                                      const signed char *__restrict__ As,
                                                      const signed char *__restrict__ Xs,
                                                      int *__restrict__ T) {
  __shared__ __attribute__((aligned(16))) signed char La[2][OZ_BUF];
  __shared__ __attribute__((aligned(16))) signed char Lb[2][OZ_BUF];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int tiles = np / OZ_TILE, ntiles = tiles * tiles;
  // The tiles (in 4 x 4 blocks) fall into 8 contiguous groups; the order runs
  // group by group, each group's tiles once per d (largest d first), and each
  // XCD takes a contiguous eighth of it: one group's tiles at every level, so
  // the workgroups resident on an XCD share their digit panels in its L2 and
  // every XCD gets the same mix of levels.
  const int w = xcd_contiguous((int)blockIdx.x, (OZ_DMAX - 1) * ntiles);
  const int g = w / ((OZ_DMAX - 1) * ((ntiles + 7) / 8));
  const int g0 = g * ((ntiles + 7) / 8), gsz = min((ntiles + 7) / 8, ntiles - g0);
  const int r = w - g * (OZ_DMAX - 1) * ((ntiles + 7) / 8);
  const int d = OZ_DMAX - r / gsz;
  int tm, tn;
  full_tile_blocked(g0 + r % gsz, tiles, tiles, tm, tn);
  const int m0 = tm * OZ_TILE, n0 = tn * OZ_TILE;
  const int sub_total = kp / OZ_KB;                    // 64-byte k slabs
  const int nbig = (sub_total + OZ_SUB - 1) / OZ_SUB;  // steps per s
  // chunk p of this thread: c = tid + 256 p -> slab q = c / 512, row, 16-byte part lc
  struct Stage {
    v4i a[4], b[4];
  };
  const int nsteps = (d - 1) * nbig;  // step q: s = 1 + q / nbig, k chunk q % nbig
  // chunk c = tid + 256 p of a step: block rb = c / 128 of the tile's eight
  // 16-row blocks, slab sq = (c / 64) % 2, 16-byte slot c % 64 -- one wave moves
  // one 1 KB run, from global memory and into LDS alike
  const size_t runs_per_block = (size_t)sub_total;
  auto load = [&](int q, Stage &g) {
    const int s = 1 + q / nbig, kb = q - (s - 1) * nbig;
    const signed char *pa = As + ((size_t)(s - 1) * (np >> 4) + (m0 >> 4)) * runs_per_block * OZ_BLK;
    const signed char *pb = Xs + ((size_t)(d - s - 1) * (np >> 4) + (n0 >> 4)) * runs_per_block * OZ_BLK;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int c = tid + 256 * p, rb = c >> 7, sq = (c >> 6) & 1, slot = c & 63;
      const int sub = kb * OZ_SUB + sq;
      if (sub < sub_total) {
        const size_t off = ((size_t)rb * runs_per_block + sub) * OZ_BLK + 16 * slot;
        g.a[p] = *(const v4i *)(pa + off);
        g.b[p] = *(const v4i *)(pb + off);
      } else {
        g.a[p] = g.b[p] = (v4i){0, 0, 0, 0};
      }
    }
  };
  auto store = [&](int buf, const Stage &g) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int c = tid + 256 * p, rb = c >> 7, sq = (c >> 6) & 1, slot = c & 63;
      const int off = (sq * (OZ_TILE / 16) + rb) * OZ_BLK + 16 * slot;
      *(v4i *)(La[buf] + off) = g.a[p];
      *(v4i *)(Lb[buf] + off) = g.b[p];
    }
  };
  v4i iacc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) iacc[i][j] = (v4i){0, 0, 0, 0};
  auto compute = [&](int q, int buf) {
    const int kb = q % nbig;
    const int nsub = min(OZ_SUB, sub_total - kb * OZ_SUB);
#pragma unroll
    for (int sq = 0; sq < OZ_SUB; ++sq) {
      if (sq < nsub) {
        v4i af[4], bf[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
          af[i] = *(const v4i *)(La[buf] + (sq * (OZ_TILE / 16) + wm * 4 + i) * OZ_BLK + 16 * lane);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          bf[j] = *(const v4i *)(Lb[buf] + (sq * (OZ_TILE / 16) + wn * 4 + j) * OZ_BLK + 16 * lane);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            iacc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bf[j], iacc[i][j], 0, 0, 0);
      }
    }
  };
  // two steps in flight: step q is in LDS buffer q % 2 when it is multiplied,
  // step q + 1 waits in registers, step q + 2 is being loaded
  Stage st0, st1;
  load(0, st0);
  if (nsteps > 1) load(1, st1);
  store(0, st0);
  __syncthreads();
  for (int q = 0; q < nsteps; q += 2) {
    if (q + 2 < nsteps) load(q + 2, st0);
    compute(q, 0);
    if (q + 1 >= nsteps) break;
    store(1, st1);
    __syncthreads();
    if (q + 3 < nsteps) load(q + 3, st1);
    compute(q + 1, 1);
    if (q + 2 >= nsteps) break;
    store(0, st0);
    __syncthreads();
  }
  int *Td = T + (size_t)(d - 2) * np * np;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + wn * 64 + j * 16 + (lane & 15);
      const int row = m0 + wm * 64 + i * 16 + 4 * (lane >> 4);
      *(v4i *)(Td + (size_t)row + (size_t)col * np) = iacc[i][j];
    }
}

// R = ((I - T_2') - T_3') - ... - T_DMAX',  T_d' = 2^(e_i + f_j - 7d + 2) T_d
__global__ __launch_bounds__(256) void k_oz_fold(int n, int np, const int *__restrict__ T,
                                                 const int *__restrict__ ea,
                                                 const int *__restrict__ ex,
                                                 double *__restrict__ R) {
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  const int e = ea[i] + ex[j];
  const size_t nn = (size_t)np * np, at = (size_t)i + (size_t)j * np;
  double r = i == j ? 1.0 : 0.0;
#pragma unroll
  for (int d = 2; d <= OZ_DMAX; ++d) r = r - ldexp((double)T[(size_t)(d - 2) * nn + at], e - 7 * d + 2);
  R[(size_t)i + (size_t)j * n] = r;
}

}  // namespace

size_t residual_oz_work_bytes(int n) {
  if (n <= 0) return 0;
  const size_t np = oz_np(n);
  return 2 * (size_t)OZ_S * np * oz_kp(n) + (size_t)(OZ_DMAX - 1) * np * np * sizeof(int) +
         2 * np * sizeof(int) + 2 * (size_t)n * sizeof(unsigned long long) + 1024;
}

hipError_t residual_oz(int n, const double *A, const double *X, double *R, void *work,
                       hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (n > RESIDUAL_OZ_MAX_N) return hipErrorInvalidValue;
  if (!work || ((uintptr_t)work & 15) != 0) return hipErrorInvalidValue;
  const size_t np = oz_np(n), kp = oz_kp(n), dig = np * kp;
  signed char *As = (signed char *)work;
  signed char *Xs = As + OZ_S * dig;
  int *T = (int *)(Xs + OZ_S * dig);
  int *ea = T + (size_t)(OZ_DMAX - 1) * np * np;
  int *ex = ea + np;
  unsigned long long *amax = (unsigned long long *)(((uintptr_t)(ex + np) + 15) & ~(uintptr_t)15);
  unsigned long long *xmax = amax + n;
  // column maxima (A is symmetric: its row maxima)
  hipError_t e = col_absmax(n, A, amax, st);
  if (e != hipSuccess) return e;
  e = col_absmax(n, X, xmax, st);
  if (e != hipSuccess) return e;
  const dim3 sg((unsigned)cdiv((long long)kp / 8, 256), (unsigned)np);
  hipLaunchKernelGGL(k_oz_slice, sg, dim3(256), 0, st, n, (int)np, (int)kp, A, amax, As, ea);
  hipLaunchKernelGGL(k_oz_slice, sg, dim3(256), 0, st, n, (int)np, (int)kp, X, xmax, Xs, ex);
  const int tiles = (int)(np / OZ_TILE);
  hipLaunchKernelGGL(k_oz_levels, dim3((OZ_DMAX - 1) * tiles * tiles), dim3(256), 0, st, (int)np,
                     (int)kp, As, Xs, T);
  hipLaunchKernelGGL(k_oz_fold, dim3(cdiv(n, 256), n), dim3(256), 0, st, n, (int)np, T, ea, ex, R);
  return hipGetLastError();
}

}  // namespace unc
