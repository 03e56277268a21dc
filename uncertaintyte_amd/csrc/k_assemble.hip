// k_assemble.hip -- (a) Jacobian decode + J^T J block assembly with column
// scaling, and (b) batched FP64 3x3 point-block inverses, for gfx950.
//
// Reference correspondence:
//   decode / fixPts row dropping      compute.cpp:283-368, :370-380
//   column scale d = 1/||J_:,j||      compute.cpp:76-83, ScaledSparseMatrix.cpp:334-358
//   U, W, V blocks of J_s^T J_s       ScaledSparseMatrix.cpp:384-388, :459-465, :238-299
//   3x3 cofactor inverse              ScaledSparseMatrix.cpp:210-230
//
// Arithmetic contract (built with -ffp-contract=off): each block entry is a
// sequential chain of explicit fma() over the group's observations -- a
// camera's in point order, a point's in row order -- exactly the order of the
// CPU oracle, so U_s, d, V_s^-1 and W_s are bit-identical to it.  Scaling is applied to the assembled blocks
// (M_s = D M D) rather than to J, so J is read once.
#include "common.h"

namespace unc {

// ---------------------------------------------------------------------------
// decode: one thread per observation (two CRS rows)
// ---------------------------------------------------------------------------
// error bits, fixed-point drop and compact point id of observation o
__device__ __forceinline__ void decode_finish(const Structure &s, int o, unsigned int e, int a,
                                              int p) {
  if (e) {
    atomicOr(s.err, e);
    s.obs_cam[o] = -1;
    s.obs_pt[o] = -1;
    return;
  }
  const bool dropped = (p == s.fix[0] || p == s.fix[1] || p == s.fix[2]);
  if (dropped) {
    s.obs_cam[o] = -1;
    s.obs_pt[o] = -1;
    return;
  }
  // compact id after removing the three fixed points' columns (newIds, compute.cpp:301-311)
  int pcmp = p - (s.fix[0] < p) - (s.fix[1] < p) - (s.fix[2] < p);
  s.obs_cam[o] = a;
  s.obs_pt[o] = pcmp;
  atomicAdd(&s.pt_cnt[pcmp], 1);
}

// the first camera column cam0 and first point column pc (minus N) of an
// observation must be block starts of an existing camera and point
__device__ __forceinline__ unsigned int decode_range(const Structure &s, int cam0, int pc, int &a,
                                                     int &p) {
  unsigned int e = 0;
  a = cam0 / s.cp;
  if (cam0 < 0 || cam0 % s.cp != 0 || a >= s.nc) e |= ERRBIT_RANGE;
  p = pc / 3;
  if (pc < 0 || pc % 3 != 0 || p >= s.np) e |= ERRBIT_RANGE;
  return e;
}

__global__ void k_decode(Structure s) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= s.nobs) return;
  const int cp = s.cp, rl = cp + 3, N = s.N;
  const int r0 = 2 * o;
  unsigned int e = 0;
  if (s.rows[r0] != r0 * rl || s.rows[r0 + 1] != (r0 + 1) * rl || s.rows[r0 + 2] != (r0 + 2) * rl)
    e |= ERRBIT_LAYOUT;
  int a = -1, p = -1;
  if (!e) {
    const int *c0 = s.cols + (size_t)r0 * rl;
    const int *c1 = c0 + rl;
    const int cam0 = c0[0], pc = c0[cp] - N;
    e |= decode_range(s, cam0, pc, a, p);
    if (!e) {
      for (int j = 0; j < cp; ++j)
        if (c0[j] != cam0 + j || c1[j] != cam0 + j) e |= ERRBIT_LAYOUT;
      for (int j = 0; j < 3; ++j)
        if (c0[cp + j] != N + pc + j || c1[cp + j] != N + pc + j) e |= ERRBIT_LAYOUT;
    }
  }
  decode_finish(s, o, e, a, p);
}

// the same from the host-checked ids (the layout already holds)
__global__ void k_decode_ids(Structure s) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= s.nobs) return;
  int a = -1, p = -1;
  const unsigned int e = decode_range(s, s.ids[2 * o], s.ids[2 * o + 1], a, p);
  decode_finish(s, o, e, a, p);
}

__global__ void k_init_lists(Structure s) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < s.npc) { s.pt_cnt[i] = 0; s.pt_cursor[i] = 0; }
  if (i == 0) s.err[0] = 0u;
}

hipError_t launch_decode(const Structure &s, hipStream_t st) {
  int n = s.npc > s.nc ? s.npc : s.nc;
  hipLaunchKernelGGL(k_init_lists, dim3(cdiv(n + 1, 256)), dim3(256), 0, st, s);
  if (s.nobs > 0) {
    if (s.ids)
      hipLaunchKernelGGL(k_decode_ids, dim3(cdiv(s.nobs, 256)), dim3(256), 0, st, s);
    else
      hipLaunchKernelGGL(k_decode, dim3(cdiv(s.nobs, 256)), dim3(256), 0, st, s);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// exclusive scan of pt_cnt -> pt_off (npc + 1 entries), three-phase
// ---------------------------------------------------------------------------
constexpr int SCAN_BLOCK = 1024;  // elements per block (256 threads x 4)

__device__ inline int block_exclusive_scan_256(int v, int *sh, int *total) {
  // Hillis-Steele inclusive scan in LDS over 256 values.
  int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int x = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  int incl = sh[t];
  if (total) *total = sh[255];
  __syncthreads();
  return incl - v;
}

__global__ void k_scan_blocks(const int *in, int n, int *out, int *block_sums) {
  __shared__ int sh[256];
  int base = blockIdx.x * SCAN_BLOCK + threadIdx.x * 4;
  int v[4];
  int sum = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    sum += v[k];
  }
  int tot;
  int ex = block_exclusive_scan_256(sum, sh, &tot);
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ void k_scan_sums(int *block_sums, int nb, int *total_out) {
  __shared__ int sh[256];
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += 256) {
    int i = b0 + threadIdx.x;
    int v = (i < nb) ? block_sums[i] : 0;
    int tot;
    int ex = block_exclusive_scan_256(v, sh, &tot);
    if (i < nb) block_sums[i] = ex + carry;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ void k_scan_add(int *out, int n, const int *block_sums) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += block_sums[i / SCAN_BLOCK];
}

__global__ void k_fill_points(Structure s) {
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= s.nobs) return;
  int p = s.obs_pt[o];
  if (p < 0) return;
  int pos = s.pt_off[p] + atomicAdd(&s.pt_cursor[p], 1);
  s.pt_list[pos] = o;
}

// atomics place observations in arbitrary order inside a point's segment;
// sorting each (short) segment by observation index restores row order
__global__ void k_sort_segments(Structure s) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= s.npc) return;
  int t0 = s.pt_off[p], t1 = s.pt_off[p + 1];
  for (int i = t0 + 1; i < t1; ++i) {
    int v = s.pt_list[i];
    int j = i - 1;
    while (j >= t0 && s.pt_list[j] > v) {
      s.pt_list[j + 1] = s.pt_list[j];
      --j;
    }
    s.pt_list[j + 1] = v;
  }
}

static void exclusive_scan(const int *cnt, int n, int *off, int *scan_tmp, hipStream_t st) {
  int nb = (int)cdiv(n, SCAN_BLOCK);
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, st, cnt, n, off, scan_tmp);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, st, scan_tmp, nb, off + n);
  hipLaunchKernelGGL(k_scan_add, dim3(cdiv(n, 256)), dim3(256), 0, st, off, n, scan_tmp);
}


hipError_t launch_point_lists(const Structure &s, int *scan_tmp, hipStream_t st) {
  int n = s.npc;
  exclusive_scan(s.pt_cnt, n, s.pt_off, scan_tmp, st);
  if (s.nobs > 0) hipLaunchKernelGGL(k_fill_points, dim3(cdiv(s.nobs, 256)), dim3(256), 0, st, s);
  hipLaunchKernelGGL(k_sort_segments, dim3(cdiv(n, 256)), dim3(256), 0, st, s);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// camera blocks: one wave per camera, lane = upper-triangle entry (i, j), the
// camera's observations in point order
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_cam_assemble(Structure s, Blocks b) {
  // 64 observations at a time are staged in LDS (lane t loads observation t's
  // two camera rows), then every lane runs its fma chain over them in order.
  __shared__ double stage[64][18];
  __shared__ double U[81];
  __shared__ double dsh[9];
  const int a = blockIdx.x, l = threadIdx.x;
  const int cp = s.cp, rl = cp + 3;
  const int npair = cp * (cp + 1) / 2;
  int i = 0, j = 0;
  {
    int r = l;
    for (int ii = 0; ii < cp; ++ii) {
      int len = cp - ii;
      if (r < len) { i = ii; j = ii + r; break; }
      r -= len;
    }
  }
  const int t0 = s.cam_off[a], t1 = s.cam_off[a + 1];
  double acc = 0.0;
  for (int tb = t0; tb < t1; tb += 64) {
    const int nb = min(64, t1 - tb);
    if (l < nb) {
      const double *r0 = s.vals + (size_t)s.cam_list[tb + l] * 2 * rl;
      for (int q = 0; q < cp; ++q) {
        stage[l][q] = r0[q];
        stage[l][9 + q] = r0[rl + q];
      }
    }
    __syncthreads();
    if (l < npair) {
      for (int q = 0; q < nb; ++q) {
        acc = fma(stage[q][i], stage[q][j], acc);
        acc = fma(stage[q][9 + i], stage[q][9 + j], acc);
      }
    }
    __syncthreads();
  }
  if (l < npair) {
    U[i * cp + j] = acc;
    U[j * cp + i] = acc;
  }
  __syncthreads();
  if (l < cp) {
    double u = U[l * cp + l];
    if (!(u > 0.0)) atomicOr(s.err, ERRBIT_ZERO_COLUMN);
    double dv = 1.0 / sqrt(u);
    dsh[l] = dv;
    b.d[a * cp + l] = dv;
  }
  __syncthreads();
  for (int e = l; e < cp * cp; e += 64) {
    int ii = e / cp, jj = e % cp;
    b.Us[(size_t)a * cp * cp + e] = (U[e] * dsh[ii]) * dsh[jj];
  }
}

// inv3x3blockSymmDiag (ScaledSparseMatrix.cpp:210-230) with the reference's
// expression order; reads the upper entries only.
__device__ inline void inv3x3_ref(const double v[9], double o[9]) {
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

// ---------------------------------------------------------------------------
// point blocks, two passes:
//  * one thread per compact point: V_p over its observations in list order,
//    d_p, V_s^-1;
//  * one thread per observation in row order (coalesced J rows, W_s streamed
//    out): W_s = D_c (J_c^T J_p) D_p.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_pt_blocks(Structure s, Blocks b) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= s.npc) return;
  const int cp = s.cp, rl = cp + 3, N = s.N;
  const int t0 = s.pt_off[p], t1 = s.pt_off[p + 1];
  double *dp = b.d + N + 3 * p;
  double *Vi = b.Vi + (size_t)p * 9;
  if (t0 == t1) {
    b.pt_has[p] = 0;
    dp[0] = dp[1] = dp[2] = 0.0;
    for (int k = 0; k < 9; ++k) Vi[k] = 0.0;
    return;
  }
  double v00 = 0, v01 = 0, v02 = 0, v11 = 0, v12 = 0, v22 = 0;
  for (int t = t0; t < t1; ++t) {
    int o = s.pt_list[t];
    const double *q0 = s.vals + (size_t)o * 2 * rl + cp;
    const double *q1 = q0 + rl;
    double a0 = q0[0], a1 = q0[1], a2 = q0[2], c0 = q1[0], c1 = q1[1], c2 = q1[2];
    v00 = fma(a0, a0, v00); v00 = fma(c0, c0, v00);
    v01 = fma(a0, a1, v01); v01 = fma(c0, c1, v01);
    v02 = fma(a0, a2, v02); v02 = fma(c0, c2, v02);
    v11 = fma(a1, a1, v11); v11 = fma(c1, c1, v11);
    v12 = fma(a1, a2, v12); v12 = fma(c1, c2, v12);
    v22 = fma(a2, a2, v22); v22 = fma(c2, c2, v22);
  }
  double V[9] = {v00, v01, v02, v01, v11, v12, v02, v12, v22};
  double dd[3];
  bool bad = false;
  for (int k = 0; k < 3; ++k) {
    if (!(V[k * 4] > 0.0)) bad = true;
    dd[k] = 1.0 / sqrt(V[k * 4]);
    dp[k] = dd[k];
  }
  double Vs[9];
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l) Vs[k * 3 + l] = (V[k * 3 + l] * dd[k]) * dd[l];
  double iv[9];
  inv3x3_ref(Vs, iv);
  for (int k = 0; k < 9; ++k) {
    if (!isfinite(iv[k])) bad = true;
    Vi[k] = iv[k];
  }
  if (bad) atomicOr(s.err, ERRBIT_SINGULAR_V);
  b.pt_has[p] = 1;
}

// Per observation: W_s = D_c J_c^T J_p D_p (CP x 3).  (Y = W_s V_s^-1 is not
// stored: the Schur kernel forms it where it is used.)  A workgroup takes 64
// consecutive observations: their Jacobian rows (one contiguous run) are
// staged through LDS with every load issued up front, one thread per
// observation forms its block into LDS, and the contiguous output run is
// written back coalesced (observations dropped by fixPts keep their old
// contents).
constexpr int OB_OBS = 64, OB_THREADS = 128;

template <int CP>
__global__ __launch_bounds__(OB_THREADS) void k_obs_blocks(Structure s, Blocks b) {
  constexpr int RL = CP + 3, IN = OB_OBS * 2 * RL, OUT = OB_OBS * CP * 3;
  constexpr int NIN = (IN + OB_THREADS - 1) / OB_THREADS;
  __shared__ double jin[IN], wout[OUT];
  __shared__ int valid[OB_OBS];
  const int o0 = blockIdx.x * OB_OBS, tid = threadIdx.x;
  const int nob = min(OB_OBS, s.nobs - o0);
  {
    const double *src = s.vals + (size_t)o0 * 2 * RL;
    const int nin = nob * 2 * RL;
    double v[NIN];
#pragma unroll
    for (int r = 0; r < NIN; ++r) {
      const int x = tid + OB_THREADS * r;
      v[r] = x < nin ? src[x] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NIN; ++r) {
      const int x = tid + OB_THREADS * r;
      if (x < IN) jin[x] = v[r];
    }
  }
  __syncthreads();
  if (tid < nob) {
    const int o = o0 + tid;
    const int a = s.obs_cam[o], p = s.obs_pt[o];
    valid[tid] = a >= 0 && p >= 0;
    if (a >= 0 && p >= 0) {
      const double *r0 = jin + tid * 2 * RL;
      const double *r1 = r0 + RL;
      const double *dc = b.d + a * CP;
      const double *dpp = b.d + s.N + 3 * p;
      double dd[3], jp0[3], jp1[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        dd[k] = dpp[k];
        jp0[k] = r0[CP + k];
        jp1[k] = r1[CP + k];
      }
      double *Ws = wout + tid * CP * 3;
#pragma unroll
      for (int i = 0; i < CP; ++i) {
        const double ci0 = r0[i], ci1 = r1[i], dci = dc[i];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double wv = fma(ci1, jp1[k], ci0 * jp0[k]);
          Ws[i * 3 + k] = (wv * dci) * dd[k];
        }
      }
    }
  }
  __syncthreads();
  double *Wg = b.Ws + (size_t)o0 * CP * 3;
  const int nout = nob * CP * 3;
  for (int x = tid; x < nout; x += OB_THREADS)
    if (valid[x / (CP * 3)]) Wg[x] = wout[x];
}

template <int CP>
static void launch_obs_blocks_cp(const Structure &s, const Blocks &b, hipStream_t st) {
  hipLaunchKernelGGL(k_obs_blocks<CP>, dim3(cdiv(s.nobs, OB_OBS)), dim3(OB_THREADS), 0, st, s, b);
}

hipError_t launch_cam_blocks(const Structure &s, const Blocks &b, hipStream_t st) {
  hipLaunchKernelGGL(k_cam_assemble, dim3(s.nc), dim3(64), 0, st, s, b);
  return hipGetLastError();
}

hipError_t launch_point_blocks(const Structure &s, const Blocks &b, hipStream_t st) {
  if (s.npc > 0) hipLaunchKernelGGL(k_pt_blocks, dim3(cdiv(s.npc, 128)), dim3(128), 0, st, s, b);
  return hipGetLastError();
}

hipError_t launch_obs_blocks(const Structure &s, const Blocks &b, hipStream_t st) {
  if (s.nobs > 0) {
    switch (s.cp) {
      case 1: launch_obs_blocks_cp<1>(s, b, st); break;
      case 2: launch_obs_blocks_cp<2>(s, b, st); break;
      case 3: launch_obs_blocks_cp<3>(s, b, st); break;
      case 4: launch_obs_blocks_cp<4>(s, b, st); break;
      case 5: launch_obs_blocks_cp<5>(s, b, st); break;
      case 6: launch_obs_blocks_cp<6>(s, b, st); break;
      case 7: launch_obs_blocks_cp<7>(s, b, st); break;
      case 8: launch_obs_blocks_cp<8>(s, b, st); break;
      case 9: launch_obs_blocks_cp<9>(s, b, st); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

}  // namespace unc
