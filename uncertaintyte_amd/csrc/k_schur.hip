// k_schur.hip -- (c) Schur complement Z_s = U_s - W_s V_s^-1 W_s^T for gfx950.
//
// Reference: composeZ (compute.cpp:107-112) forms Y = W iV, Y Wt by sparse
// products and U - YWt with addSSM (ScaledSparseMatrix.cpp:554-641), then
// densifies (ScaledDenseMatrix.cpp:54-66).
//
// Design: every camera pair (a, b >= a) that shares a point contributes one
// 9x9 block S_ab = sum over pairs (o1 of a, o2 of b, same point) of
// Y_o1 W_o2^T.  The pairs are listed in the order (a, o1 in row order, o2 in
// point order) -- position = exclusive scan of per-observation pair counts --
// grouped by block with a stable radix sort on the key (a, b) carrying the
// packed (o1, o2), and cut into runs of equal key.  One workgroup per run stages
// the pairs' Y and W rows in LDS and
// thread (i, j) forms element (i, j) as one sequential fma chain in list order.
// That is the CPU oracle's chain, so Z is bit-identical to it and does not
// depend on scheduling (no atomics on values).  Upper blocks are written with
// their mirrors, so Z is exactly symmetric; blocks without a shared point keep
// the zero written before the launch.
#include "common.h"

namespace unc {

// pairs contributed by the observation at camera-sorted position t
__global__ void k_pair_count(Structure s, int *cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= s.nobs) return;
  const int nk = s.cam_off[s.nc];
  int c = 0;
  if (t < nk) {
    const int o1 = s.cam_list[t];
    const int a = s.obs_cam[o1], p = s.obs_pt[o1];
    for (int u = s.pt_off[p]; u < s.pt_off[p + 1]; ++u) c += (s.obs_cam[s.pt_list[u]] >= a) ? 1 : 0;
  }
  cnt[t] = c;
}

// cps[a] = first pair of camera a (a <= nc); cps[nc] = total
__global__ void k_cam_pair_start(Structure s, const int *pofs, const int *cnt, int *cps) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a > s.nc) return;
  const int total = s.nobs > 0 ? pofs[s.nobs - 1] + cnt[s.nobs - 1] : 0;
  const int t = s.cam_off[a];
  cps[a] = (t < s.nobs) ? pofs[t] : total;
}

// pairs of positions [t0, t1) (cameras [a0, a1)), written at pofs[t] - base:
// key = (a - a0) nc + b, value = o1 << 32 | o2
__global__ void k_pair_emit(Structure s, int t0, int t1, int a0, int base, const int *pofs,
                            unsigned int *keys, unsigned long long *vals) {
  const int t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= t1) return;
  const int o1 = s.cam_list[t];
  const int a = s.obs_cam[o1], p = s.obs_pt[o1];
  int w = pofs[t] - base;
  const unsigned int row = (unsigned int)(a - a0) * (unsigned int)s.nc;
  for (int u = s.pt_off[p]; u < s.pt_off[p + 1]; ++u) {
    const int o2 = s.pt_list[u];
    const int b = s.obs_cam[o2];
    if (b < a) continue;
    keys[w] = row + (unsigned int)b;
    vals[w] = ((unsigned long long)(unsigned int)o1 << 32) | (unsigned int)o2;
    ++w;
  }
}

constexpr int SR_THREADS = 128;
constexpr int SR_ROW = 27;  // cp * 3 for cp <= 9

// one run (block (a, b)) per workgroup iteration; thread e owns element (i, j)
// (upper triangle only on diagonal blocks).  Per chunk of SR_CHUNK pairs: the pairs'
// (o1, o2) are staged first, then every thread issues its share of the 2 x SR_CHUNK
// Y / W rows as independent loads.
template <int SR_CHUNK>
__global__ __launch_bounds__(SR_THREADS) void k_schur_reduce(Structure s, Blocks bl, double *Z,
                                                             int a0, const unsigned int *ukeys,
                                                             const int *counts, const int *offs,
                                                             const int *nruns,
                                                             const unsigned long long *pairs) {
  __shared__ double ys[SR_CHUNK][SR_ROW], ws[SR_CHUNK][SR_ROW];
  __shared__ int po1[SR_CHUNK], po2[SR_CHUNK];
  const int tid = threadIdx.x, cp = s.cp, nc = s.nc, N = s.N, r3 = cp * 3;
  const int nr = *nruns;
  for (int seg = blockIdx.x; seg < nr; seg += gridDim.x) {
    const unsigned int key = ukeys[seg];
    const int a = a0 + (int)(key / (unsigned int)nc), b = (int)(key % (unsigned int)nc);
    const int off = offs[seg], cnt = counts[seg];
    const bool diag = a == b;
    int i = 0, j = 0;
    bool act;
    if (diag) {
      int r = tid;
      act = tid < cp * (cp + 1) / 2;
      if (act) {
        while (r >= cp - i) {
          r -= cp - i;
          ++i;
        }
        j = i + r;
      }
    } else {
      act = tid < cp * cp;
      i = tid / cp;
      j = tid % cp;
    }
    double acc = 0.0;
    for (int c0 = 0; c0 < cnt; c0 += SR_CHUNK) {
      const int m = min(SR_CHUNK, cnt - c0);
      __syncthreads();
      if (tid < m) {
        const unsigned long long v = pairs[off + c0 + tid];
        po1[tid] = (int)(v >> 32);
        po2[tid] = (int)(v & 0xffffffffull);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < (SR_CHUNK * SR_ROW + SR_THREADS - 1) / SR_THREADS; ++k) {
        const int x = tid + k * SR_THREADS;
        const int q = x / SR_ROW, r = x % SR_ROW;
        if (q < m && r < r3) {
          ys[q][r] = bl.Y[(size_t)po1[q] * r3 + r];
          ws[q][r] = bl.Ws[(size_t)po2[q] * r3 + r];
        }
      }
      __syncthreads();
      if (act) {
        for (int q = 0; q < m; ++q) {
          acc = fma(ys[q][i * 3 + 0], ws[q][j * 3 + 0], acc);
          acc = fma(ys[q][i * 3 + 1], ws[q][j * 3 + 1], acc);
          acc = fma(ys[q][i * 3 + 2], ws[q][j * 3 + 2], acc);
        }
      }
    }
    if (act) {
      const size_t r = (size_t)a * cp + i, c = (size_t)b * cp + j;
      const double v = diag ? bl.Us[((size_t)a * cp + i) * cp + j] - acc : -acc;
      Z[r + c * N] = v;
      Z[c + r * N] = v;
    }
  }
}

hipError_t launch_pair_count(const Structure &s, int *cnt, hipStream_t st) {
  if (s.nobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_count, dim3(cdiv(s.nobs, 256)), dim3(256), 0, st, s, cnt);
  return hipGetLastError();
}

hipError_t launch_cam_pair_start(const Structure &s, const int *pofs, const int *cnt, int *cps,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_cam_pair_start, dim3(cdiv(s.nc + 1, 256)), dim3(256), 0, st, s, pofs, cnt,
                     cps);
  return hipGetLastError();
}

hipError_t launch_pair_emit(const Structure &s, int t0, int t1, int a0, int base, const int *pofs,
                            unsigned int *keys, unsigned long long *vals, hipStream_t st) {
  if (t1 <= t0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_emit, dim3(cdiv(t1 - t0, 256)), dim3(256), 0, st, s, t0, t1, a0, base,
                     pofs, keys, vals);
  return hipGetLastError();
}

static int g_schur_variant = 0;
void set_schur_variant(int v) { g_schur_variant = v; }

// npairs: pairs in this launch.  Large launches (long runs) take 64-pair chunks
// (twice the loads in flight per round trip, 3 waves per SIMD); small ones 32-pair
// chunks (5 waves per SIMD).  Measured: rings200 (1.22 M pairs) 0.74 -> 0.68 ms with
// 32, walk1500 (16.4 M pairs) 11.3 -> 16.7 ms.
hipError_t launch_schur_reduce(const Structure &s, const Blocks &b, double *Z, int a0,
                               const unsigned int *ukeys, const int *counts, const int *offs,
                               const int *nruns, const unsigned long long *pairs, long long npairs,
                               hipStream_t st) {
  if (s.cp > 9) return hipErrorInvalidValue;
  const bool wide = g_schur_variant == 1 || (g_schur_variant == 0 && npairs >= (1ll << 22));
  if (wide)
    hipLaunchKernelGGL(k_schur_reduce<64>, dim3(2048), dim3(SR_THREADS), 0, st, s, b, Z, a0, ukeys,
                       counts, offs, nruns, pairs);
  else
    hipLaunchKernelGGL(k_schur_reduce<32>, dim3(2560), dim3(SR_THREADS), 0, st, s, b, Z, a0, ukeys,
                       counts, offs, nruns, pairs);
  return hipGetLastError();
}

}  // namespace unc
