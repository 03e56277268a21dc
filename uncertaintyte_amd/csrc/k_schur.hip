// k_schur.hip -- (c) Schur complement Z_s = U_s - W_s V_s^-1 W_s^T for gfx950.
//
// Reference: composeZ (compute.cpp:107-112) forms Y = W iV, Y Wt by sparse
// products and U - YWt with addSSM (ScaledSparseMatrix.cpp:554-641), then
// densifies (ScaledDenseMatrix.cpp:54-66).
//
// Design: one workgroup owns camera a (its 9-row slab of Z).  The cameras b >= a
// that share a point with a get an LDS slot (81 FP64 accumulators) through a
// co-visibility bitmap + popcount prefix.  Observations o1 of camera a are swept
// in row order in batches: Y_o1 is staged in LDS and the (o1, o2) pairs of the
// batch are listed in order; every accumulator element is owned by one thread,
// which applies its pairs in list order with explicit fma().  Hence each Z entry
// is the same sequential fma chain as the CPU oracle (bit-identical) and the
// result is deterministic (no atomics on values).  Upper blocks are computed
// and mirrored, so Z is exactly symmetric.  Blocks that share no point stay at
// the zero written before the launch.  Workgroups are mapped to cameras so that
// consecutive cameras (which share points) land on the same XCD and its L2.
#include "common.h"

namespace unc {

constexpr int SCHUR_THREADS = 256;
constexpr int SCHUR_BATCH = 64;      // observations of camera a staged per step (wave 0)
constexpr int SCHUR_MAXP = 1024;     // (o1, o2) pairs staged per step
constexpr int SCHUR_GROUPS = 21;     // off-diagonal groups of 9 threads (threads 64..252)
constexpr int SCHUR_PREF = 12;       // W rows a group lane loads ahead per round

struct Pair {
  int o2;
  short t;     // index of o1 inside the batch
  short slot;  // slot relative to the current chunk
};

__device__ inline int popc(unsigned int x) { return __popc(x); }

// Co-visibility pre-pass: bit b of row a is set when cameras a <= b share a
// point; K_a = popcount(row a) is the number of LDS slots camera a needs.
__global__ __launch_bounds__(256) void k_covis(Structure s, unsigned int *covis, int *kcnt,
                                               int *kmax) {
  const int a = blockIdx.x, tid = threadIdx.x;
  const int ncw = (s.nc + 31) / 32;
  unsigned int *row = covis + (size_t)a * ncw;
  for (int w = tid; w < ncw; w += 256) row[w] = 0u;
  __syncthreads();
  const int c0 = s.cam_off[a], c1 = s.cam_off[a + 1];
  for (int t1 = c0 + tid; t1 < c1; t1 += 256) {
    const int p = s.obs_pt[s.cam_list[t1]];
    for (int t = s.pt_off[p]; t < s.pt_off[p + 1]; ++t) {
      const int b = s.obs_cam[s.pt_list[t]];
      if (b >= a) atomicOr(&row[b >> 5], 1u << (b & 31));
    }
  }
  __syncthreads();
  __shared__ int red[256];
  int c = 0;
  for (int w = tid; w < ncw; w += 256) c += popc(row[w]);
  red[tid] = c;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off) red[tid] += red[tid + off];
    __syncthreads();
  }
  if (tid == 0) {
    kcnt[a] = red[0];
    atomicMax(kmax, red[0]);
  }
}

hipError_t launch_covis(const Structure &s, unsigned int *covis, int *kcnt, int *kmax,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_covis, dim3(s.nc), dim3(256), 0, st, s, covis, kcnt, kmax);
  return hipGetLastError();
}

// Workgroup (camera a, chunk c) accumulates the blocks (a, b) whose slot lies in
// [c*S, c*S + S).  Work split inside the workgroup:
//  * wave 0 stages the batch and owns the upper triangle of the diagonal block
//    (chunk 0 only): one element per lane, accumulated in a register from
//    staged Y_o1 and W_o1.
//  * threads 64..252 form 21 groups of 9; group g owns the off-diagonal slots
//    s with (s - 1) % 21 == g; lane j owns column j of those blocks (9 LDS
//    accumulators per block) and loads the W_o2 rows of up to 12 of its pairs
//    at once before applying them in list order.
// Every element sees its (o1, o2) pairs in list order, so each Z entry is the
// oracle's fma chain.
__global__ __launch_bounds__(SCHUR_THREADS) void k_schur(Structure s, Blocks bl, double *Z,
                                                         const unsigned int *covis,
                                                         const int *kcnt, int slots_per_chunk) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nc = s.nc, cp = s.cp, N = s.N;
  const int tid = threadIdx.x, lane = tid & 63;
  // XCD-aware camera mapping: workgroup x runs on XCD x % 8
  const int per_xcd = (nc + 7) / 8;
  const int a = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (a >= nc) return;
  const int K = kcnt[a];
  const int s0 = blockIdx.y * slots_per_chunk;
  if (s0 >= K) return;
  const int S = min(slots_per_chunk, K - s0);

  const int ncw = (nc + 31) / 32;
  unsigned int *bits = (unsigned int *)smem;
  int *wpre = (int *)(bits + ncw);
  double *ybuf = (double *)(smem + ((8 * ncw + 15) / 16) * 16);
  double *wbuf = ybuf + SCHUR_BATCH * 27;
  Pair *pairs = (Pair *)(wbuf + SCHUR_BATCH * 27);
  int *cnt = (int *)(pairs + SCHUR_MAXP);  // SCHUR_BATCH + 2 ints
  int *o1buf = cnt + SCHUR_BATCH + 2;      // SCHUR_BATCH ints
  double *acc = (double *)(((uintptr_t)(o1buf + SCHUR_BATCH) + 15) & ~(uintptr_t)15);

  const unsigned int *row = covis + (size_t)a * ncw;
  for (int w = tid; w < ncw; w += SCHUR_THREADS) bits[w] = row[w];
  __syncthreads();
  if (tid < 64) {  // word prefix of popcounts, wave 0
    int carry = 0;
    for (int w0 = 0; w0 < ncw; w0 += 64) {
      int w = w0 + lane;
      int v = (w < ncw) ? popc(bits[w]) : 0;
      int incl = v;
      for (int off = 1; off < 64; off <<= 1) {
        int x = __shfl_up(incl, off);
        if (lane >= off) incl += x;
      }
      if (w < ncw) wpre[w] = carry + incl - v;
      carry += __shfl(incl, 63);
    }
  }
  for (int x = tid; x < S * cp * cp; x += SCHUR_THREADS) acc[x] = 0.0;
  __syncthreads();

  const int c0 = s.cam_off[a], c1 = s.cam_off[a + 1];
  const int ne = cp * cp;
  const int nupper = cp * (cp + 1) / 2;
  int di = 0, dj = 0;
  {
    int r = tid;
    for (int ii = 0; ii < cp; ++ii) {
      int len = cp - ii;
      if (r < len) { di = ii; dj = ii + r; break; }
      r -= len;
    }
  }
  const bool diag_owner = s0 == 0 && tid < nupper;
  const int og = (tid - 64) / 9, oj = (tid - 64) % 9;
  const bool off_owner = tid >= 64 && og < SCHUR_GROUPS && oj < cp;
  double dacc = 0.0;

  int tb = c0;
  while (tb < c1) {
    // ---- stage (wave 0): pair counts, wave prefix, Y/W rows, ordered pairs
    int myc = 0, o1 = -1, p = -1;
    const bool valid = tid < SCHUR_BATCH && tb + tid < c1;
    if (valid) {
      o1 = s.cam_list[tb + tid];
      p = s.obs_pt[o1];
      for (int t = s.pt_off[p]; t < s.pt_off[p + 1]; ++t) {
        int b = s.obs_cam[s.pt_list[t]];
        if (b < a) continue;
        int sl = wpre[b >> 5] + popc(bits[b >> 5] & ((1u << (b & 31)) - 1u));
        if (sl >= s0 && sl < s0 + S) ++myc;
      }
    }
    if (tid < 64) {
      int incl = myc;
      for (int off = 1; off < 64; off <<= 1) {
        int x = __shfl_up(incl, off);
        if (lane >= off) incl += x;
      }
      // truncate the batch so its pairs fit the staging buffer
      unsigned long long fits = __ballot(valid && incl <= SCHUR_MAXP);
      int nb = __popcll(fits);  // fits is a prefix of the valid lanes
      cnt[lane] = incl - myc;
      if (lane == 0) cnt[SCHUR_BATCH + 1] = nb;
      if (lane == nb - 1 || (nb == 0 && lane == 0)) cnt[SCHUR_BATCH] = nb ? incl : 0;
      o1buf[lane] = o1;
    }
    __syncthreads();
    const int nb = cnt[SCHUR_BATCH + 1];
    if (nb == 0) {  // one observation's track alone exceeds the staging buffer
      if (tid == 0) atomicOr(s.err, ERRBIT_SCHUR_OVERFLOW);
      return;
    }
    const int npairs = cnt[SCHUR_BATCH];
    if (valid && tid < nb) {
      const double *Y = bl.Y + (size_t)o1 * cp * 3;
      const double *W = bl.Ws + (size_t)o1 * cp * 3;
      for (int k = 0; k < cp * 3; ++k) {
        ybuf[tid * 27 + k] = Y[k];
        wbuf[tid * 27 + k] = W[k];
      }
      int w = cnt[tid];
      for (int t = s.pt_off[p]; t < s.pt_off[p + 1]; ++t) {
        int o2 = s.pt_list[t];
        int b = s.obs_cam[o2];
        if (b < a) continue;
        int sl = wpre[b >> 5] + popc(bits[b >> 5] & ((1u << (b & 31)) - 1u));
        if (sl >= s0 && sl < s0 + S) {
          Pair pr;
          pr.o2 = o2;
          pr.t = (short)tid;
          pr.slot = (short)(sl - s0);
          pairs[w++] = pr;
        }
      }
    }
    __syncthreads();
    // ---- accumulate
    if (diag_owner) {
      for (int t = 0; t < nb; ++t) {
        const int qe = (t + 1 < nb) ? cnt[t + 1] : npairs;
        const double *y = ybuf + t * 27 + di * 3;
        for (int q = cnt[t]; q < qe; ++q) {
          const Pair pr = pairs[q];
          if (pr.slot != 0) continue;
          const double *w = (pr.o2 == o1buf[t]) ? wbuf + t * 27 + dj * 3
                                                : bl.Ws + (size_t)pr.o2 * cp * 3 + dj * 3;
          dacc = fma(y[0], w[0], dacc);
          dacc = fma(y[1], w[1], dacc);
          dacc = fma(y[2], w[2], dacc);
        }
      }
    } else if (off_owner) {
      int q = 0;
      while (q < npairs) {
        int qs[SCHUR_PREF];
        double wv[SCHUR_PREF][3];
        int n = 0;
        for (; q < npairs && n < SCHUR_PREF; ++q) {
          const int sl = pairs[q].slot + s0;
          if (sl != 0 && (sl - 1) % SCHUR_GROUPS == og) qs[n++] = q;
        }
#pragma unroll
        for (int k = 0; k < SCHUR_PREF; ++k)
          if (k < n) {
            const double *w = bl.Ws + (size_t)pairs[qs[k]].o2 * cp * 3 + oj * 3;
            wv[k][0] = w[0];
            wv[k][1] = w[1];
            wv[k][2] = w[2];
          }
#pragma unroll
        for (int k = 0; k < SCHUR_PREF; ++k)
          if (k < n) {
            const Pair pr = pairs[qs[k]];
            const double *y = ybuf + pr.t * 27;
            double *ac = acc + pr.slot * ne + oj;
            for (int i = 0; i < cp; ++i) {
              double v = ac[i * cp];
              v = fma(y[i * 3 + 0], wv[k][0], v);
              v = fma(y[i * 3 + 1], wv[k][1], v);
              v = fma(y[i * 3 + 2], wv[k][2], v);
              ac[i * cp] = v;
            }
          }
      }
    }
    __syncthreads();
    tb += nb;
  }
  if (diag_owner) acc[di * cp + dj] = dacc;
  __syncthreads();
  // ---- write the chunk's blocks (and their mirrors)
  for (int b = a + tid; b < nc; b += SCHUR_THREADS) {
    if (!((bits[b >> 5] >> (b & 31)) & 1u)) continue;
    const int sl = wpre[b >> 5] + popc(bits[b >> 5] & ((1u << (b & 31)) - 1u));
    if (sl < s0 || sl >= s0 + S) continue;
    const double *A = acc + (sl - s0) * ne;
    if (b == a) {
      const double *U = bl.Us + (size_t)a * ne;
      for (int j = 0; j < cp; ++j)
        for (int i = 0; i < cp; ++i) {
          double v = (i <= j) ? U[i * cp + j] - A[i * cp + j] : U[j * cp + i] - A[j * cp + i];
          Z[(size_t)(a * cp + i) + (size_t)(a * cp + j) * N] = v;
        }
    } else {
      for (int j = 0; j < cp; ++j)
        for (int i = 0; i < cp; ++i) {
          double v = -A[i * cp + j];
          Z[(size_t)(a * cp + i) + (size_t)(b * cp + j) * N] = v;
          Z[(size_t)(b * cp + j) + (size_t)(a * cp + i) * N] = v;
        }
    }
  }
}

int schur_slots_per_chunk(int nc, int cp, int max_lds_bytes, size_t *bytes_out) {
  const int ncw = (nc + 31) / 32;
  size_t fixed = ((8 * (size_t)ncw + 15) / 16) * 16 + 2 * SCHUR_BATCH * 27 * sizeof(double) +
                 SCHUR_MAXP * sizeof(Pair) + (2 * SCHUR_BATCH + 2) * sizeof(int) + 32;
  const int ne = cp * cp;
  long long avail = (long long)max_lds_bytes - (long long)fixed;
  int slots = (int)(avail / (ne * (long long)sizeof(double)));
  if (slots < 1) return 0;
  if (slots > 32767) slots = 32767;
  if (bytes_out) *bytes_out = fixed + (size_t)slots * ne * sizeof(double);
  return slots;
}

hipError_t launch_schur(const Structure &s, const Blocks &b, double *Z, const unsigned int *covis,
                        const int *kcnt, int kmax, hipStream_t st, int max_lds_bytes) {
  size_t bytes = 0;
  const int slots = schur_slots_per_chunk(s.nc, s.cp, max_lds_bytes, &bytes);
  if (slots < 1) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute((const void *)k_schur,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  if (kmax <= 0) return hipSuccess;
  dim3 grid(8u * (unsigned int)((s.nc + 7) / 8), (unsigned int)((kmax + slots - 1) / slots));
  hipLaunchKernelGGL(k_schur, grid, dim3(SCHUR_THREADS), bytes, st, s, b, Z, covis, kcnt, slots);
  return hipGetLastError();
}

}  // namespace unc
