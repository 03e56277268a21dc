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
constexpr int SCHUR_BATCH = 64;      // observations of camera a staged per step
constexpr int SCHUR_MAXP = 2048;     // (o1, o2) pairs staged per step

struct Pair {
  int o2;
  short t;     // index of o1 inside the batch
  short slot;  // slot relative to the current chunk
};

__device__ inline int popc(unsigned int x) { return __popc(x); }

__global__ __launch_bounds__(SCHUR_THREADS) void k_schur(Structure s, Blocks bl, double *Z,
                                                         int slots_per_chunk) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nc = s.nc, cp = s.cp, N = s.N;
  const int tid = threadIdx.x;
  // XCD-aware camera mapping: workgroup i runs on XCD i % 8
  const int per_xcd = (nc + 7) / 8;
  const int a = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (a >= nc) return;

  const int ncw = (nc + 31) / 32;
  unsigned int *bits = (unsigned int *)smem;
  int *wpre = (int *)(bits + ncw);
  double *ybuf = (double *)(smem + ((8 * ncw + 15) / 16) * 16);
  Pair *pairs = (Pair *)(ybuf + SCHUR_BATCH * 27);
  int *cnt = (int *)(pairs + SCHUR_MAXP);           // SCHUR_BATCH + 1 ints
  double *acc = (double *)(((uintptr_t)(cnt + SCHUR_BATCH + 2) + 15) & ~(uintptr_t)15);

  const int first = s.cam_first[a], last = s.cam_last[a];
  for (int w = tid; w < ncw; w += SCHUR_THREADS) bits[w] = 0u;
  __syncthreads();
  for (int o1 = first + tid; o1 <= last; o1 += SCHUR_THREADS) {
    if (s.obs_cam[o1] < 0) continue;
    int p = s.obs_pt[o1];
    for (int t = s.pt_off[p]; t < s.pt_off[p + 1]; ++t) {
      int b = s.obs_cam[s.pt_list[t]];
      if (b >= a) atomicOr(&bits[b >> 5], 1u << (b & 31));
    }
  }
  __syncthreads();
  if (tid == 0) {
    int sum = 0;
    for (int w = 0; w < ncw; ++w) {
      wpre[w] = sum;
      sum += popc(bits[w]);
    }
    cnt[SCHUR_BATCH + 1] = sum;
  }
  __syncthreads();
  const int K = cnt[SCHUR_BATCH + 1];
  const int ne = cp * cp;
  const int G = SCHUR_THREADS / ne;  // element-owner groups
  const int g = tid / ne, e = tid % ne;
  const int ei = e / cp, ej = e % cp;
  const bool owner = tid < G * ne;

  for (int s0 = 0; s0 < K; s0 += slots_per_chunk) {
    const int S = min(slots_per_chunk, K - s0);
    for (int x = tid; x < S * ne; x += SCHUR_THREADS) acc[x] = 0.0;
    __syncthreads();
    for (int ob = first; ob <= last; ob += SCHUR_BATCH) {
      // stage: per-o1 pair counts (first wave), Y rows, then the ordered pair list
      int myc = 0, o1 = ob + tid, p = -1;
      bool valid = tid < SCHUR_BATCH && o1 <= last && s.obs_cam[o1] >= 0;
      if (valid) {
        p = s.obs_pt[o1];
        for (int t = s.pt_off[p]; t < s.pt_off[p + 1]; ++t) {
          int b = s.obs_cam[s.pt_list[t]];
          if (b < a) continue;
          int sl = wpre[b >> 5] + popc(bits[b >> 5] & ((1u << (b & 31)) - 1u));
          if (sl >= s0 && sl < s0 + S) ++myc;
        }
      }
      if (tid < SCHUR_BATCH) cnt[tid] = myc;
      __syncthreads();
      if (tid == 0) {
        int run = 0;
        for (int t = 0; t < SCHUR_BATCH; ++t) {
          int c = cnt[t];
          cnt[t] = run;
          run += c;
        }
        cnt[SCHUR_BATCH] = run;
      }
      __syncthreads();
      const int npairs = cnt[SCHUR_BATCH];
      if (npairs > SCHUR_MAXP) {  // cannot stage this batch; reported as unsupported
        if (tid == 0) atomicOr(s.err, ERRBIT_SCHUR_OVERFLOW);
        return;
      }
      if (valid) {
        const double *Y = bl.Y + (size_t)o1 * cp * 3;
        for (int k = 0; k < cp * 3; ++k) ybuf[tid * 27 + k] = Y[k];
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
      if (owner) {
        for (int q = 0; q < npairs; ++q) {
          Pair pr = pairs[q];
          if (pr.slot % G != g) continue;
          if (pr.slot + s0 == 0 && ej < ei) continue;  // diagonal block: upper triangle only
          const double *y = ybuf + pr.t * 27 + ei * 3;
          const double *w = bl.Ws + (size_t)pr.o2 * cp * 3 + ej * 3;
          double v = acc[pr.slot * ne + e];
          v = fma(y[0], w[0], v);
          v = fma(y[1], w[1], v);
          v = fma(y[2], w[2], v);
          acc[pr.slot * ne + e] = v;
        }
      }
      __syncthreads();
    }
    // write the chunk's blocks (and their mirrors)
    for (int b = a + tid; b < nc; b += SCHUR_THREADS) {
      if (!((bits[b >> 5] >> (b & 31)) & 1u)) continue;
      int sl = wpre[b >> 5] + popc(bits[b >> 5] & ((1u << (b & 31)) - 1u));
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
    __syncthreads();
  }
}

hipError_t launch_schur(const Structure &s, const Blocks &b, double *Z, hipStream_t st,
                        int max_lds_bytes) {
  const int ncw = (s.nc + 31) / 32;
  size_t fixed = ((8 * (size_t)ncw + 15) / 16) * 16 + SCHUR_BATCH * 27 * sizeof(double) +
                 SCHUR_MAXP * sizeof(Pair) + (SCHUR_BATCH + 2) * sizeof(int) + 32;
  const int ne = s.cp * s.cp;
  long long avail = (long long)max_lds_bytes - (long long)fixed;
  int slots = (int)(avail / (ne * (long long)sizeof(double)));
  if (slots < 1) return hipErrorInvalidValue;
  if (slots > 32767) slots = 32767;
  size_t bytes = fixed + (size_t)slots * ne * sizeof(double);
  hipError_t e = hipFuncSetAttribute((const void *)k_schur,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  unsigned int grid = 8u * (unsigned int)((s.nc + 7) / 8);
  hipLaunchKernelGGL(k_schur, dim3(grid), dim3(SCHUR_THREADS), bytes, st, s, b, Z, slots);
  return hipGetLastError();
}

}  // namespace unc
