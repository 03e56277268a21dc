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
constexpr int SCHUR_GROUPS = 21;     // off-diagonal groups of 9 threads (threads 64..252)

struct Pair {
  int o2;
  short t;     // index of o1 inside the batch
  short slot;  // slot relative to the current chunk
};

__device__ inline int popc(unsigned int x) { return __popc(x); }

// Work split inside the workgroup:
//  * wave 0 owns the upper triangle of the diagonal block (a, a): one element
//    per lane, accumulated in a register; its operands (Y_o1, W_o1) are staged.
//  * threads 64..252 form 21 groups of 9; group g owns the off-diagonal slots
//    s with (s - 1) % 21 == g, lane j of the group owns column j of those
//    blocks (9 accumulators in LDS) and reads W_o2 row j from global memory,
//    one pair ahead.
// Each element still sees its (o1, o2) pairs in list order, so the fma chain is
// the oracle's.
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
  double *wbuf = ybuf + SCHUR_BATCH * 27;
  Pair *pairs = (Pair *)(wbuf + SCHUR_BATCH * 27);
  int *cnt = (int *)(pairs + SCHUR_MAXP);  // SCHUR_BATCH + 2 ints
  int *o1buf = cnt + SCHUR_BATCH + 2;      // SCHUR_BATCH ints
  double *acc = (double *)(((uintptr_t)(o1buf + SCHUR_BATCH) + 15) & ~(uintptr_t)15);

  const int c0 = s.cam_off[a], c1 = s.cam_off[a + 1];
  for (int w = tid; w < ncw; w += SCHUR_THREADS) bits[w] = 0u;
  __syncthreads();
  for (int t1 = c0 + tid; t1 < c1; t1 += SCHUR_THREADS) {
    int p = s.obs_pt[s.cam_list[t1]];
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
  // diagonal-block role (wave 0): upper element (di, dj)
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
  const bool diag_owner = tid < nupper;
  // off-diagonal role
  const int og = (tid - 64) / 9, oj = (tid - 64) % 9;
  const bool off_owner = tid >= 64 && og < SCHUR_GROUPS && oj < cp;

  for (int s0 = 0; s0 < K; s0 += slots_per_chunk) {
    const int S = min(slots_per_chunk, K - s0);
    for (int x = tid; x < S * ne; x += SCHUR_THREADS) acc[x] = 0.0;
    double dacc = 0.0;
    __syncthreads();
    for (int tb = c0; tb < c1; tb += SCHUR_BATCH) {
      // stage: per-o1 pair counts, Y and W rows of o1, then the ordered pair list
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
      if (tid < SCHUR_BATCH) {
        cnt[tid] = myc;
        o1buf[tid] = o1;
      }
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
      if (s0 == 0 && diag_owner) {
        // diagonal block: slot 0 (camera a itself)
        const int nb = min(SCHUR_BATCH, c1 - tb);
        for (int t = 0; t < nb; ++t) {
          const double *y = ybuf + t * 27 + di * 3;
          for (int q = cnt[t]; q < cnt[t + 1]; ++q) {
            Pair pr = pairs[q];
            if (pr.slot != 0) continue;
            const double *w = (pr.o2 == o1buf[t]) ? wbuf + t * 27 + dj * 3
                                                  : bl.Ws + (size_t)pr.o2 * cp * 3 + dj * 3;
            dacc = fma(y[0], w[0], dacc);
            dacc = fma(y[1], w[1], dacc);
            dacc = fma(y[2], w[2], dacc);
          }
        }
      } else if (off_owner) {
        auto match = [&](int q) {
          while (q < npairs) {
            int sl = pairs[q].slot + s0;
            if (sl != 0 && (sl - 1) % SCHUR_GROUPS == og) break;
            ++q;
          }
          return q;
        };
        int q = match(0);
        double w0 = 0, w1 = 0, w2 = 0;
        if (q < npairs) {
          const double *w = bl.Ws + (size_t)pairs[q].o2 * cp * 3 + oj * 3;
          w0 = w[0]; w1 = w[1]; w2 = w[2];
        }
        while (q < npairs) {
          const int qn = match(q + 1);
          double n0 = 0, n1 = 0, n2 = 0;
          if (qn < npairs) {
            const double *w = bl.Ws + (size_t)pairs[qn].o2 * cp * 3 + oj * 3;
            n0 = w[0]; n1 = w[1]; n2 = w[2];
          }
          const Pair pr = pairs[q];
          const double *y = ybuf + pr.t * 27;
          double *ac = acc + pr.slot * ne + oj;
          for (int i = 0; i < cp; ++i) {
            double v = ac[i * cp];
            v = fma(y[i * 3 + 0], w0, v);
            v = fma(y[i * 3 + 1], w1, v);
            v = fma(y[i * 3 + 2], w2, v);
            ac[i * cp] = v;
          }
          w0 = n0; w1 = n1; w2 = n2;
          q = qn;
        }
      }
      __syncthreads();
    }
    if (s0 == 0 && diag_owner) acc[di * cp + dj] = dacc;
    __syncthreads();
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
  size_t fixed = ((8 * (size_t)ncw + 15) / 16) * 16 + 2 * SCHUR_BATCH * 27 * sizeof(double) +
                 SCHUR_MAXP * sizeof(Pair) + (2 * SCHUR_BATCH + 2) * sizeof(int) + 32;
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
