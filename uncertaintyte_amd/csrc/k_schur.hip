// k_schur.hip -- (c) Schur complement Z_s = U_s - W_s V_s^-1 W_s^T for gfx950.
//
// Reference: composeZ (compute.cpp:107-112) forms Y = W iV, Y Wt by sparse
// products and U - YWt with addSSM (ScaledSparseMatrix.cpp:554-641), then
// densifies (ScaledDenseMatrix.cpp:54-66).
//
// Design: a camera pair (a, b >= a) that shares points p_0 < p_1 < ... (point
// ids ascending) contributes the 9x9 block
//     S_ab = sum_k Y_{a,p_k} W_{b,p_k}^T,   Y_{a,p} = W_{a,p} V_p^-1.
// One workgroup per camera a (optionally split over `split` workgroups that
// take every split-th neighbour):
//   1. camera a's observations (sorted by point) go into an LDS hash table
//      point -> observation;
//   2. the cameras b >= a that share a point with a are marked in an LDS
//      bitmap (by walking the point lists of a's points);
//   3. each wave takes neighbours b in turn, scans b's point-sorted
//      observation list 64 entries at a time, looks each point up in the hash
//      table and appends the matches (o1 of a, o2 of b, point) to a wave-private
//      LDS buffer in point order;
//   4. the matches are consumed seven at a time: lane (g, r, c) of the wave
//      (g < 7 streams, r, c < 3) adds match 7s + g to its 3 x 3 sub-block
//      (rows 3r.., columns 3c..) of S_ab, with Y's rows formed on the fly from
//      W_s and V_s^-1 (the same expression as the per-observation Y);
//   5. at the end of the block the seven streams are added in order,
//      S = ((s_0 + s_1) + ...) + s_6, and U_s - S (diagonal block, upper part
//      mirrored) or -S is written to both (a, b) and (b, a).
// Every element therefore follows one fixed order -- match k into stream
// k mod 7 by a sequential fma chain over its three terms, then the fold --
// which the CPU oracle restates, so Z is bit-identical to it and independent
// of scheduling (no atomics on values).  Blocks without a shared point keep
// the zero written before the launch.
//
// Cameras with more observations than the LDS table holds (SC_CAP) look their
// points up in a global table instead (4 slots per observation of the
// workspace, at 4 * cam_off[a]), built by a small kernel before the launch.
// Each wave requests the next 64 entries of b's list before looking up the
// current ones.  A camera that observes a point twice is
// reported (ERRBIT_DUPLICATE): the order above assumes one observation per
// camera and point.
#include <algorithm>

#include "common.h"

namespace unc {

constexpr int SC_THREADS = 512;                 // 8 waves, two workgroups per CU
constexpr int SC_WAVES = SC_THREADS / 64;
constexpr int SC_CAP = 4096;                    // observations hashed in LDS
constexpr int SC_SLOTS = 2 * SC_CAP;            // open addressing, load <= 1/2
constexpr int SC_SLOT_BITS = 13;                // log2(SC_SLOTS)
constexpr int SC_WIN = 8192;                    // cameras per neighbour window
constexpr int SC_WORDS = SC_WIN / 32;
constexpr int SC_STREAMS = 7;                   // 9-lane groups per wave
constexpr int SC_BUF = 64 + SC_STREAMS;         // per-wave match buffer
__device__ __forceinline__ unsigned sc_hash(int p, int bits) {
  return ((unsigned)p * 2654435761u) >> (32 - bits);
}

__device__ __forceinline__ int sc_global_bits(int na) {
  int bits = 1;
  while ((1 << bits) < 2 * na) ++bits;
  return bits;
}

// visibility of LDS writes to the other lanes of the same wave
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int SC_PF_BITS = 16;                  // membership prefilter: 2^16 bits

__device__ __forceinline__ unsigned sc_pf(int p) {
  return ((unsigned)p * 0x9e3779b1u) >> (32 - SC_PF_BITS);
}

struct ScLds {
  int keys[SC_CAP];                 // point of a's idx-th observation (LDS table)
  int o1s[SC_CAP];                  // the observation itself
  unsigned slots[SC_SLOTS / 2];     // two 16-bit slots (idx) per word, 0xffff = empty
  unsigned pf[(1 << SC_PF_BITS) / 32];  // bit sc_pf(p) set for every point of a
  unsigned bitmap[SC_WORDS];        // neighbours b = w0 + bit
  int wpre[SC_WORDS + 1];           // exclusive prefix of the word popcounts
  int bo[SC_WAVES][3][SC_BUF];      // matches: o1, o2, point
};

// Global tables of the cameras with more than SC_CAP observations, built
// before the Schur launch so that several workgroups can share one: slots
// hold int positions into the camera's point-sorted list (-1 = empty), at
// 4 * cam_off[a] (2^bits <= 4 na slots).
__global__ __launch_bounds__(256) void k_schur_tables(Structure s, const int *cam_pts, int *ghash) {
  const int a = blockIdx.x;
  const int ca0 = s.cam_off[a], na = s.cam_off[a + 1] - ca0;
  if (na <= SC_CAP) return;
  const int bits = sc_global_bits(na);
  const unsigned mask = (1u << bits) - 1u;
  int *g = ghash + 4 * (size_t)ca0;
  const int *keys = cam_pts + ca0;
  for (int x = threadIdx.x; x < (1 << bits); x += blockDim.x) g[x] = -1;
  __threadfence();
  __syncthreads();
  for (int x = threadIdx.x; x < na; x += blockDim.x) {
    const int p = keys[x];
    unsigned q = sc_hash(p, bits);
    while (true) {
      const int held = atomicCAS(&g[q], -1, x);
      if (held == -1) break;
      if (keys[held] == p) {
        atomicOr(s.err, ERRBIT_DUPLICATE);
        break;
      }
      q = (q + 1) & mask;
    }
  }
}

// point -> observation of camera a, or -1: the prefilter first (one LDS read
// settles most misses), then the LDS table, or the global one of a large
// camera (read through L2 with atomic loads)
struct ScTable {
  ScLds *L;
  bool global;
  const int *g;            // global slots
  const int *gkeys, *go1;  // a's point-sorted points / observations
  int bits;

  __device__ int lookup(int p) const {
    const unsigned f = sc_pf(p);
    if (!((L->pf[f >> 5] >> (f & 31)) & 1u)) return -1;
    unsigned s = sc_hash(p, bits);
    const unsigned mask = (1u << bits) - 1u;
    while (true) {
      if (global) {
        const int v = __hip_atomic_load(&g[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v < 0) return -1;
        if (gkeys[v] == p) return go1[v];
      } else {
        const int v = (int)((L->slots[s >> 1] >> ((s & 1) * 16)) & 0xffffu);
        if (v == 0xffff) return -1;
        if (L->keys[v] == p) return L->o1s[v];
      }
      s = (s + 1) & mask;
    }
  }

  // LDS table: false when the point is already present (duplicate observation)
  __device__ bool insert(int idx, int p) {
    unsigned s = sc_hash(p, bits);
    const unsigned mask = (1u << bits) - 1u;
    while (true) {
      unsigned *w = &L->slots[s >> 1];
      const int sh = (s & 1) * 16;
      unsigned old = *w;
      int held = -1;
      while (true) {
        const unsigned cur = (old >> sh) & 0xffffu;
        if (cur != 0xffffu) {
          held = (int)cur;
          break;
        }
        const unsigned nw = (old & ~(0xffffu << sh)) | ((unsigned)idx << sh);
        const unsigned prev = atomicCAS(w, old, nw);
        if (prev == old) return true;
        old = prev;
      }
      if (L->keys[held] == p) return false;
      s = (s + 1) & mask;
    }
  }
};

__global__ __launch_bounds__(SC_THREADS, 4) void k_schur_rows(Structure s, Blocks bl, double *Z,
                                                              const int *cam_pts,
                                                              const int *pt_cams,
                                                              const int *ghash, int split) {
  __shared__ ScLds L;
  const int a = blockIdx.x / split, part = blockIdx.x % split;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cp = s.cp, nc = s.nc, N = s.N, r3 = cp * 3;
  const int ca0 = s.cam_off[a], na = s.cam_off[a + 1] - ca0;
  if (na == 0) return;

  ScTable T;
  T.L = &L;
  T.global = na > SC_CAP;
  T.gkeys = cam_pts + ca0;
  T.go1 = s.cam_list + ca0;
  for (int x = tid; x < (1 << SC_PF_BITS) / 32; x += SC_THREADS) L.pf[x] = 0u;
  if (T.global) {
    T.bits = sc_global_bits(na);
    T.g = ghash + 4 * (size_t)ca0;
  } else {
    T.bits = SC_SLOT_BITS;
    T.g = nullptr;
    for (int x = tid; x < SC_SLOTS / 2; x += SC_THREADS) L.slots[x] = 0xffffffffu;
    for (int x = tid; x < na; x += SC_THREADS) {
      L.keys[x] = cam_pts[ca0 + x];
      L.o1s[x] = s.cam_list[ca0 + x];
    }
  }
  __syncthreads();
  for (int x = tid; x < na; x += SC_THREADS) {
    const int p = cam_pts[ca0 + x];
    const unsigned f = sc_pf(p);
    atomicOr(&L.pf[f >> 5], 1u << (f & 31));
    if (!T.global && !T.insert(x, p) && part == 0) atomicOr(s.err, ERRBIT_DUPLICATE);
  }
  __syncthreads();

  // lane (g, r, c): stream g of the wave, sub-block rows 3r.., columns 3c..
  const int g = lane / 9, rc = lane % 9, r = rc / 3, c = rc % 3;
  const bool slane = g < SC_STREAMS;

  for (int w0 = a; w0 < nc; w0 += SC_WIN) {
    for (int x = tid; x < SC_WORDS; x += SC_THREADS) L.bitmap[x] = 0u;
    __syncthreads();
    for (int x = tid; x < na; x += SC_THREADS) {
      const int p = T.gkeys[x];
      const int u1 = s.pt_off[p + 1];
      for (int u = s.pt_off[p]; u < u1; ++u) {
        const int b = pt_cams[u] - w0;
        if (b >= 0 && b < SC_WIN) atomicOr(&L.bitmap[b >> 5], 1u << (b & 31));
      }
    }
    __syncthreads();
    if (wave == 0) {  // exclusive prefix of popcounts over the SC_WORDS words
      int carry = 0;
      for (int x0 = 0; x0 < SC_WORDS; x0 += 64) {
        const int v = __popc(L.bitmap[x0 + lane]);
        int incl = v;
        for (int off = 1; off < 64; off <<= 1) {
          const int y = __shfl_up(incl, off);
          if (lane >= off) incl += y;
        }
        L.wpre[x0 + lane] = carry + incl - v;
        carry += __shfl(incl, 63);
      }
      if (lane == 0) L.wpre[SC_WORDS] = carry;
    }
    __syncthreads();
    const int nnb = L.wpre[SC_WORDS];
    int(*bo)[SC_BUF] = L.bo[wave];
    int *bo1 = bo[0], *bo2 = bo[1], *bp = bo[2];
    for (int k = part + split * wave; k < nnb; k += split * SC_WAVES) {
      // neighbour k -> camera b: word by binary search, then the bit
      int lo = 0, hi = SC_WORDS - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (L.wpre[mid] <= k) lo = mid;
        else hi = mid - 1;
      }
      unsigned bits = L.bitmap[lo];
      for (int q = k - L.wpre[lo]; q > 0; --q) bits &= bits - 1u;
      const int b = w0 + (lo << 5) + __builtin_ctz(bits);
      const int cb0 = s.cam_off[b], cb1 = s.cam_off[b + 1];

      double acc[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[i][j] = 0.0;
      // the seven matches [e0, e0 + cnt) of the buffer, one per stream: lane
      // (g, r, c) loads match g's W_o1 rows 3r.., W_o2 rows 3c.. and V_p^-1
      auto consume_step = [&](int e0, int cnt) {
        if (!slane || g >= cnt) return;
        const int o1 = bo1[e0 + g], o2 = bo2[e0 + g], p = bp[e0 + g];
        const double *w1 = bl.Ws + (size_t)o1 * r3;
        const double *w2 = bl.Ws + (size_t)o2 * r3;
        const double *vi = bl.Vi + (size_t)p * 9;
        // Y's row i = W_o1 row i times V^-1, added into row ii of the sub-block
        double v[9], x2[3][3];
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
          const int j = 3 * c + jj;
#pragma unroll
          for (int m = 0; m < 3; ++m) x2[jj][m] = j < cp ? w2[j * 3 + m] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) v[q] = vi[q];
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) {
          const int i = 3 * r + ii;
          double wa0 = 0.0, wa1 = 0.0, wa2 = 0.0;
          if (i < cp) {
            wa0 = w1[i * 3 + 0];
            wa1 = w1[i * 3 + 1];
            wa2 = w1[i * 3 + 2];
          }
          double y[3];
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            double t = wa0 * v[0 * 3 + m];
            t = fma(wa1, v[1 * 3 + m], t);
            t = fma(wa2, v[2 * 3 + m], t);
            y[m] = t;
          }
#pragma unroll
          for (int jj = 0; jj < 3; ++jj) {
            double t = acc[ii][jj];
            t = fma(y[0], x2[jj][0], t);
            t = fma(y[1], x2[jj][1], t);
            t = fma(y[2], x2[jj][2], t);
            acc[ii][jj] = t;
          }
        }
      };
      auto consume = [&](int steps, int last) {
#pragma unroll 1
        for (int q = 0; q < steps; ++q)
          consume_step(q * SC_STREAMS, q + 1 == steps ? last : SC_STREAMS);
      };
      int nbuf = 0;
      int n2 = cb0 + lane < cb1 ? s.cam_list[cb0 + lane] : -1;
      int np2 = cb0 + lane < cb1 ? cam_pts[cb0 + lane] : -1;
      for (int t0 = cb0; t0 < cb1; t0 += 64) {
        const int o2 = n2, p = np2;
        // the next chunk is requested before this one's lookups
        const int tn = t0 + 64 + lane;
        n2 = tn < cb1 ? s.cam_list[tn] : -1;
        np2 = tn < cb1 ? cam_pts[tn] : -1;
        const int o1 = o2 >= 0 ? T.lookup(p) : -1;
        const unsigned long long m = __ballot(o1 >= 0);
        if (o1 >= 0) {
          const int pos = nbuf + (int)__popcll(m & ((1ull << lane) - 1ull));
          bo1[pos] = o1;
          bo2[pos] = o2;
          bp[pos] = p;
        }
        nbuf += (int)__popcll(m);
        wave_sync();
        const int full = nbuf / SC_STREAMS;
        if (full > 0) consume(full, SC_STREAMS);
        const int rem = nbuf - full * SC_STREAMS;
        if (full > 0 && rem > 0) {  // carry the remainder to the buffer front
          int x1 = 0, x2 = 0, x3 = 0;
          if (lane < rem) {
            x1 = bo1[full * SC_STREAMS + lane];
            x2 = bo2[full * SC_STREAMS + lane];
            x3 = bp[full * SC_STREAMS + lane];
          }
          wave_sync();
          if (lane < rem) {
            bo1[lane] = x1;
            bo2[lane] = x2;
            bp[lane] = x3;
          }
        }
        wave_sync();
        nbuf = rem;
      }
      if (nbuf > 0) consume(1, nbuf);
      // S = ((s_0 + s_1) + ...) + s_6 into the lanes of stream 0
      double S[3][3];
#pragma unroll
      for (int ii = 0; ii < 3; ++ii)
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) S[ii][jj] = acc[ii][jj];
#pragma unroll 1
      for (int gg = 1; gg < SC_STREAMS; ++gg)
#pragma unroll
        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
          for (int jj = 0; jj < 3; ++jj) {
            const double o = __shfl(acc[ii][jj], gg * 9 + rc);
            S[ii][jj] = S[ii][jj] + o;
          }
      if (g == 0) {
        const bool diag = b == a;
#pragma unroll
        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
          for (int jj = 0; jj < 3; ++jj) {
            const int i = 3 * r + ii, j = 3 * c + jj;
            if (i >= cp || j >= cp || (diag && j < i)) continue;
            const size_t R = (size_t)a * cp + i, C = (size_t)b * cp + j;
            const double v = diag ? bl.Us[((size_t)a * cp + i) * cp + j] - S[ii][jj] : -S[ii][jj];
            Z[R + C * N] = v;
            Z[C + R * N] = v;
          }
      }
      wave_sync();
    }
    __syncthreads();
  }
}

hipError_t launch_schur(const Structure &s, const Blocks &b, double *Z, const int *cam_pts,
                        const int *pt_cams, int *ghash, hipStream_t st) {
  if (s.cp > 9) return hipErrorInvalidValue;
  if (s.nc <= 0) return hipSuccess;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    cus = 256;
  }
  hipLaunchKernelGGL(k_schur_tables, dim3(s.nc), dim3(256), 0, st, s, cam_pts, ghash);
  // few cameras: several workgroups per camera, each taking every split-th
  // neighbour, so that every CU holds two
  const int split = std::max(1, std::min(8, (2 * cus + s.nc - 1) / s.nc));
  hipLaunchKernelGGL(k_schur_rows, dim3(s.nc * split), dim3(SC_THREADS), 0, st, s, b, Z, cam_pts,
                     pt_cams, ghash, split);
  return hipGetLastError();
}

}  // namespace unc
