// k_crt_gemm.hip -- C = alpha op(A) op(B) + beta C in FP64 from one exact integer
// product, formed on the int8 matrix cores (v_mfma_i32_16x16x64_i8) modulo a set
// of coprime moduli and rebuilt by the Chinese remainder theorem.
//
// Method (Ozaki, Uchino, Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of
// floating-point matrix multiplication using an integer modular technique",
// 2025):
//  1. every row i of op(A) is scaled by 2^a_i and every column j of op(B) by
//     2^b_j so that their 2-norms stay below 2^(q-1), and rounded to integers
//     A', B' (|A'_ik|, |B'_kj| < 2^(q-1));
//  2. by Cauchy-Schwarz |(A'B')_ij| <= |A'_i.| |B'_.j| < 2^(2q-2) < P/2 with
//     P = m_1 ... m_s, so A'B' is fixed by its residues modulo the m_l;
//  3. for each m_l the residues of A' and B' (symmetric, in [-128, 127]) are
//     multiplied exactly in int32 (K 128^2 < 2^31 for K < 131072) and reduced
//     modulo m_l;
//  4. (A'B')_ij = sum_l r_l E_l mod P in 128-bit integers (E_l = P_l (P_l^-1 mod
//     m_l), P_l = P / m_l), taken in the symmetric range, converted to FP64 and
//     scaled back by 2^-(a_i + b_j).
// The only rounding is the cut of step 1: |C - A B|_ij <= 2^(-q+2) sqrt(K)
// |A_i.| |B_.j| (plus the FP64 rounding of the result).  With s moduli,
// q = floor((log2 P - 1) / 2): 10 moduli give q = 39, 13 give q = 50.
//
// Operands are cut into residue planes laid out in the order the MFMA reads
// them: a plane holds blocks of 16 rows x 64 k, each a contiguous 1 KB in which
// lane l's 16 bytes (row l % 16, k 16 (l / 16) .. + 15) sit at 16 l, so every
// global load, LDS store and LDS read of a wave is one contiguous 1 KB.
#include <algorithm>
#include <cmath>

#include "common.h"
#include "gemm_tile.h"

namespace unc {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int CRT_T = 256;      // rows of an output tile; operand rows are padded to it
constexpr int CRT_KB = 64;      // k per block and per pipeline step
constexpr int CRT_BLK = 1024;   // bytes of one 16 x 64 residue block
constexpr int CRT_NCH = 16;     // K chunks of the row-norm partial sums (gathered operands)
constexpr int CRT_THREADS = 256;
constexpr int CRT_NARROW_Q = 51;  // up to q bits the cut integers stay below 2^52

// pairwise coprime moduli <= 256; their residues fit int8 in the symmetric range
constexpr int kModuli[CRT_MAX_MOD] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                      223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

constexpr int CRT_LIMBS = 10;  // 16-bit limbs of E_l < P < 2^160 (20 moduli: 2^155.4)
constexpr int CRT_WORDS = 6;   // 32-bit words of P and of the rebuilt sums

struct CrtConst {
  int nmod, q, nlimb;
  int m[CRT_MAX_MOD];
  int c16[CRT_MAX_MOD];                      // 2^16 mod m
  double minv[CRT_MAX_MOD];
  unsigned int e16[CRT_MAX_MOD][CRT_LIMBS];  // E_l in 16-bit limbs, least significant first
  unsigned int p[CRT_WORDS], ph[CRT_WORDS];  // P and P / 2 in 32-bit words
  double pinv;                               // 1 / P (rounded)
};

// Non-negative integers below 2^256 as eight 32-bit words (host only)
struct Big {
  unsigned int w[8] = {};
  void mul(unsigned int f) {
    unsigned long long c = 0;
    for (unsigned int &x : w) {
      c += (unsigned long long)x * f;
      x = (unsigned int)c;
      c >>= 32;
    }
  }
  unsigned int divmod(unsigned int d) {  // *this /= d, returns the remainder
    unsigned long long r = 0;
    for (int k = 7; k >= 0; --k) {
      const unsigned long long cur = (r << 32) | w[k];
      w[k] = (unsigned int)(cur / d);
      r = cur % d;
    }
    return (unsigned int)r;
  }
  int bits() const {
    for (int k = 7; k >= 0; --k)
      if (w[k]) return 32 * k + 32 - __builtin_clz(w[k]);
    return 0;
  }
  double to_double() const {
    double v = 0.0;
    for (int k = 7; k >= 0; --k) v = v * 4294967296.0 + (double)w[k];
    return v;
  }
};

inline unsigned int inv_mod(unsigned int a, unsigned int m) {
  for (unsigned int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 0;
}

CrtConst make_const(int nmod) {
  CrtConst c{};
  c.nmod = nmod;
  Big P;
  P.w[0] = 1;
  for (int l = 0; l < nmod; ++l) P.mul((unsigned int)kModuli[l]);
  for (int l = 0; l < nmod; ++l) {
    const unsigned int m = (unsigned int)kModuli[l];
    Big E = P;
    (void)E.divmod(m);          // P_l = P / m (exact)
    Big t = E;
    const unsigned int y = inv_mod(t.divmod(m), m);
    E.mul(y);                   // E_l = P_l (P_l^-1 mod m) < P
    c.m[l] = (int)m;
    c.c16[l] = (int)(65536u % m);
    c.minv[l] = 1.0 / (double)m;
    for (int k = 0; k < CRT_LIMBS; ++k) c.e16[l][k] = (E.w[k / 2] >> (16 * (k & 1))) & 0xffffu;
  }
  for (int k = 0; k < CRT_WORDS; ++k) {
    c.p[k] = P.w[k];
    c.ph[k] = (P.w[k] >> 1) | (k + 1 < 8 ? P.w[k + 1] << 31 : 0u);
  }
  c.nlimb = (P.bits() + 15) / 16;
  const double pd = P.to_double();
  c.pinv = 1.0 / pd;
  c.q = (int)std::floor((std::log2(pd) - 1.0) / 2.0);
  return c;
}

__device__ __forceinline__ bool crt_keep(int mask, int r, int k) {
  return mask == 0 || (mask == 1 ? k <= r : k >= r);
}

// part[ch * R + r] = sum over the ch-th of CRT_NCH chunks of k of P(r, k)^2 for an
// operand whose element (r, k) sits at P[r + k ld] (rows gathered across k);
// 64 rows x 4 interleaved k phases per workgroup, phases added in order.  Each
// thread keeps four loads in flight (four partial sums); masked elements are
// loaded and dropped by a select, never multiplied.
__global__ __launch_bounds__(256) void k_crt_norm_gather(int R, int K, const double *__restrict__ P,
                                                          int ld, int mask, double *part) {
  __shared__ double s4[4][64];
  const int rr = threadIdx.x & 63, ph = threadIdx.x >> 6, r = blockIdx.x * 64 + rr;
  const int per = (K + CRT_NCH - 1) / CRT_NCH;
  const int kb = blockIdx.y * per, ke = min(K, kb + per);
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  if (r < R) {
    int k = kb + ph;
    for (; k + 12 < ke; k += 16) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = P[(size_t)r + (size_t)(k + 4 * u) * ld];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double x = crt_keep(mask, r, k + 4 * u) ? v[u] : 0.0;
        s[u] = fma(x, x, s[u]);
      }
    }
    for (; k < ke; k += 4) {
      const double v = P[(size_t)r + (size_t)k * ld];
      const double x = crt_keep(mask, r, k) ? v : 0.0;
      s[0] = fma(x, x, s[0]);
    }
  }
  s4[ph][rr] = (s[0] + s[1]) + (s[2] + s[3]);
  __syncthreads();
  if (ph == 0 && r < R)
    part[(size_t)blockIdx.y * R + r] = ((s4[0][rr] + s4[1][rr]) + s4[2][rr]) + s4[3][rr];
}

// part[r] = sum_k P(r, k)^2 for an operand whose element (r, k) sits at
// P[k + r ld] (rows contiguous in k): one wave per row, four loads in flight
// per lane (four partial sums), then a fixed butterfly.
__global__ __launch_bounds__(256) void k_crt_norm_contig(int R, int K, const double *__restrict__ P,
                                                          int ld, int mask, double *part) {
  const int lane = threadIdx.x & 63, r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= R) return;
  const double *row = P + (size_t)r * ld;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  int k = lane;
  for (; k + 192 < K; k += 256) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = row[k + 64 * u];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double x = crt_keep(mask, r, k + 64 * u) ? v[u] : 0.0;
      s[u] = fma(x, x, s[u]);
    }
  }
  for (; k < K; k += 64) {
    const double v = row[k];
    const double x = crt_keep(mask, r, k) ? v : 0.0;
    s[0] = fma(x, x, s[0]);
  }
  double t = (s[0] + s[1]) + (s[2] + s[3]);
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
  if (lane == 0) part[r] = t;
}

// Symmetric residue of the integer x (|x| < 2^52, held exactly as a double;
// WIDE: any |x| < 2^100, in two steps: x - m rint(x / m) is exact and below
// 2^25 even where x / m rounds, and the second step leaves |r| <= m / 2).
template <bool WIDE>
__device__ __forceinline__ int crt_residue(double x, int m, double minv) {
  if (WIDE) {
    const double r1 = fma(-(double)m, rint(x * minv), x);
    int r = (int)fma(-(double)m, rint(r1 * minv), r1);
    const int hi = (m - 1) >> 1, lo = hi - m + 1;
    if (r > hi) r -= m;
    else if (r < lo) r += m;
    return r;
  }
  if (m == 256) {
    const int r = (int)((long long)x & 255);
    return r >= 128 ? r - 256 : r;
  }
  const double t = rint(x * minv);
  int r = (int)fma(-(double)m, t, x);  // exact: m t < 2^50 and the result is a small integer
  const int h = m >> 1;
  if (r > h) r -= m;
  else if (r < -h) r += m;
  return r;
}

// The scale exponent of row r from its norm partials (nch of them, summed in
// order): |row| 2^e < 2^(q-2).
__device__ __forceinline__ int crt_row_exponent(int r, int R, const double *__restrict__ part,
                                                int nch, int q) {
  if (r >= R) return 0;
  double ss = 0.0;
  for (int c = 0; c < nch; ++c) ss += part[(size_t)c * R + r];
  if (!(ss > 0.0)) return 0;
  int E;
  (void)frexp(ss, &E);              // ss < 2^E, so |row| < 2^ceil(E/2)
  return q - 2 - ((E + 1) >> 1);
}

// The residues of the 16 integers x (a lane's row and 16 k of one 16 x 64 block)
// for every modulus, 16 bytes per plane at the lane's slot of the block.
template <bool WIDE>
__device__ __forceinline__ void crt_emit(const double (&x)[16], const CrtConst &cc,
                                         signed char *__restrict__ planes, size_t plane_bytes,
                                         size_t at) {
  for (int l = 0; l < cc.nmod; ++l) {
    const int m = cc.m[l];
    const double mi = cc.minv[l];
    int w[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      unsigned int u = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        u |= (unsigned int)(unsigned char)(signed char)crt_residue<WIDE>(x[4 * g + b], m, mi) << (8 * b);
      w[g] = (int)u;
    }
    *(v4i *)(planes + (size_t)l * plane_bytes + at) = (v4i){w[0], w[1], w[2], w[3]};
  }
}

// Residue planes of an operand with R rows over K whose element (r, k) sits at
// P[r + k ld] (rows gathered across k; `mask` keeps k <= r (1) or k >= r (2)):
// one wave per 16 x 64 block, lane l taking row l % 16 and k 16 (l / 16) .. + 15,
// so each load reads 16 consecutive rows.  expo[r] = the row's scale exponent.
template <bool WIDE>
__global__ __launch_bounds__(256) void k_crt_split_gather(int R, int K, int Rp, int Kp,
                                                          const double *__restrict__ P, int ld,
                                                          int mask,
                                                          const double *__restrict__ part, int nch,
                                                          CrtConst cc,
                                                          signed char *__restrict__ planes,
                                                          size_t plane_bytes,
                                                          int *__restrict__ expo) {
  const int lane = threadIdx.x & 63;
  const int kblocks = Kp / CRT_KB;
  const long long bi = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (bi >= (long long)(Rp / 16) * kblocks) return;
  const int rb = (int)(bi / kblocks), kb = (int)(bi - (long long)rb * kblocks);
  const int r = rb * 16 + (lane & 15), k0 = kb * CRT_KB + 16 * (lane >> 4);
  const int e = crt_row_exponent(r, R, part, nch, cc.q);
  if (kb == 0 && lane < 16) expo[r] = e;
  // every load issued before the first is used: clamped into the operand,
  // the elements outside it or masked dropped by a select
  const int rc = min(r, R - 1);
  double x[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) x[t] = P[(size_t)rc + (size_t)min(k0 + t, K - 1) * ld];
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int k = k0 + t;
    const double v = (r < R && k < K && crt_keep(mask, r, k)) ? x[t] : 0.0;
    x[t] = rint(ldexp(v, e));
  }
  crt_emit<WIDE>(x, cc, planes, plane_bytes,
                 ((size_t)rb * kblocks + kb) * CRT_BLK + 16 * (size_t)lane);
}

// The same for an operand whose rows are contiguous in k (element (r, k) at
// P[k + r ld]): a workgroup stages 16 rows x 256 k through LDS, each load
// instruction reading 128 consecutive bytes of four rows, then each wave cuts one
// 16 x 64 block from LDS.  One norm partial per row.
template <bool WIDE>
__global__ __launch_bounds__(256) void k_crt_split_contig(int R, int K, int Rp, int Kp,
                                                          const double *__restrict__ P, int ld,
                                                          int mask,
                                                          const double *__restrict__ part,
                                                          CrtConst cc,
                                                          signed char *__restrict__ planes,
                                                          size_t plane_bytes,
                                                          int *__restrict__ expo) {
  __shared__ double T[16][4 * CRT_KB + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kblocks = Kp / CRT_KB, kgroups = (kblocks + 3) / 4;
  const int rb = blockIdx.x / kgroups, kg = blockIdx.x - rb * kgroups;
  {
    const int lr = tid >> 4, c = tid & 15, r = rb * 16 + lr, rc = min(r, R - 1);
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i)
      v[i] = P[(size_t)min(kg * 4 * CRT_KB + c + 16 * i, K - 1) + (size_t)rc * ld];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int kl = c + 16 * i, k = kg * 4 * CRT_KB + kl;
      T[lr][kl] = (r < R && k < K && crt_keep(mask, r, k)) ? v[i] : 0.0;
    }
  }
  __syncthreads();
  const int kb = kg * 4 + wave;
  if (kb >= kblocks) return;
  const int r = rb * 16 + (lane & 15), kl0 = wave * CRT_KB + 16 * (lane >> 4);
  const int e = crt_row_exponent(r, R, part, 1, cc.q);
  if (kb == 0 && lane < 16) expo[r] = e;
  double x[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) x[t] = rint(ldexp(T[lane & 15][kl0 + t], e));
  crt_emit<WIDE>(x, cc, planes, plane_bytes,
                 ((size_t)rb * kblocks + kb) * CRT_BLK + 16 * (size_t)lane);
}

// Residue (in [0, m)) of an int32 product sum.
__device__ __forceinline__ unsigned int crt_reduce(int c, int m, double minv) {
  if (m == 256) return (unsigned int)c & 255u;
  const int t = (int)floor((double)c * minv);
  int r = c - m * t;
  if (r < 0) r += m;
  else if (r >= m) r -= m;
  return (unsigned int)r;
}

// One 256 x 128 tile of one modulus' product per workgroup, two workgroups per
// CU (one covers the other's barrier and load waits): 4 waves (one per SIMD),
// wave (wm, wn) = (w & 1, w >> 1) owning rows 128 wm .. + 127 and columns
// 64 wn .. + 63 as 8 x 4 blocks of 16 x 16 (the int32 sums in the accumulation
// registers).  A k step (64) is 24 residue blocks of 1 KB (16 of A, 8 of B),
// copied straight from global memory into one of CRT_NBUF LDS buffers (each
// wave moves 6 blocks, one global_load_lds_dwordx4 each).  Each step reads its
// 12 fragments per wave from LDS once, while the 32 MFMAs run.  The residues of
// the int32 sums go to the C plane of the modulus in 16 x 16 blocks.
// Grid: nmod x the tiles (the lower ones only when GEMM_C_LOWER), modulus-major,
// each XCD taking a contiguous range so that the workgroups sharing a modulus'
// panels share its L2.
// Measured on the 6784-row products of a 13500-row recursion (12 moduli): 256 x
// 256 tiles, one workgroup per CU, 3.9 ms against 3.5 for these; the same tiles
// on v_mfma_i32_32x32x32_i8 (32 x 32 blocks, half the instructions) 3.7 ms: the
// kernel waits on its operand loads (L2 hit rate ~65 %), not on MFMA issue.
typedef __attribute__((address_space(3))) void lds_void;

// s_waitcnt immediate for vmcnt(n), everything else unconstrained (gfx9:
// vmcnt in bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8)
constexpr int vmcnt_imm(int n) { return 0x0F70 | (n & 15) | (((n >> 4) & 3) << 14); }

constexpr int CRT_BN = 128;                // columns of the workgroup tile
constexpr int CRT_NBUF = 3;                // LDS buffers (k steps in flight + 1)
constexpr int CRT_JN = CRT_BN / 32;        // 16-column blocks per wave
constexpr int CRT_NBL = 16 + CRT_BN / 16;  // residue blocks per k step (A then B)
constexpr int CRT_PER = CRT_NBL / 4;       // of them copied by each wave
constexpr int CRT_TPR = CRT_T / CRT_BN;    // tile columns per 256 columns

__global__ __launch_bounds__(CRT_THREADS, 2) void k_crt_gemm(
    int Mp, int Np, int Kp, int tiles_m, int tiles_n, int ntiles, int tri, CrtConst cc,
    const signed char *__restrict__ PA, const signed char *__restrict__ PB, size_t plane_a,
    size_t plane_b, unsigned char *__restrict__ PC, size_t plane_c) {
  constexpr int JN = CRT_JN, PER = CRT_PER, R = CRT_TPR;
  __shared__ __attribute__((aligned(16))) signed char L[CRT_NBUF][CRT_NBL * CRT_BLK];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int w = xcd_contiguous((int)blockIdx.x, cc.nmod * ntiles);
  const int mod = w / ntiles;
  int t = w - mod * ntiles, tm, tn;
  if (tri & GEMM_C_LOWER) {  // lower tiles row by row: row tm holds R (tm + 1) of them
    auto before = [](int r) { return R * r * (r + 1) / 2; };
    tm = (int)((sqrt(8.0 * t / R + 1.0) - 1.0) * 0.5);
    while (before(tm + 1) <= t) ++tm;
    while (before(tm) > t) --tm;
    tn = t - before(tm);
  } else {  // groups of four tile rows, column by column inside a group
    const int per = 4 * tiles_n, g0 = (t / per) * 4, gs = min(4, tiles_m - g0);
    tm = g0 + (t % per) % gs;
    tn = (t % per) / gs;
  }
  const int m0 = tm * CRT_T, n0 = tn * CRT_BN;
  int kbeg = 0, kend = Kp;
  if (tri & GEMM_A_LOWER) kend = min(kend, m0 + CRT_T);
  if (tri & GEMM_BT_LOWER) kend = min(kend, n0 + CRT_BN);
  if (tri & GEMM_B_LOWER) kbeg = n0;
  const int kblocks = Kp / CRT_KB;
  const int s0 = kbeg / CRT_KB, nsteps = max(0, (kend - kbeg + CRT_KB - 1) / CRT_KB);
  // this wave's blocks of a step: PER of the NBL (A blocks 0..15, then B)
  const signed char *pa = PA + (size_t)mod * plane_a + (size_t)(m0 / 16) * kblocks * CRT_BLK;
  const signed char *pb = PB + (size_t)mod * plane_b + (size_t)(n0 / 16) * kblocks * CRT_BLK;
  const signed char *src[PER];
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int blk = wave * PER + b;
    src[b] = (blk < 16 ? pa + (size_t)blk * kblocks * CRT_BLK
                       : pb + (size_t)(blk - 16) * kblocks * CRT_BLK) + 16 * lane;
  }
  auto issue = [&](int q) {
    signed char *dst = L[q % CRT_NBUF] + PER * wave * CRT_BLK;
#pragma unroll
    for (int b = 0; b < PER; ++b)
      __builtin_amdgcn_global_load_lds((const void *)(src[b] + (size_t)(s0 + q) * CRT_BLK),
                                       (lds_void *)(dst + b * CRT_BLK), 16, 0, 0);
  };
  v4i acc[8][JN];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) acc[i][j] = (v4i){0, 0, 0, 0};
  // Steps run through the three LDS buffers: while step q is multiplied, step
  // q + 1 is being copied in; step r + 2's copy is issued at the barrier before
  // step r is read, into the buffer of step r - 1, whose fragment reads every
  // wave finished before that barrier.
  auto sync_step = [&](int r) {
    // this wave's fragment reads are done and its copies of step r have landed
    // (vector memory operations complete in order; the PER of step r + 1 may
    // be in flight); the barrier then makes every wave's visible
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    if (r + 1 < nsteps) __builtin_amdgcn_s_waitcnt(vmcnt_imm(PER));
    else __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
    __builtin_amdgcn_s_barrier();
    if (r + CRT_NBUF - 1 < nsteps) issue(r + CRT_NBUF - 1);
  };
#pragma unroll
  for (int q = 0; q < CRT_NBUF - 1; ++q)
    if (q < nsteps) issue(q);
  for (int q = 0; q < nsteps; ++q) {
    sync_step(q);
    const signed char *Lb = L[q % CRT_NBUF];
    // the A fragments first, then each B fragment read while the eight MFMAs
    // of the one before it run
    v4i af[8], bq[2];
#pragma unroll
    for (int i = 0; i < 8; ++i) af[i] = *(const v4i *)(Lb + (wm * 8 + i) * CRT_BLK + 16 * lane);
    bq[0] = *(const v4i *)(Lb + (16 + wn * JN) * CRT_BLK + 16 * lane);
    __builtin_amdgcn_sched_group_barrier(0x100, 9, 0);  // 9 LDS reads
#pragma unroll
    for (int j = 0; j < JN; ++j) {
      if (j + 1 < JN) {
        bq[(j + 1) & 1] = *(const v4i *)(Lb + (16 + wn * JN + j + 1) * CRT_BLK + 16 * lane);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // 1 LDS read
      }
#pragma unroll
      for (int i = 0; i < 8; ++i)
        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bq[j & 1], acc[i][j], 0, 0, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);  // 8 MFMAs
    }
  }
  const int m = cc.m[mod];
  const double mi = cc.minv[mod];
  unsigned char *pc = PC + (size_t)mod * plane_c;
  // C plane in 16 x 16 blocks (column-block major), each the 256 bytes of one
  // MFMA output block in lane order (the word at 4 p holds rows 4 (p / 16) ..
  // + 3 of column p % 16): one contiguous store per block and wave
  const size_t rblocks = (size_t)Mp / 16;
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) {
      const size_t rb = (size_t)(m0 + wm * 128 + i * 16) / 16;
      const size_t cb = (size_t)(n0 + wn * (CRT_BN / 2) + j * 16) / 16;
      const unsigned int u = crt_reduce(acc[i][j][0], m, mi) |
                             crt_reduce(acc[i][j][1], m, mi) << 8 |
                             crt_reduce(acc[i][j][2], m, mi) << 16 |
                             crt_reduce(acc[i][j][3], m, mi) << 24;
      *(unsigned int *)(pc + (cb * rblocks + rb) * 256 + 4 * lane) = u;
    }
}

// 2^s mod m (s >= 0) by 16-bit steps: c16 = 2^16 mod m
__device__ __forceinline__ unsigned int crt_pow2(int s, int m, double minv, int c16) {
  unsigned int v = crt_reduce(1 << (s & 15), m, minv);
  for (int t = s >> 4; t > 0; --t) v = crt_reduce((int)v * c16, m, minv);
  return v;
}

// C_ij = alpha 2^-(ea_i + eb_j) (crt(r_1 .. r_s) - D_ij 2^(ea_i + eb_j)) + beta C_ij,
// D_ij = 1 where dshift and i + doff == j (the identity, subtracted exactly in
// the residues, so that I - A X keeps the digits its cancellation leaves), else 0.
// One thread per four rows of one column, read from the blocked C planes as one
// 32-bit word per plane (the lane's word of a 16 x 16 block); 256 threads take
// four row blocks of one column block.  The CRT sum sum_l r_l E_l is formed on
// NL 16-bit limbs of E_l: s_k = sum_l r_l e_lk < 20 * 255 * 65535 < 2^29 with
// one 24-bit integer multiply-add per term (full rate), carried into NW = (NL +
// 2) / 2 32-bit words, reduced modulo P (quotient from FP64, then exact word
// arithmetic) and taken in the symmetric range.  With `lower` only i >= j is
// written.
template <int NL>
__global__ __launch_bounds__(256) void k_crt_rebuild(int M, int N, int Mp, int lower, double alpha,
                                                     double beta, int dshift, int doff,
                                                     CrtConst cc,
                                                     const unsigned char *__restrict__ PC,
                                                     size_t plane_c, const int *__restrict__ ea,
                                                     const int *__restrict__ eb, double *C,
                                                     int ldc) {
  constexpr int NW = (NL + 2) / 2;
  const int lane = threadIdx.x & 63;
  const int rb = blockIdx.x * 4 + (threadIdx.x >> 6), cb = blockIdx.y;
  const int j = cb * 16 + (lane & 15), i0 = rb * 16 + 4 * (lane >> 4);
  if (i0 >= M || j >= N) return;
  if (lower && i0 + 3 < j) return;
  const size_t at = ((size_t)cb * (Mp / 16) + rb) * 256 + 4 * lane;
  const int ej = eb[j];
  // the diagonal element among this thread's four rows, if any, and its scale
  const int bd = dshift ? j - doff - i0 : -1;
  const bool has_d = bd >= 0 && bd < 4 && i0 + bd < M;
  const int sd = has_d ? ea[i0 + bd] + ej : -1;
  const bool shift_res = has_d && sd >= 0;
  unsigned int sk[4][NL] = {};  // [row][limb]
  for (int l = 0; l < cc.nmod; ++l) {
    unsigned int u = *(const unsigned int *)(PC + (size_t)l * plane_c + at);
    if (shift_res) {  // r <- r - 2^sd mod m in the diagonal row's byte
      const int m = cc.m[l];
      const unsigned int r = __builtin_amdgcn_ubfe(u, 8 * bd, 8);
      const unsigned int p2 = crt_pow2(sd, m, cc.minv[l], cc.c16[l]);
      const unsigned int nr = r >= p2 ? r - p2 : r + (unsigned int)m - p2;
      u = (u & ~(0xffu << (8 * bd))) | (nr << (8 * bd));
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const unsigned int r = __builtin_amdgcn_ubfe(u, 8 * b, 8);
#pragma unroll
      for (int k = 0; k < NL; ++k) sk[b][k] += __umul24(r, cc.e16[l][k]);
    }
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int i = i0 + b;
    if (i >= M || (lower && i < j)) continue;
    // carry the limb sums into NW words
    unsigned int x[NW] = {};
    unsigned int c = 0;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      c += sk[b][k];
      x[k >> 1] |= (c & 0xffffu) << (16 * (k & 1));
      c >>= 16;
    }
    x[NL >> 1] |= c << (16 * (NL & 1));
    // x mod P: the quotient (< 2^14) from FP64, one below to stay non-negative
    double xd = 0.0;
#pragma unroll
    for (int k = NW - 1; k >= 0; --k) xd = xd * 4294967296.0 + (double)x[k];
    // (the estimate of floor(x / P) is off by at most one, so 0 <= x - q P < 3 P)
    const double qd = floor(xd * cc.pinv) - 1.0;
    const unsigned long long q = qd > 0.0 ? (unsigned long long)qd : 0ull;
    long long borrow = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const long long t = (long long)x[k] - (long long)(q * cc.p[k]) - borrow;
      x[k] = (unsigned int)t;
      borrow = -(t >> 32);
    }
    // at most two more subtractions of P, branch-free: the difference replaces
    // x where it did not borrow
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      unsigned int t[NW], br = 0;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        const unsigned long long d = (unsigned long long)x[k] - cc.p[k] - br;
        t[k] = (unsigned int)d;
        br = (unsigned int)(d >> 63);
      }
#pragma unroll
      for (int k = 0; k < NW; ++k) x[k] = br ? x[k] : t[k];
    }
    // symmetric range: x >= P / 2 stands for x - P = -(P - x)
    unsigned int y[NW], by = 0, bh = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const unsigned long long d = (unsigned long long)cc.p[k] - x[k] - by;
      y[k] = (unsigned int)d;
      by = (unsigned int)(d >> 63);
      const unsigned long long e = (unsigned long long)x[k] - cc.ph[k] - bh;
      bh = (unsigned int)(e >> 63);
    }
    const bool neg = bh == 0;
    double vx = 0.0, vy = 0.0;
#pragma unroll
    for (int k = NW - 1; k >= 0; --k) {
      vx = vx * 4294967296.0 + (double)x[k];
      vy = vy * 4294967296.0 + (double)y[k];
    }
    const double v = neg ? -vy : vx;
    const int sc = ea[i] + ej;
    double out;
    if (has_d && b == bd && !shift_res) out = alpha * (ldexp(v, -sc) - 1.0);
    else out = alpha * ldexp(v, -sc);
    double *cp = C + (size_t)i + (size_t)j * ldc;
    if (beta != 0.0) out = fma(beta, *cp, out);
    *cp = out;
  }
}

size_t r256(size_t b) { return (b + 255) / 256 * 256; }

const CrtConst &crt_const(int nmod) {
  static const CrtConst table[CRT_MAX_MOD + 1] = {
      make_const(1),  make_const(1),  make_const(2),  make_const(3),  make_const(4),
      make_const(5),  make_const(6),  make_const(7),  make_const(8),  make_const(9),
      make_const(10), make_const(11), make_const(12), make_const(13), make_const(14),
      make_const(15), make_const(16), make_const(17), make_const(18), make_const(19),
      make_const(20)};
  return table[nmod];
}

// an operand's area: nmod residue planes (Rp x Kp), Rp exponents, norm partials
struct OperandLayout {
  size_t Rp, Kp, plane, planes, expo, part, total;
};

OperandLayout operand_layout(int R, int K, int nmod) {
  OperandLayout L{};
  L.Rp = (size_t)cdiv(R, CRT_T) * CRT_T;
  L.Kp = (size_t)cdiv(K, CRT_KB) * CRT_KB;
  L.plane = L.Rp * L.Kp;
  size_t o = 0;
  L.planes = o; o += r256(nmod * L.plane);
  L.expo = o; o += r256(sizeof(int) * L.Rp);
  L.part = o; o += r256(sizeof(double) * CRT_NCH * (size_t)R);
  L.total = o;
  return L;
}

size_t cplanes_bytes(int M, int N, int nmod) {
  return r256((size_t)nmod * (size_t)cdiv(M, CRT_T) * CRT_T * (size_t)cdiv(N, CRT_T) * CRT_T);
}

}  // namespace

int crt_bits(int nmod) {
  if (nmod < CRT_MIN_MOD || nmod > CRT_MAX_MOD) return 0;
  return crt_const(nmod).q;
}

size_t crt_operand_bytes(int R, int K, int nmod) {
  if (R <= 0 || K <= 0 || nmod < CRT_MIN_MOD || nmod > CRT_MAX_MOD) return 0;
  return operand_layout(R, K, nmod).total;
}

size_t crt_product_bytes(int M, int N, int nmod) {
  if (M <= 0 || N <= 0 || nmod < CRT_MIN_MOD || nmod > CRT_MAX_MOD) return 0;
  return cplanes_bytes(M, N, nmod);
}

hipError_t crt_split(int R, int K, const double *P, int ld, bool gather, int mask, int nmod,
                     void *area, CrtOperand *out, hipStream_t st) {
  if (R <= 0 || K <= 0 || K > CRT_MAX_K || nmod < CRT_MIN_MOD || nmod > CRT_MAX_MOD || !area ||
      ((uintptr_t)area & 255) != 0 || mask < 0 || mask > 2)
    return hipErrorInvalidValue;
  const OperandLayout Ly = operand_layout(R, K, nmod);
  const CrtConst &cc = crt_const(nmod);
  char *a = (char *)area;
  CrtOperand o;
  o.planes = (signed char *)(a + Ly.planes);
  o.expo = (int *)(a + Ly.expo);
  o.R = R; o.K = K; o.Rp = (int)Ly.Rp; o.Kp = (int)Ly.Kp; o.nmod = nmod;
  double *part = (double *)(a + Ly.part);
  if (gather) {
    const unsigned sg = cdiv((long long)(Ly.Rp / 16) * (Ly.Kp / CRT_KB), 4);
    hipLaunchKernelGGL(k_crt_norm_gather, dim3(cdiv(R, 64), CRT_NCH), dim3(256), 0, st, R, K, P,
                       ld, mask, part);
    if (cc.q > CRT_NARROW_Q)
      hipLaunchKernelGGL(k_crt_split_gather<true>, dim3(sg), dim3(256), 0, st, R, K, (int)Ly.Rp,
                         (int)Ly.Kp, P, ld, mask, part, CRT_NCH, cc, o.planes, Ly.plane, o.expo);
    else
      hipLaunchKernelGGL(k_crt_split_gather<false>, dim3(sg), dim3(256), 0, st, R, K, (int)Ly.Rp,
                         (int)Ly.Kp, P, ld, mask, part, CRT_NCH, cc, o.planes, Ly.plane, o.expo);
  } else {
    const unsigned sg = (unsigned)(Ly.Rp / 16) * cdiv((long long)(Ly.Kp / CRT_KB), 4);
    hipLaunchKernelGGL(k_crt_norm_contig, dim3(cdiv(R, 4)), dim3(256), 0, st, R, K, P, ld, mask,
                       part);
    if (cc.q > CRT_NARROW_Q)
      hipLaunchKernelGGL(k_crt_split_contig<true>, dim3(sg), dim3(256), 0, st, R, K, (int)Ly.Rp,
                         (int)Ly.Kp, P, ld, mask, part, cc, o.planes, Ly.plane, o.expo);
    else
      hipLaunchKernelGGL(k_crt_split_contig<false>, dim3(sg), dim3(256), 0, st, R, K, (int)Ly.Rp,
                         (int)Ly.Kp, P, ld, mask, part, cc, o.planes, Ly.plane, o.expo);
  }
  *out = o;
  return hipGetLastError();
}

hipError_t crt_product(double alpha, const CrtOperand &A, const CrtOperand &B, double beta,
                       double *C, int ldc, int tri, void *cplanes, size_t cplanes_size,
                       hipStream_t st, int doff) {
  const int M = A.R, N = B.R, nmod = A.nmod;
  if (M <= 0 || N <= 0) return hipSuccess;
  if (A.K != B.K || A.Kp != B.Kp || A.nmod != B.nmod || nmod < CRT_MIN_MOD ||
      nmod > CRT_MAX_MOD || !cplanes || ((uintptr_t)cplanes & 255) != 0 ||
      cplanes_size < cplanes_bytes(M, N, nmod) || ((tri & GEMM_C_LOWER) && M != N) ||
      (doff >= 0 && (tri & GEMM_C_LOWER)))
    return hipErrorInvalidValue;
  const CrtConst &cc = crt_const(nmod);
  unsigned char *pc = (unsigned char *)cplanes;
  const size_t plane_a = (size_t)A.Rp * A.Kp, plane_b = (size_t)B.Rp * B.Kp;
  const size_t plane_c = (size_t)A.Rp * B.Rp;
  const int tm = A.Rp / CRT_T, tn = B.Rp / CRT_BN;
  const int ntiles = (tri & GEMM_C_LOWER) ? CRT_TPR * tm * (tm + 1) / 2 : tm * tn;
  hipLaunchKernelGGL(k_crt_gemm, dim3(nmod * ntiles), dim3(CRT_THREADS), 0, st, A.Rp, B.Rp, A.Kp,
                     tm, tn, ntiles, tri, cc, A.planes, B.planes, plane_a, plane_b, pc, plane_c);
#define UNC_CRT_REBUILD(NL)                                                                \
  hipLaunchKernelGGL(k_crt_rebuild<NL>, dim3(A.Rp / 64, B.Rp / 16), dim3(256), 0, st, M, N, A.Rp, \
                     (tri & GEMM_C_LOWER) ? 1 : 0, alpha, beta, doff >= 0 ? 1 : 0, doff, cc, pc, \
                     plane_c, A.expo, B.expo, C, ldc)
  switch (cc.nlimb) {
    case 3: UNC_CRT_REBUILD(3); break;
    case 4: UNC_CRT_REBUILD(4); break;
    case 5: UNC_CRT_REBUILD(5); break;
    case 6: UNC_CRT_REBUILD(6); break;
    case 7: UNC_CRT_REBUILD(7); break;
    case 8: UNC_CRT_REBUILD(8); break;
    case 9: UNC_CRT_REBUILD(9); break;
    case 10: UNC_CRT_REBUILD(10); break;
    default: return hipErrorInvalidValue;
  }
#undef UNC_CRT_REBUILD
  return hipGetLastError();
}

size_t crt_gemm_work_bytes(int M, int N, int K, int nmod) {
  if (M <= 0 || N <= 0 || K <= 0 || nmod < CRT_MIN_MOD || nmod > CRT_MAX_MOD) return 0;
  return crt_operand_bytes(M, K, nmod) + crt_operand_bytes(N, K, nmod) +
         crt_product_bytes(M, N, nmod);
}

hipError_t gemm_crt(int M, int N, int K, double alpha, const double *A, int lda, const double *B,
                    int ldb, bool trans_b, double beta, double *C, int ldc, int tri, int nmod,
                    void *work, size_t work_bytes, hipStream_t st) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (nmod < CRT_MIN_MOD || nmod > CRT_MAX_MOD || K <= 0 || K > CRT_MAX_K)
    return hipErrorInvalidValue;
  const size_t ba = crt_operand_bytes(M, K, nmod), bb = crt_operand_bytes(N, K, nmod);
  const size_t bc = crt_product_bytes(M, N, nmod);
  if (!work || work_bytes < ba + bb + bc || ((uintptr_t)work & 255) != 0)
    return hipErrorInvalidValue;
  char *w = (char *)work;
  CrtOperand oa, ob;
  // op(A): rows i, element (i, k) at A[i + k lda]
  hipError_t e = crt_split(M, K, A, lda, true, (tri & GEMM_A_LOWER) ? 1 : 0, nmod, w, &oa, st);
  if (e != hipSuccess) return e;
  // op(B) by columns j: B^T -> element (j, k) at B[j + k ldb]; B -> at B[k + j ldb]
  if (trans_b)
    e = crt_split(N, K, B, ldb, true, (tri & GEMM_BT_LOWER) ? 1 : 0, nmod, w + ba, &ob, st);
  else
    e = crt_split(N, K, B, ldb, false, (tri & GEMM_B_LOWER) ? 2 : 0, nmod, w + ba, &ob, st);
  if (e != hipSuccess) return e;
  return crt_product(alpha, oa, ob, beta, C, ldc, tri, w + ba + bb, bc, st);
}

// R = I - A X for symmetric A: rows of A are its columns, so both operands are
// cut from contiguous columns (crt_split, gather = false); A in row blocks of h
// rows so that the residue planes fit `work` (CRT_RESID_MOD moduli: q = 77, the
// exact product of the cut operands differs from A X by at most
// 2^-74 sqrt(n) |A_i.| |X_.j|, and the identity is taken off in the residues).
size_t residual_crt_work_bytes(int n, int blocks) {
  if (n <= 0 || blocks <= 0) return 0;
  const int h = (int)cdiv(n, blocks);
  return crt_operand_bytes(n, n, CRT_RESID_MOD) + crt_operand_bytes(h, n, CRT_RESID_MOD) +
         crt_product_bytes(h, n, CRT_RESID_MOD);
}

hipError_t residual_crt(int n, const double *A, const double *X, double *R, void *work,
                        size_t work_bytes, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (n > CRT_MAX_K || !work || ((uintptr_t)work & 255) != 0) return hipErrorInvalidValue;
  int blocks = 1;
  while (blocks <= 64 && residual_crt_work_bytes(n, blocks) > work_bytes) ++blocks;
  if (blocks > 64) return hipErrorInvalidValue;
  const int h = (int)cdiv(n, blocks);
  char *w = (char *)work;
  const size_t bx = crt_operand_bytes(n, n, CRT_RESID_MOD);
  const size_t ba = crt_operand_bytes(h, n, CRT_RESID_MOD);
  const size_t bc = crt_product_bytes(h, n, CRT_RESID_MOD);
  CrtOperand ox, oa;
  hipError_t e = crt_split(n, n, X, n, false, 0, CRT_RESID_MOD, w, &ox, st);
  for (int r0 = 0; e == hipSuccess && r0 < n; r0 += h) {
    const int hr = std::min(h, n - r0);
    e = crt_split(hr, n, A + (size_t)r0 * n, n, false, 0, CRT_RESID_MOD, w + bx, &oa, st);
    if (e == hipSuccess)
      e = crt_product(-1.0, oa, ox, 0.0, R + r0, n, 0, w + bx + ba, bc, st, r0);
  }
  return e;
}

}  // namespace unc

