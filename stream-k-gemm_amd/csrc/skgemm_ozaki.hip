// Split-precision FP64 GEMM / SYRK on the INT8 matrix cores (Ozaki scheme
// II: Ozaki, Uchino & Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of
// floating-point matrix multiplication using an integer modular technique").
// No counterpart in the reference, whose path is FP64 only.
//
// C = A B (A: M x K, B: K x N; for SYRK A = Psi^T, B = Psi):
//  1. scale: sigma_o = h - 1 - ceil(e_o / 2) with ||row/column o||_2^2 =
//     f 2^e_o, so that A'_o = rint(A_o 2^sigma_o) has ||A'_o||_2 <= 2^h and
//     |(A'B')_ij| <= 2^(2h) <= P/4, P = product of the moduli (Cauchy-Schwarz);
//  2. split: for L pairwise-coprime moduli m_l <= 256, the symmetric
//     residues of A' and B' (in [-128, 127]) as INT8 planes, K contiguous;
//  3. L INT8 products on v_mfma_i32_32x32x32_i8 with exact INT32
//     accumulation (|sum| <= Kpad * 2^14 < 2^31), reduced mod m_l, in one
//     persistent Stream-K launch over (modulus, tile) units;
//  4. CRT: Garner's mixed-radix digits (symmetric), exact 128-bit Horner,
//     one rounding to FP64, exact power-of-two unscaling, epilogue
//     alpha*C + beta*D + gamma.
// Integer steps are exact, so a result does not depend on the partition of
// the work (bit-identical for any grid).  The only approximation is step 1's
// rounding to integers; its a-posteriori bound is returned with the result.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/skgemm.h"
#include "skgemm_ctx.h"
#include "skgemm_internal.h"

namespace skg {
namespace oz {

constexpr int kMaxMod = 20;
// Pairwise coprime: 256 = 2^8, 255 = 3*5*17, 253 = 11*23, 247 = 13*19,
// 217 = 7*31, the others prime.
static const int kModuli[kMaxMod] = {256, 255, 253, 251, 247, 241, 239,
                                     233, 229, 227, 223, 217, 211, 199,
                                     197, 193, 191, 181, 179, 173};
constexpr int kMaxKpad = 131072;  // Kpad * 128^2 <= 2^31

struct Tables {
  int L;
  int h;
  int m[kMaxMod];
  double md[kMaxMod];     // m_l
  double inv_m[kMaxMod];  // fl(1 / m_l)
  double c26[kMaxMod];    // 2^26 mod m_l, symmetric
  double c52[kMaxMod];    // 2^52 mod m_l, symmetric
  uint8_t cinv[kMaxMod][kMaxMod];  // m_j^{-1} mod m_l (j < l), Garner form
  // Direct CRT (L <= 16, P < 2^128): x = sum_l r_l W_l - q P (mod 2^128),
  // W_l = (P / m_l) ((P / m_l)^{-1} mod m_l), q = floor(sum_l r_l w_l + 1/2)
  // with w_l = W_l / P rounded to FP64.
  uint64_t W_lo[kMaxMod], W_hi[kMaxMod];
  double wfrac[kMaxMod];
  uint64_t P_lo, P_hi;
  int direct;
};

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(1))) int gint;

constexpr int kRsrcWord3 = 0x00020000;
constexpr int kNumRecords = 0x7FFFFF00;
// Cache-policy operand of buffer loads/stores: sc1, the device-coherent
// form of a relaxed agent-scope atomic load/store on gfx950 (it bypasses
// the non-coherent per-CU cache and is served by the XCD L2).  The buffer
// form takes a scalar offset per element instead of a 64-bit address.
constexpr int kAgentCoherent = 16;

__device__ __forceinline__ int64_t op_addr(const Operand& op, int o, int k) {
  const int pm = (1 << op.pshift) - 1;
  return (int64_t)(o >> op.pshift) * op.pstride + (int64_t)(o & pm) * op.os +
         (int64_t)k * op.ks;
}

__device__ __forceinline__ int64_t out_addr(const Output& d, int m, int n) {
  const int pm = (1 << d.pshift) - 1;
  return (int64_t)(n >> d.pshift) * d.pstride + (int64_t)(n & pm) * d.ns +
         (int64_t)m * d.ms;
}

__device__ __forceinline__ int sigma_of(double norm2, int h) {
  if (!(norm2 > 0.0)) return 0;
  int e;
  frexp(norm2, &e);
  return h - 1 - ((e + 1) >> 1);  // ceil(e / 2) for any sign of e
}

// Rows whose squared norm leaves [2^-1000, 2^1000] (or is not finite, or
// underflows to zero with non-zero entries) cannot be scaled into the
// integer range; such a call is computed by the plain FP64 path instead.
__device__ __forceinline__ bool needs_fallback(double norm2, double maxabs) {
  const double lo = 9.332636185032189e-302;  // 2^-1000
  const double hi = 1.0715086071862673e+301;  // 2^1000
  if (!(norm2 <= hi)) return true;             // inf or NaN
  if (maxabs != 0.0 && !(norm2 >= lo)) return true;
  return false;
}

__device__ __forceinline__ void finish_norm(double s, double mx, int o, int h,
                                            double* norm2, int* sigma,
                                            int* fallback) {
  norm2[o] = s;
  sigma[o] = sigma_of(s, h);
  if (needs_fallback(s, mx)) atomicOr(fallback, 1);
}

// ---- 1. norms and scale exponents -------------------------------------------

// Operand contiguous along k: one workgroup per free index o, k ascending
// (the residue pass walks k descending so that its first reads hit what
// this pass left in the memory-side cache).
__global__ __launch_bounds__(256) void norm_k_kernel(Operand op, int K, int h,
                                                     double* norm2, int* sigma,
                                                     int* fallback) {
  __shared__ double red[2][256];
  const int o = blockIdx.x;
  const double* row = op.ptr + op_addr(op, o, 0);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, mx = 0.0;
  int k = threadIdx.x;
  for (; k + 768 < K; k += 1024) {
    const double x0 = row[k], x1 = row[k + 256], x2 = row[k + 512],
                 x3 = row[k + 768];
    s0 = fma(x0, x0, s0);
    s1 = fma(x1, x1, s1);
    s2 = fma(x2, x2, s2);
    s3 = fma(x3, x3, s3);
    mx = fmax(mx, fmax(fmax(fabs(x0), fabs(x1)), fmax(fabs(x2), fabs(x3))));
  }
  for (; k < K; k += 256) {
    const double x = row[k];
    s0 = fma(x, x, s0);
    mx = fmax(mx, fabs(x));
  }
  red[0][threadIdx.x] = (s0 + s1) + (s2 + s3);
  red[1][threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) finish_norm(red[0][0], red[1][0], o, h, norm2, sigma,
                                    fallback);
}

// Operand contiguous along o: partial sums over k ranges (grid.y), then a
// fixed-order finish, so the norms do not depend on scheduling.
__global__ __launch_bounds__(256) void norm_o_partial_kernel(
    Operand op, int K, int kseg, double* part, double* partmax) {
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o >= op.extent) return;
  const int k0 = blockIdx.y * kseg;
  const int k1 = min(K, k0 + kseg);
  double s = 0.0, mx = 0.0;
  for (int k = k0; k < k1; ++k) {
    const double x = op.ptr[op_addr(op, o, k)];
    s = fma(x, x, s);
    mx = fmax(mx, fabs(x));
  }
  part[(int64_t)blockIdx.y * op.extent + o] = s;
  partmax[(int64_t)blockIdx.y * op.extent + o] = mx;
}

__global__ __launch_bounds__(256) void norm_o_finish_kernel(
    int O, int nseg, int h, const double* part, const double* partmax,
    double* norm2, int* sigma, int* fallback) {
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o >= O) return;
  double s = 0.0, mx = 0.0;
  for (int j = 0; j < nseg; ++j) {
    s += part[(int64_t)j * O + o];
    mx = fmax(mx, partmax[(int64_t)j * O + o]);
  }
  finish_norm(s, mx, o, h, norm2, sigma, fallback);
}

// ---- 2. residue planes -----------------------------------------------------------

// Residues of an exact integer x (|x| < 2^54, as FP64) modulo every m_l.
// x = xh 2^26 + xl exactly (|xl| <= 2^25, |xh| <= 2^28); then
// y = xh (2^26 mod m) + xl is exact with |y| < 2^36, and for odd m the FP64
// quotient rint(y * fl(1/m)) is the nearest integer to y / m (its error is
// below 2^-24 while y / m lies at least 1/(2m) from a half-integer), so
// r = y - q m is the symmetric residue in [-(m-1)/2, (m-1)/2].  For m = 256
// the result is in [-128, 128]; +128 and -128 share the byte 0x80, which is
// the same residue.  Five FP64 operations per residue, all exact.
// With more than 14 moduli h exceeds 54 and |x| reaches 2^76: then
// x = x2 2^52 + xh 2^26 + xl (all limbs <= 2^25 in magnitude) and
// y = x2 (2^52 mod m) + xh (2^26 mod m) + xl, |y| < 2^33, one more FMA.
struct Split {
  double x2, xh, xl;
};

template <bool WIDE>
__device__ __forceinline__ Split split26(double x) {
  Split p;
  p.x2 = 0.0;
  if (WIDE) {
    p.x2 = rint(x * 2.220446049250313e-16);  // 2^-52
    x = fma(-p.x2, 4503599627370496.0, x);    // 2^52
  }
  p.xh = rint(x * 1.4901161193847656e-08);  // 2^-26
  p.xl = fma(-p.xh, 67108864.0, x);         // 2^26
  return p;
}

template <bool WIDE>
__device__ __forceinline__ int residue(const Split& p, double m, double inv,
                                       double c, double c2) {
  double y = fma(p.xh, c, p.xl);
  if (WIDE) y = fma(p.x2, c2, y);
  const double q = rint(y * inv);
  return (int)fma(-q, m, y);
}

__device__ __forceinline__ uint32_t pack4(int r0, int r1, int r2, int r3) {
  return (uint32_t)(r0 & 255) | ((uint32_t)(r1 & 255) << 8) |
         ((uint32_t)(r2 & 255) << 16) | ((uint32_t)r3 << 24);
}

// planes[l][o][k] (row pitch kpad bytes) for o < opad, k < kpad; zeros past
// the matrix.  Operand contiguous along k: workgroup (o, k segment of
// 4096), 4 consecutive k per thread and step, segments taken from high k
// to low k.
constexpr int kSplitSeg = 4096;

template <bool WIDE>
__global__ __launch_bounds__(256) void split_k_kernel(
    Operand op, int K, const int* sigma, Tables T, int8_t* planes,
    int64_t plane, int kpad) {
  const int nseg = (kpad + kSplitSeg - 1) / kSplitSeg;
  const int o = blockIdx.x / nseg;
  const int seg = nseg - 1 - (blockIdx.x - o * nseg);
  const bool in_o = o < op.extent;
  const int s = in_o ? sigma[o] : 0;
  const double* row = in_o ? op.ptr + op_addr(op, o, 0) : op.ptr;
  int8_t* dst = planes + (int64_t)o * kpad;
  for (int step = 3; step >= 0; --step) {
    const int k0 = seg * kSplitSeg + step * 1024 + threadIdx.x * 4;
    if (k0 >= kpad) continue;
    Split p[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + j;
      const double a = (in_o && k < K) ? row[k] : 0.0;
      p[j] = split26<WIDE>(rint(ldexp(a, s)));
    }
    for (int l = 0; l < T.L; ++l) {
      const double m = T.md[l], inv = T.inv_m[l], c = T.c26[l], c2 = T.c52[l];
      const uint32_t v = pack4(residue<WIDE>(p[0], m, inv, c, c2),
                               residue<WIDE>(p[1], m, inv, c, c2),
                               residue<WIDE>(p[2], m, inv, c, c2),
                               residue<WIDE>(p[3], m, inv, c, c2));
      *reinterpret_cast<uint32_t*>(dst + l * plane + k0) = v;
    }
  }
}

// Operand contiguous along o: workgroup of 64 o x 128 k; thread (o, 32
// consecutive k) loads along o (coalesced) and writes 32 bytes per plane,
// so each workgroup writes whole 128-byte rows of every plane.
template <bool WIDE>
__global__ __launch_bounds__(256) void split_o_kernel(
    Operand op, int K, const int* sigma, Tables T, int8_t* planes,
    int64_t plane, int kpad, int opad) {
  const int kblocks = kpad / 128;
  const int ob = blockIdx.x / kblocks;
  const int kb = kblocks - 1 - (blockIdx.x - ob * kblocks);
  const int o = ob * 64 + (threadIdx.x & 63);
  const int k0 = kb * 128 + (threadIdx.x >> 6) * 32;
  if (o >= opad) return;
  const bool in_o = o < op.extent;
  const int s = in_o ? sigma[o] : 0;
  Split p[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    const int k = k0 + j;
    const double a = (in_o && k < K) ? op.ptr[op_addr(op, o, k)] : 0.0;
    p[j] = split26<WIDE>(rint(ldexp(a, s)));
  }
  int8_t* dst = planes + (int64_t)o * kpad + k0;
  for (int l = 0; l < T.L; ++l) {
    const double m = T.md[l], inv = T.inv_m[l], c = T.c26[l], c2 = T.c52[l];
    uint32_t w[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
      w[q] = pack4(residue<WIDE>(p[4 * q], m, inv, c, c2),
                   residue<WIDE>(p[4 * q + 1], m, inv, c, c2),
                   residue<WIDE>(p[4 * q + 2], m, inv, c, c2),
                   residue<WIDE>(p[4 * q + 3], m, inv, c, c2));
    uint4* d4 = reinterpret_cast<uint4*>(dst + l * plane);
    d4[0] = make_uint4(w[0], w[1], w[2], w[3]);
    d4[1] = make_uint4(w[4], w[5], w[6], w[7]);
  }
}

// ---- 3. INT8 GEMM over (modulus, tile) units ----------------------------------

struct GemmLaunch {
  const int8_t* A;
  const int8_t* B;
  int64_t a_plane, b_plane;  // bytes per plane
  int kpad;                  // bytes per row
  int M, N, syrk;
  int tiles_m, tiles_n;
  int64_t ntiles;  // per plane
  int ipt;         // kpad / BK
  int L;
  int mods[kMaxMod];
  double inv_mods[kMaxMod];
  uint8_t* R;  // [L][M][N] residues in [0, m)
  int64_t r_plane;
  int grid;
  int dp;          // 1: each workgroup takes whole units (no fix-up)
  int64_t units;   // L * ntiles
  int64_t total;   // units * ipt (Stream-K mode)
  int* ws;
  int* flags;
  int* status;
  int64_t spin_limit;
};

__device__ __forceinline__ int64_t wg_start(int w, int64_t total, int G) {
  const int64_t base = total / G, extra = total % G;
  return (int64_t)w * base + ((int64_t)w < extra ? (int64_t)w : extra);
}

__device__ __forceinline__ void tri_tile(int64_t t, int& ti, int& tj) {
  int64_t i = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (i * (i + 1) / 2 > t) --i;
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  ti = (int)i;
  tj = (int)(t - i * (i + 1) / 2);
}

// LDS row of BK bytes in 16-byte chunks; chunk c of row r is stored at
// chunk c ^ f(r) so that the 8 rows one 128-byte LDS cycle serves (a
// ds_read_b128 of 32 rows at one chunk index, or a row-wise store) fall in
// 8 distinct 16-byte bank groups.
template <int BK>
__device__ __forceinline__ int lds_off(int r, int c) {
  if (BK == 128) return r * 128 + ((c ^ (r & 7)) << 4);
  return r * 64 + ((c ^ ((r >> 1) & 3)) << 4);
}

template <int BM, int BN, int WM, int WN, int BK>
__global__ __launch_bounds__(WM* WN * 64) void gemm_i8_kernel(GemmLaunch G) {
  constexpr int NT = WM * WN * 64;
  constexpr int TM = BM / WM, TN = BN / WN;
  constexpr int MB = TM / 32, NB = TN / 32;
  constexpr int CPR = BK / 16;                         // chunks per row
  constexpr int CA = BM * CPR / NT, CB = BN * CPR / NT;  // chunks / thread
  static_assert(BM * CPR % NT == 0 && BN * CPR % NT == 0, "tile vs threads");
  extern __shared__ __align__(16) int8_t lds[];  // [2][(BM + BN) * BK]

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave % WM, wn = wave / WM;
  const int Gn = G.grid, g = blockIdx.x;
  const int per = Gn >> 3, rem = Gn & 7, xcd = g & 7;
  const int w = xcd * per + (xcd < rem ? xcd : rem) + (g >> 3);

  int rowA[CA], chA[CA], rowB[CB], chB[CB];
#pragma unroll
  for (int p = 0; p < CA; ++p) {
    const int q = p * NT + tid;
    rowA[p] = q / CPR;
    chA[p] = q % CPR;
  }
#pragma unroll
  for (int p = 0; p < CB; ++p) {
    const int q = p * NT + tid;
    rowB[p] = q / CPR;
    chB[p] = q % CPR;
  }

  v16i acc[MB][NB];
  v4i ra[CA], rb[CB];

  // MAC steps [kb, ke) of unit `unit` into acc.
  auto mac = [&](int64_t unit, int kb, int ke, int& l, int& m0, int& n0) {
    l = (int)(unit / G.ntiles);
    const int64_t tile = unit - (int64_t)l * G.ntiles;
    int ti, tj;
    if (G.syrk) {
      tri_tile(tile, ti, tj);
    } else {
      ti = (int)(tile / G.tiles_n);
      tj = (int)(tile % G.tiles_n);
    }
    m0 = ti * BM;
    n0 = tj * BN;
    const int8_t* baseA = G.A + l * G.a_plane + (int64_t)m0 * G.kpad;
    const int8_t* baseB = G.B + l * G.b_plane + (int64_t)n0 * G.kpad;
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[i][j][v] = 0;

    auto load = [&](int kk) {
      const __amdgpu_buffer_rsrc_t sa = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<int8_t*>(baseA + (int64_t)kk * BK), (short)0,
          kNumRecords, kRsrcWord3);
      const __amdgpu_buffer_rsrc_t sb = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<int8_t*>(baseB + (int64_t)kk * BK), (short)0,
          kNumRecords, kRsrcWord3);
#pragma unroll
      for (int p = 0; p < CA; ++p)
        ra[p] = __builtin_amdgcn_raw_buffer_load_b128(
            sa, rowA[p] * G.kpad + chA[p] * 16, 0, 0);
#pragma unroll
      for (int p = 0; p < CB; ++p)
        rb[p] = __builtin_amdgcn_raw_buffer_load_b128(
            sb, rowB[p] * G.kpad + chB[p] * 16, 0, 0);
    };
    auto store = [&](int buf) {
      int8_t* base = lds + buf * (BM + BN) * BK;
#pragma unroll
      for (int p = 0; p < CA; ++p)
        *reinterpret_cast<v4i*>(base + lds_off<BK>(rowA[p], chA[p])) = ra[p];
#pragma unroll
      for (int p = 0; p < CB; ++p)
        *reinterpret_cast<v4i*>(base + BM * BK + lds_off<BK>(rowB[p], chB[p])) =
            rb[p];
    };
    load(kb);
    store(0);
    __syncthreads();
    int cur = 0;
    for (int kk = kb; kk < ke; ++kk) {
      const bool more = kk + 1 < ke;
      if (more) load(kk + 1);
      const int8_t* la = lds + cur * (BM + BN) * BK;
      const int8_t* lb = la + BM * BK;
#pragma unroll
      for (int s = 0; s < BK / 32; ++s) {
        const int c = 2 * s + (lane >> 5);
        v4i a[MB], b[NB];
#pragma unroll
        for (int i = 0; i < MB; ++i)
          a[i] = *reinterpret_cast<const v4i*>(
              la + lds_off<BK>(wm * TM + i * 32 + (lane & 31), c));
#pragma unroll
        for (int j = 0; j < NB; ++j)
          b[j] = *reinterpret_cast<const v4i*>(
              lb + lds_off<BK>(wn * TN + j * 32 + (lane & 31), c));
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i], b[j],
                                                              acc[i][j], 0, 0, 0);
      }
      if (more) store(cur ^ 1);
      __syncthreads();
      cur ^= 1;
    }
  };

  // acc mod m_l -> R[l] (lower triangle only for SYRK).  acc is exact in
  // FP64 and floor(acc * fl(1/m)) is off by at most one from floor(acc / m)
  // (only when acc / m is an integer), so one correction each way suffices.
  auto write_residues = [&](int l, int m0, int n0) {
    const int m = G.mods[l];
    const double inv = G.inv_mods[l];
    uint8_t* rp = G.R + l * G.r_plane;
    // Accumulator element v of lane x is row 8(v/4) + 4(x/32) + v%4,
    // column x%32 of its 32x32 block (measured on gfx950).
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int col = n0 + wn * TN + j * 32 + (lane & 31);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int row = m0 + wm * TM + i * 32 + (v >> 2) * 8 +
                          (lane >> 5) * 4 + (v & 3);
          if (row < G.M && col < G.N && (!G.syrk || row >= col)) {
            const int a = acc[i][j][v];
            int r = a - m * (int)floor((double)a * inv);
            r += (r < 0) ? m : 0;
            r -= (r >= m) ? m : 0;
            rp[(int64_t)row * G.N + col] = (uint8_t)r;
          }
        }
      }
  };

  if (G.dp) {
    // Whole units, round-robin: in each round the workgroups of one XCD
    // take consecutive units, i.e. tiles of one plane that share panels in
    // that XCD's L2 and advance through K together.
    for (int64_t u = w; u < G.units; u += Gn) {
      int l, m0, n0;
      mac(u, 0, G.ipt, l, m0, n0);
      write_residues(l, m0, n0);
    }
    return;
  }

  int64_t it = wg_start(w, G.total, Gn);
  const int64_t end = wg_start(w + 1, G.total, Gn);
  while (it < end) {
    const int64_t unit = it / G.ipt;
    const int64_t t0 = unit * G.ipt;
    const int kb = (int)(it - t0);
    const int ke = (int)((end - t0) < G.ipt ? (end - t0) : G.ipt);
    int l, m0, n0;
    mac(unit, kb, ke, l, m0, n0);

    // Stream-K fix-up on exact INT32 partials (device-coherent sc1 traffic,
    // no L2-wide fences; same protocol as the FP64 kernel).
    bool write_out = true;
    if (kb != 0) {
      const __amdgpu_buffer_rsrc_t slot = __builtin_amdgcn_make_buffer_rsrc(
          G.ws + (int64_t)w * (BM * BN), (short)0, kNumRecords, kRsrcWord3);
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int v = 0; v < 16; ++v)
            __builtin_amdgcn_raw_buffer_store_b32(
                acc[i][j][v], slot, tid * 4, ((i * NB + j) * 16 + v) * NT * 4,
                kAgentCoherent);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (tid == 0)
        __hip_atomic_store(&G.flags[w], 1, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      write_out = false;
    } else if (ke < G.ipt) {
      const int64_t unit_end = t0 + G.ipt;
      for (int w2 = w + 1; w2 < Gn && wg_start(w2, G.total, Gn) < unit_end;
           ++w2) {
        if (tid == 0) {
          int64_t n = 0;
          while (__hip_atomic_load(&G.flags[w2], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT) == 0) {
            __builtin_amdgcn_s_sleep(8);
            if (++n > G.spin_limit) {
              atomicOr(G.status, 1);
              break;
            }
          }
          __hip_atomic_store(&G.flags[w2], 0, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const __amdgpu_buffer_rsrc_t slot = __builtin_amdgcn_make_buffer_rsrc(
            G.ws + (int64_t)w2 * (BM * BN), (short)0, kNumRecords, kRsrcWord3);
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v)
              acc[i][j][v] += __builtin_amdgcn_raw_buffer_load_b32(
                  slot, tid * 4, ((i * NB + j) * 16 + v) * NT * 4,
                  kAgentCoherent);
      }
    }
    if (write_out) write_residues(l, m0, n0);
    it = t0 + ke;
  }
}

// ---- 4. CRT reconstruction and epilogue ----------------------------------------

struct CrtLaunch {
  const uint8_t* R;
  int64_t r_plane;
  Tables T;
  const int* sig_a;
  const int* sig_b;
  const int* fallback;  // set: the plain FP64 path writes this result
  Output d;
  int M, N, syrk;
  double alpha, beta, gamma;
  double* fro2;  // += ||C||_F^2 (SYRK: of the full symmetric matrix)
};

typedef unsigned __int128 u128;

__device__ __forceinline__ __int128 crt_direct(const CrtLaunch& p, int64_t idx) {
  u128 S = 0;
  double f = 0.5;
  for (int l = 0; l < p.T.L; ++l) {
    const uint32_t r = p.R[l * p.r_plane + idx];
    const u128 W = ((u128)p.T.W_hi[l] << 64) | p.T.W_lo[l];
    S += W * r;
    f = fma((double)r, p.T.wfrac[l], f);
  }
  const uint64_t q = (uint64_t)floor(f);
  const u128 P = ((u128)p.T.P_hi << 64) | p.T.P_lo;
  return (__int128)(S - P * q);
}

__device__ __forceinline__ __int128 crt_garner(const CrtLaunch& p, int64_t idx) {
  int v[kMaxMod];
  for (int l = 0; l < p.T.L; ++l) {
    const int m = p.T.m[l];
    int t = p.R[l * p.r_plane + idx];
    for (int j = 0; j < l; ++j) {
      t = (t - v[j]) % m;
      if (t < 0) t += m;
      t = (t * (int)p.T.cinv[j][l]) % m;
    }
    const int hi = (m - 1) >> 1;
    v[l] = t > hi ? t - m : t;
  }
  __int128 x = v[p.T.L - 1];
  for (int l = p.T.L - 2; l >= 0; --l) x = x * p.T.m[l] + v[l];
  return x;
}

__global__ __launch_bounds__(256) void crt_kernel(CrtLaunch p) {
  __shared__ double red[256];
  if (*p.fallback != 0) return;
  const int col = blockIdx.x * 64 + (threadIdx.x & 63);
  const int row = blockIdx.y * 4 + (threadIdx.x >> 6);
  double contrib = 0.0;
  if (row < p.M && col < p.N && (!p.syrk || row >= col)) {
    const int64_t idx = (int64_t)row * p.N + col;
    const __int128 x = p.T.direct ? crt_direct(p, idx) : crt_garner(p, idx);
    const long long xh = (long long)(x >> 64);
    const unsigned long long xl = (unsigned long long)x;
    const double d = (double)xh * 18446744073709551616.0 + (double)xl;
    const double c = ldexp(d, -(p.sig_a[row] + p.sig_b[col]));
    contrib = (p.syrk && row != col) ? 2.0 * c * c : c * c;
    double* dst = p.d.ptr + out_addr(p.d, row, col);
    double val = p.alpha * c;
    if (p.beta != 0.0) val = fma(p.beta, *dst, val);
    *dst = val + p.gamma;
  }
  red[threadIdx.x] = contrib;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] != 0.0) atomicAdd(p.fro2, red[0]);
}

// ---- host ----------------------------------------------------------------------

struct GemmCfg {
  int bm, bn, wm, wn, bk;
};
// 0: 256x256, 8 waves of 64x128, BK 128 (128 KB LDS, one workgroup per CU)
// 1: 128x128, 4 waves of 64x64,  BK 128 (64 KB LDS)
// 2: 128x128, 2 waves of 64x128, BK 128 (64 KB LDS)
// 3: 256x256, 8 waves of 64x128, BK 64  (64 KB LDS)
static const GemmCfg kGemmCfgs[] = {{256, 256, 4, 2, 128},
                                    {128, 128, 2, 2, 128},
                                    {128, 128, 2, 1, 128},
                                    {256, 256, 4, 2, 64}};
constexpr int kNumGemmCfgs = 4;

template <int BM, int BN, int WM, int WN, int BK>
static const void* kfn() {
  return reinterpret_cast<const void*>(gemm_i8_kernel<BM, BN, WM, WN, BK>);
}

static const void* gemm_fn(int cfg) {
  switch (cfg) {
    case 0: return kfn<256, 256, 4, 2, 128>();
    case 1: return kfn<128, 128, 2, 2, 128>();
    case 2: return kfn<128, 128, 2, 1, 128>();
    default: return kfn<256, 256, 4, 2, 64>();
  }
}

static size_t gemm_lds(int cfg) {
  const GemmCfg& g = kGemmCfgs[cfg];
  return (size_t)2 * (g.bm + g.bn) * g.bk;
}

static hipError_t gemm_attr(int cfg) {
  static bool done[kNumGemmCfgs] = {false, false, false, false};
  if (done[cfg]) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(
      gemm_fn(cfg), hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)gemm_lds(cfg));
  if (e == hipSuccess) done[cfg] = true;
  return e;
}

static int gemm_occupancy(int cfg) {
  if (gemm_attr(cfg) != hipSuccess) return 1;
  const GemmCfg& g = kGemmCfgs[cfg];
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &nb, gemm_fn(cfg), g.wm * g.wn * 64, gemm_lds(cfg)) != hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

static hipError_t launch_gemm(int cfg, const GemmLaunch& G, hipStream_t s) {
  hipError_t e = gemm_attr(cfg);
  if (e != hipSuccess) return e;
  const GemmCfg& g = kGemmCfgs[cfg];
  const dim3 grid(G.grid), block(g.wm * g.wn * 64);
  const size_t lds = gemm_lds(cfg);
  switch (cfg) {
    case 0:
      hipLaunchKernelGGL((gemm_i8_kernel<256, 256, 4, 2, 128>), grid, block,
                         lds, s, G);
      break;
    case 1:
      hipLaunchKernelGGL((gemm_i8_kernel<128, 128, 2, 2, 128>), grid, block,
                         lds, s, G);
      break;
    case 2:
      hipLaunchKernelGGL((gemm_i8_kernel<128, 128, 2, 1, 128>), grid, block,
                         lds, s, G);
      break;
    default:
      hipLaunchKernelGGL((gemm_i8_kernel<256, 256, 4, 2, 64>), grid, block,
                         lds, s, G);
      break;
  }
  return hipGetLastError();
}

static int modinv(int a, int m) {
  a %= m;
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 0;
}

static double sym_pow2_mod(int e, int m) {
  int r = 1;
  for (int i = 0; i < e; ++i) r = (r * 2) % m;
  return (double)(r > m / 2 ? r - m : r);
}

Tables make_tables(int L) {
  Tables T{};
  T.L = L;
  double bits = 0.0;
  for (int l = 0; l < L; ++l) {
    T.m[l] = kModuli[l];
    T.md[l] = (double)kModuli[l];
    T.inv_m[l] = 1.0 / kModuli[l];
    T.c26[l] = sym_pow2_mod(26, kModuli[l]);
    T.c52[l] = sym_pow2_mod(52, kModuli[l]);
    bits += std::log2((double)kModuli[l]);
  }
  for (int l = 0; l < L; ++l)
    for (int j = 0; j < l; ++j) T.cinv[j][l] = (uint8_t)modinv(kModuli[j], kModuli[l]);
  T.h = (int)std::floor(bits / 2.0) - 1;
  T.direct = bits < 127.0 ? 1 : 0;
  if (T.direct) {
    typedef unsigned __int128 u128;
    u128 P = 1;
    for (int l = 0; l < L; ++l) P *= (u128)kModuli[l];
    T.P_lo = (uint64_t)P;
    T.P_hi = (uint64_t)(P >> 64);
    for (int l = 0; l < L; ++l) {
      const u128 Pl = P / (u128)kModuli[l];
      const int inv = modinv((int)(Pl % (u128)kModuli[l]), kModuli[l]);
      const u128 W = Pl * (u128)inv;  // < P
      T.W_lo[l] = (uint64_t)W;
      T.W_hi[l] = (uint64_t)(W >> 64);
      // W / P = inv / m_l exactly; rounded once to FP64.
      T.wfrac[l] = (double)inv / (double)kModuli[l];
    }
  }
  return T;
}

}  // namespace oz
}  // namespace skg

// ---- context extras ------------------------------------------------------------

struct skgemm_ctx::Extra {
  static constexpr int kBufs = 12;
  void* buf[kBufs] = {nullptr};
  size_t n[kBufs] = {0};
  void* get(int i, size_t bytes) {
    if (bytes > n[i]) {
      if (buf[i]) {
        hipDeviceSynchronize();
        hipFree(buf[i]);
      }
      buf[i] = nullptr;
      n[i] = 0;
      if (hipMalloc(&buf[i], bytes) != hipSuccess) return nullptr;
      n[i] = bytes;
    }
    return buf[i];
  }
};

void skgemm_ctx::free_extra() {
  if (extra == nullptr) return;
  for (int i = 0; i < Extra::kBufs; ++i)
    if (extra->buf[i]) hipFree(extra->buf[i]);
  delete extra;
  extra = nullptr;
}

using namespace skg;
using namespace skg::oz;

namespace {

enum {
  kPlanesA = 0, kPlanesB, kRes, kNormA, kNormB, kSigA, kSigB, kScalars,
  kPartA, kPartB
};

int64_t round_up(int64_t x, int64_t q) { return (x + q - 1) / q * q; }

// Scale + split one operand (free index o, contraction k) into planes.
int prepare_operand(skgemm_ctx* c, const Operand& op, int K, const Tables& T,
                    int8_t* planes, int64_t plane, int kpad, int opad,
                    double* norm2, int* sigma, int* fallback, int part_buf) {
  const int O = op.extent;
  const bool wide = T.h > 54;
  if (op.kcontig) {
    if (O > 0)
      hipLaunchKernelGGL(norm_k_kernel, dim3(O), dim3(256), 0, c->stream, op,
                         K, T.h, norm2, sigma, fallback);
    const int nseg = (kpad + kSplitSeg - 1) / kSplitSeg;
    const dim3 grid((unsigned)((int64_t)opad * nseg));
    if (wide)
      hipLaunchKernelGGL(split_k_kernel<true>, grid, dim3(256), 0, c->stream,
                         op, K, sigma, T, planes, plane, kpad);
    else
      hipLaunchKernelGGL(split_k_kernel<false>, grid, dim3(256), 0, c->stream,
                         op, K, sigma, T, planes, plane, kpad);
  } else {
    if (O > 0) {
      const int nseg = std::max(1, std::min(64, (K + 255) / 256));
      const int kseg = (K + nseg - 1) / nseg;
      double* part = (double*)c->extra->get(part_buf,
                                            sizeof(double) * 2 * nseg * O);
      if (part == nullptr) return set_error(SKGEMM_ERR_NOMEM, "norm partials");
      hipLaunchKernelGGL(norm_o_partial_kernel, dim3((O + 255) / 256, nseg),
                         dim3(256), 0, c->stream, op, K, kseg, part,
                         part + (int64_t)nseg * O);
      hipLaunchKernelGGL(norm_o_finish_kernel, dim3((O + 255) / 256), dim3(256),
                         0, c->stream, O, nseg, T.h, part,
                         part + (int64_t)nseg * O, norm2, sigma, fallback);
    }
    const dim3 grid((unsigned)((int64_t)(opad / 64) * (kpad / 128)));
    if (wide)
      hipLaunchKernelGGL(split_o_kernel<true>, grid, dim3(256), 0, c->stream,
                         op, K, sigma, T, planes, plane, kpad, opad);
    else
      hipLaunchKernelGGL(split_o_kernel<false>, grid, dim3(256), 0, c->stream,
                         op, K, sigma, T, planes, plane, kpad, opad);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(e, "ozaki split");
  return SKGEMM_OK;
}

struct OzPlan {
  int cfg, dp, grid;
};

// Tile configuration and schedule of the INT8 product.  Data-parallel whole
// units when they fill the chip in near-even rounds (L2 reuse of panels
// between the tiles of one plane on one XCD); Stream-K otherwise.
// SKGEMM_OZ_PLAN="cfg,dp[,grid]" overrides (tuning sweeps); a grid forced
// with skgemm_set_tuning selects Stream-K on that grid.
OzPlan choose_plan(skgemm_ctx* c, int kind, int M, int N, int L) {
  OzPlan best{1, 0, 0};
  auto units_for = [&](int cfg) {
    const GemmCfg& g = kGemmCfgs[cfg];
    const int64_t tm = (M + g.bm - 1) / g.bm, tn = (N + g.bn - 1) / g.bn;
    return (int64_t)L * (kind == 1 ? tm * (tm + 1) / 2 : tm * tn);
  };
  if (const char* env = std::getenv("SKGEMM_OZ_PLAN")) {
    int cfg = 1, dp = 0, grid = 0;
    if (std::sscanf(env, "%d,%d,%d", &cfg, &dp, &grid) >= 2 && cfg >= 0 &&
        cfg < kNumGemmCfgs) {
      const int cap = gemm_occupancy(cfg) * c->num_cus;
      const int64_t U = units_for(cfg);
      int g = grid > 0 ? std::min(grid, cap) : cap;
      if (dp) g = (int)std::min<int64_t>(g, U);
      return OzPlan{cfg, dp ? 1 : 0, std::max(1, g)};
    }
  }
  if (c->forced_grid > 0)  // skgemm_set_tuning: Stream-K on that many WGs
    return OzPlan{1, 0, std::min(c->forced_grid,
                                 gemm_occupancy(1) * c->num_cus)};
  double best_score = 0.0;
  for (int cfg : {0, 1}) {
    const int cap = gemm_occupancy(cfg) * c->num_cus;
    const int64_t U = units_for(cfg);
    const int64_t rounds = (U + cap - 1) / cap;
    const double eff = (double)U / (double)(rounds * cap);
    const double score = eff * (cfg == 0 ? 1.0 : 0.92);
    if (eff >= 0.75 && score > best_score) {
      best_score = score;
      best = OzPlan{cfg, 1, (int)std::min<int64_t>(cap, U)};
    }
  }
  if (best.dp) return best;
  return OzPlan{1, 0, gemm_occupancy(1) * c->num_cus};
}

// A-posteriori bound on ||C - AB||_F / ||AB||_F from the integer rounding
// of step 1: with |A 2^sigma - A'| <= 1/2 elementwise,
// |C_ij - (AB)_ij| <= 2^-(sa_i+sb_j) (sqrt(K)/2 (|A'_i| + |B'_j|) + K/4),
// |A'_i| <= 2^sa_i ||A_i|| + sqrt(K)/2; the Frobenius norm of that bound
// matrix is a sum of separable terms.
double error_bound(const std::vector<double>& na, const std::vector<int>& sa,
                   const std::vector<double>& nb, const std::vector<int>& sb,
                   int K, double fro2) {
  const long double c = 0.5L * std::sqrt((long double)K), d = 0.25L * K;
  long double A0 = 0, A1 = 0, A2 = 0, B0 = 0, B1 = 0, B2 = 0;
  for (size_t i = 0; i < na.size(); ++i) {
    const long double al = std::ldexp(1.0L, -sa[i]);
    const long double a = std::ldexp(std::sqrt((long double)na[i]), sa[i]) + c;
    A0 += al * al;
    A1 += al * al * a;
    A2 += al * al * a * a;
  }
  for (size_t j = 0; j < nb.size(); ++j) {
    const long double be = std::ldexp(1.0L, -sb[j]);
    const long double b = std::ldexp(std::sqrt((long double)nb[j]), sb[j]) + c;
    B0 += be * be;
    B1 += be * be * b;
    B2 += be * be * b * b;
  }
  const long double s = c * c * (A2 * B0 + A0 * B2 + 2 * A1 * B1) +
                        2 * c * d * (A1 * B0 + A0 * B1) + d * d * A0 * B0;
  const long double num = std::sqrt(std::max(s, 0.0L));
  const long double den = std::sqrt((long double)fro2) - num;
  if (!(den > 0)) return INFINITY;
  return (double)(num / den) + 4.5e-16;  // + final roundings
}

// Shared driver.  kind 0: C = A B; kind 1: C = Psi^T Psi (B == A).
int ozaki_run(skgemm_ctx* c, int kind, int la, int lb, const Problem& pr,
              int slices, double* err_bound) {
  const Operand& A = pr.a;
  const Operand& B = pr.b;
  const int M = pr.M, N = pr.N, K = pr.K;
  int L = slices == 0 ? 14 : slices;
  if (L < 8 || L > kMaxMod)
    return set_error(SKGEMM_ERR_ARG, "slices (moduli) must be 0 or 8..%d",
                     kMaxMod);
  const int kpad = (int)round_up(std::max(K, 1), 128);
  if (kpad > kMaxKpad)
    return set_error(SKGEMM_ERR_NOTSUP,
                     "split form needs K <= %d for exact INT32 sums", kMaxKpad);
  if (hipSetDevice(c->device) != hipSuccess)
    return set_error(SKGEMM_ERR_HIP, "hipSetDevice");
  if (c->extra == nullptr) c->extra = new skgemm_ctx::Extra();
  const Tables T = make_tables(L);
  const OzPlan op = choose_plan(c, kind, M, N, L);
  const GemmCfg& gc = kGemmCfgs[op.cfg];
  const int opa = (int)round_up(M, 256), opb = (int)round_up(N, 256);
  const int64_t pa = (int64_t)opa * kpad, pb = (int64_t)opb * kpad;
  auto* X = c->extra;
  int8_t* planesA = (int8_t*)X->get(kPlanesA, (size_t)L * pa);
  int8_t* planesB = kind == 1 ? planesA : (int8_t*)X->get(kPlanesB, (size_t)L * pb);
  uint8_t* R = (uint8_t*)X->get(kRes, (size_t)L * M * N);
  double* n2a = (double*)X->get(kNormA, sizeof(double) * opa);
  double* n2b = kind == 1 ? n2a : (double*)X->get(kNormB, sizeof(double) * opb);
  int* sga = (int*)X->get(kSigA, sizeof(int) * opa);
  int* sgb = kind == 1 ? sga : (int*)X->get(kSigB, sizeof(int) * opb);
  // [0, 8): ||C||_F^2 accumulator (double); [8, 12): fallback flag.
  char* scal = (char*)X->get(kScalars, 16);
  if (!planesA || !planesB || !R || !n2a || !n2b || !sga || !sgb || !scal)
    return set_error(SKGEMM_ERR_NOMEM, "split-form workspace");
  double* fro2 = (double*)scal;
  int* fallback = (int*)(scal + 8);

  c->record_begin();
  // Cleared by a kernel, not a memset node, so the sequence stays valid
  // inside a captured graph.
  hipError_t e = launch_zero_ints(reinterpret_cast<int*>(scal), 3, c->stream);
  if (e != hipSuccess) return hip_error(e, "clear split-form scalars");
  int rc = prepare_operand(c, A, K, T, planesA, pa, kpad, opa, n2a, sga,
                           fallback, kPartA);
  if (rc != SKGEMM_OK) return rc;
  if (kind == 0) {
    rc = prepare_operand(c, B, K, T, planesB, pb, kpad, opb, n2b, sgb,
                         fallback, kPartB);
    if (rc != SKGEMM_OK) return rc;
  }

  GemmLaunch G{};
  G.A = planesA;
  G.B = planesB;
  G.a_plane = pa;
  G.b_plane = kind == 1 ? pa : pb;
  G.kpad = kpad;
  G.M = M;
  G.N = N;
  G.syrk = kind == 1;
  G.tiles_m = (M + gc.bm - 1) / gc.bm;
  G.tiles_n = (N + gc.bn - 1) / gc.bn;
  G.ntiles = kind == 1 ? (int64_t)G.tiles_m * (G.tiles_m + 1) / 2
                       : (int64_t)G.tiles_m * G.tiles_n;
  G.ipt = kpad / gc.bk;
  G.L = L;
  for (int l = 0; l < L; ++l) {
    G.mods[l] = T.m[l];
    G.inv_mods[l] = 1.0 / T.m[l];
  }
  G.R = R;
  G.r_plane = (int64_t)M * N;
  G.units = G.ntiles * L;
  G.total = G.units * G.ipt;
  G.dp = op.dp;
  G.grid = (int)std::max<int64_t>(1, std::min<int64_t>(op.grid, G.dp ? G.units
                                                                     : G.total));
  if (!G.dp) {
    rc = c->ensure_ws((size_t)G.grid * gc.bm * gc.bn / 2 + 1, G.grid);
    if (rc != SKGEMM_OK) return rc;
  }
  G.ws = reinterpret_cast<int*>(c->ws);
  G.flags = c->flags;
  G.status = c->status;
  G.spin_limit = c->spin_limit;
  e = launch_gemm(op.cfg, G, c->stream);
  if (e != hipSuccess) return hip_error(e, "ozaki gemm launch");

  CrtLaunch P{};
  P.R = R;
  P.r_plane = (int64_t)M * N;
  P.T = T;
  P.sig_a = sga;
  P.sig_b = sgb;
  P.fallback = fallback;
  P.d = pr.d;
  P.M = M;
  P.N = N;
  P.syrk = kind == 1;
  P.alpha = pr.alpha;
  P.beta = pr.beta;
  P.gamma = pr.gamma;
  P.fro2 = fro2;
  hipLaunchKernelGGL(crt_kernel, dim3((N + 63) / 64, (M + 3) / 4), dim3(256),
                     0, c->stream, P);
  if ((e = hipGetLastError()) != hipSuccess) return hip_error(e, "crt launch");

  // Plain FP64 fallback, decided on the device: the launch does its work
  // only when a row of A or column of B could not be scaled (fallback set).
  {
    Problem fp = pr;
    const Plan pl = plan(c, kind, la, lb, M, N, K);
    rc = fill_problem(c, &fp, kind, pl, M, N, K);
    if (rc != SKGEMM_OK) return rc;
    rc = run_single(c, fp, pl, fallback, false);
    if (rc != SKGEMM_OK) return rc;
  }
  c->record_end();

  if (err_bound != nullptr) {
    std::vector<double> na(M), nb(N);
    std::vector<int> sa(M), sb(N);
    double f[2];
    if ((e = hipMemcpyAsync(na.data(), n2a, sizeof(double) * M,
                            hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(sa.data(), sga, sizeof(int) * M,
                            hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(nb.data(), n2b, sizeof(double) * N,
                            hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(sb.data(), sgb, sizeof(int) * N,
                            hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(f, scal, 16, hipMemcpyDeviceToHost, c->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(c->stream)) != hipSuccess)
      return hip_error(e, "error-bound readback");
    rc = c->check_status();
    if (rc != SKGEMM_OK) return rc;
    int fb;
    std::memcpy(&fb, &f[1], sizeof(int));
    *err_bound = fb ? -1.0 : error_bound(na, sa, nb, sb, K, f[0]);
  }
  return SKGEMM_OK;
}

}  // namespace

extern "C" int skgemm_dgemm_ozaki(skgemm_ctx* c, int la, int lb, int ld_,
                                  int M, int N, int K, double alpha,
                                  const double* A, int64_t lda,
                                  const double* B, int64_t ldb, double beta,
                                  double* D, int64_t ldd, double gamma,
                                  int slices, double* err_bound) {
  int rc = check_ctx(c);
  if (rc != SKGEMM_OK) return rc;
  Problem p{};
  rc = prepare_gemm(la, lb, ld_, M, N, K, A, lda, B, ldb, D, ldd, &p);
  if (rc != SKGEMM_OK) return rc;
  if (err_bound) *err_bound = 0.0;
  if (M == 0 || N == 0) return SKGEMM_OK;
  p.M = M;
  p.N = N;
  p.K = K;
  p.alpha = alpha;
  p.beta = beta;
  p.gamma = gamma;
  return ozaki_run(c, 0, la, lb, p, slices, err_bound);
}

extern "C" int skgemm_dsyrk_ozaki(skgemm_ctx* c, int la, int N, int K,
                                  double alpha, const double* A, int64_t lda,
                                  double beta, double* C, int64_t ldc,
                                  int slices, double* err_bound) {
  int rc = check_ctx(c);
  if (rc != SKGEMM_OK) return rc;
  Problem p{};
  rc = prepare_syrk(la, N, K, A, lda, C, ldc, &p);
  if (rc != SKGEMM_OK) return rc;
  if (err_bound) *err_bound = 0.0;
  if (N == 0) return SKGEMM_OK;
  p.M = N;
  p.N = N;
  p.K = K;
  p.alpha = alpha;
  p.beta = beta;
  p.gamma = 0.0;
  return ozaki_run(c, 1, la, la, p, slices, err_bound);
}
