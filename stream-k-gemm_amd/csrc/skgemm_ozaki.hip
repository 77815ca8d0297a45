// Split-precision FP64 GEMM / SYRK on the INT8 matrix cores (Ozaki scheme
// II: Ozaki, Uchino & Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of
// floating-point matrix multiplication using an integer modular technique").
// No counterpart in the reference, whose path is FP64 only.
//
// C = A B (A: M x K, B: K x N; for SYRK A = Psi^T, B = Psi):
//  1. scale: each row of A / column of B gets s_o = 2^(sigma - 3) m,
//     m in 8..15, the largest with ||x_o s_o|| < 2^(h-1) from its squared
//     norm (scale_code), so that A'_o = rint(A_o s_o) and
//     |(A'B')_ij| < 2^(2h-2) (1 + 2^-20) < P/4, P = product of the moduli
//     (Cauchy-Schwarz); the same pass sums the squared rounding errors
//     (fma(a, s, -rint(a s)), exact before one rounding);
//  2. split: for L pairwise-coprime moduli m_l <= 256, the symmetric
//     residues of A' and B' (in [-128, 127]) as INT8 planes, K contiguous;
//  3. L INT8 products on v_mfma_i32_32x32x32_i8 with exact INT32
//     accumulation (|sum| <= Kpad * 127^2 < 2^31 for the odd moduli; for
//     m = 256 a wrap modulo 2^32 keeps the residue), reduced mod m_l, in one
//     launch over (modulus, K chunk, tile) units, data-parallel or
//     Stream-K; K beyond 131072 is cut into chunks whose residues are added
//     modulo m_l;
//  4. CRT: x = sum_l r_l W_l - q P in exact 128-bit arithmetic (q from an
//     FP64 estimate), one rounding to FP64, division by the two scales'
//     mantissas and an exact power-of-two unscaling, epilogue
//     alpha*C + beta*D + gamma.
// Integer steps are exact, so a result does not depend on the partition of
// the work (bit-identical for any grid).  The only approximation is step 1's
// rounding to integers; its rigorous a-posteriori bound (from the norms and
// the rounding-error norms) is returned with the result.  The certifying
// mode (SKGEMM_SLICES_AUTO) starts at 14 moduli and adds moduli until that
// bound is at most 1e-13, else computes in plain FP64.

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

// 16 moduli keep P < 2^128, so the reconstruction is exact in 128-bit
// integers (|x| <= P/4 < 2^126).
constexpr int kMaxMod = 16;
// Pairwise coprime: 256 = 2^8, 255 = 3*5*17, 253 = 11*23, 247 = 13*19,
// 217 = 7*31, the others prime.
static const int kModuli[kMaxMod] = {256, 255, 253, 251, 247, 241, 239, 233,
                                     229, 227, 223, 217, 211, 199, 197, 193};
// Products one INT32 sum may hold: kMaxKpad * 128^2 <= 2^31.  Longer K is
// cut into chunks of at most this many (their residues are added modulo m).
constexpr int kMaxKpad = 131072;

struct Tables {
  int L;
  int h;
  int m[kMaxMod];
  double md[kMaxMod];     // m_l
  double inv_m[kMaxMod];  // fl(1 / m_l)
  double c26[kMaxMod];    // 2^26 mod m_l, symmetric
  double c52[kMaxMod];    // 2^52 mod m_l, symmetric
  // CRT: x = sum_l r_l W_l - q P (mod 2^128),
  // W_l = (P / m_l) ((P / m_l)^{-1} mod m_l), q = floor(sum_l r_l w_l + 1/2)
  // with w_l = W_l / P rounded to FP64.  W_l and P in limbs (exact, least
  // significant first): three of 37 bits for L <= kLimb37Max, else four of
  // 32 bits.
  double W[kMaxMod][4];
  double wfrac[kMaxMod];
  double P[4];
};
constexpr int kLimb37Max = 14;  // product of the first 14 moduli < 2^111

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

// Scale of a row from its squared norm: s = 2^(sigma - 3) m, m in 8..15,
// the largest with norm2 s^2 <= 2^(2h-2) (1 - 2^-19), i.e. ||x s|| stays
// below 2^(h-1) with room for the rounding of the norm, stored as the code
// 8 sigma + (m - 8).  With norm2 = g 2^e (g in [0.5, 1)): e even gives
// sigma = h - 1 - e/2 and the test fl(g m^2) <= 64 (1 - 2^-19), e odd
// sigma = h - (e + 3)/2 and fl(g m^2) <= 128 (1 - 2^-19); m = 8 is the
// power-of-two scale 2^sigma, and the eight steps put every row within 1/8
// of an octave of the bound.  Exact operations and one rounded product,
// so host and device codes agree for the same norm2.  A zero row takes
// code 0 (s = 1).
__host__ __device__ inline int scale_code(double norm2, int h) {
  if (!(norm2 > 0.0)) return 0;
  int e;
  const double g = frexp(norm2, &e);
  const bool odd = (e & 1) != 0;
  int sig = odd ? h - ((e + 3) >> 1) : h - 1 - (e >> 1);
  double lim = odd ? 128.0 * (1.0 - 0x1p-19) : 64.0 * (1.0 - 0x1p-19);
  if (!odd && !(g * 64.0 <= lim)) {  // g within 2^-19 of 1: one step down
    --sig;
    lim = 256.0 * (1.0 - 0x1p-19);
  }
  int m = 15;
  while (m > 8 && !(g * (double)(m * m) <= lim)) --m;
  return 8 * sig + (m - 8);
}
__host__ __device__ inline int code_exp(int code) { return code >> 3; }
__host__ __device__ inline int code_mant(int code) { return 8 + (code & 7); }
// s = 2^(sigma - 3) m, exact.
__host__ __device__ inline double scale_of(int code) {
  return ldexp((double)code_mant(code), code_exp(code) - 3);
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
  sigma[o] = scale_code(s, h);
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
// below 2^-23 while y / m lies at least 1/(2m) from a half-integer), so
// r = y - q m is the symmetric residue in [-(m-1)/2, (m-1)/2].  For m = 256
// the result is in [-128, 128]; +128 and -128 share the byte 0x80, which is
// the same residue.  Five FP64 operations per residue, all exact.
// With more than 14 moduli h exceeds 55 and |x| reaches 2^76: then
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
// to low k; a launch covers segments seg0 .. seg0 + nsegc - 1 (all of them,
// or one K chunk of the pipelined SYRK).
constexpr int kSplitSeg = 4096;

// Residues of row elements k0 .. k0 + 3 (zero past K or outside the
// operand), one dword per plane.
// x' = rint(fl(a s)) and the rounding error a s - x', formed by one FMA
// (the exact error, rounded once); *d2 accumulates its square.
__device__ __forceinline__ double scaled_int(double a, double sc, double* d2) {
  const double xr = rint(a * sc);
  const double d = fma(a, sc, -xr);
  *d2 = fma(d, d, *d2);
  return xr;
}

template <bool WIDE>
__device__ __forceinline__ void split4(const double* row, bool in_o, int K,
                                       double sc, const Tables& T, int8_t* dst,
                                       int64_t plane, int k0, double* d2) {
  Split p[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = k0 + j;
    const double a = (in_o && k < K) ? row[k] : 0.0;
    p[j] = split26<WIDE>(scaled_int(a, sc, d2));
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

// Sum over the workgroup of one value per thread (fixed order), for
// thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  return red[0];
}

// dpart[seg][o]: the squared rounding error of row o over segment seg.
template <bool WIDE>
__global__ __launch_bounds__(256) void split_k_kernel(
    Operand op, int K, const int* sigma, Tables T, int8_t* planes,
    int64_t plane, int kpad, int seg0, int nsegc, double* dpart, int opad) {
  __shared__ double red[256];
  const int o = blockIdx.x / nsegc;
  const int seg = seg0 + nsegc - 1 - (blockIdx.x - o * nsegc);
  const bool in_o = o < op.extent;
  const double sc = scale_of(in_o ? sigma[o] : 0);
  const double* row = in_o ? op.ptr + op_addr(op, o, 0) : op.ptr;
  int8_t* dst = planes + (int64_t)o * kpad;
  double d2 = 0.0;
  for (int step = 3; step >= 0; --step) {
    const int k0 = seg * kSplitSeg + step * 1024 + threadIdx.x * 4;
    if (k0 >= kpad) continue;
    split4<WIDE>(row, in_o, K, sc, T, dst, plane, k0, &d2);
  }
  const double tot = block_sum<256>(d2, red);
  if (threadIdx.x == 0) dpart[(int64_t)seg * opad + o] = tot;
}

// dd2[o] = sum of the parts of row o in part order.
__global__ __launch_bounds__(256) void delta_finish_kernel(
    int O, int nparts, const double* dpart, int opad, double* dd2) {
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o >= O) return;
  double t = 0.0;
  for (int q = 0; q < nparts; ++q) t += dpart[(int64_t)q * opad + o];
  dd2[o] = t;
}

// Operand contiguous along k, norm and residues in one kernel: one
// 1024-thread workgroup per free index o (padding rows included).  The
// norm uses norm_k_kernel's grouping and order exactly (thread t + 256 q
// runs that kernel's partial sum q of thread t; the four are added as
// (s0 + s1) + (s2 + s3), then the same tree), so norms, scales and
// residues are bit-identical to norm_k_kernel followed by split_k_kernel.
// The row is then split from high k to low k while it is recent in the
// caches.
template <bool WIDE>
__global__ __launch_bounds__(1024) void norm_split_k_kernel(
    Operand op, int K, int h, Tables T, int8_t* planes, int64_t plane,
    int kpad, double* norm2, int* sigma, int* fallback, double* dd2) {
  __shared__ double red[2][1024];
  __shared__ int s_sig;
  const int o = blockIdx.x;
  const int tid = threadIdx.x;
  const bool in_o = o < op.extent;  // uniform over the workgroup
  const double* row = in_o ? op.ptr + op_addr(op, o, 0) : op.ptr;
  int s = 0;
  if (in_o) {
    const int t = tid & 255, q = tid >> 8;
    double acc = 0.0, mx = 0.0;
    int k = t;
    // Four loads in flight per thread; the sum stays one ordered chain.
    for (; k + 3 * 1024 + 768 < K; k += 4 * 1024) {
      const double x0 = row[k + 256 * q], x1 = row[k + 1024 + 256 * q],
                   x2 = row[k + 2048 + 256 * q], x3 = row[k + 3072 + 256 * q];
      acc = fma(x0, x0, acc);
      acc = fma(x1, x1, acc);
      acc = fma(x2, x2, acc);
      acc = fma(x3, x3, acc);
      mx = fmax(mx, fmax(fmax(fabs(x0), fabs(x1)), fmax(fabs(x2), fabs(x3))));
    }
    for (; k + 768 < K; k += 1024) {
      const double x = row[k + 256 * q];
      acc = fma(x, x, acc);
      mx = fmax(mx, fabs(x));
    }
    if (q == 0)
      for (; k < K; k += 256) {
        const double x = row[k];
        acc = fma(x, x, acc);
        mx = fmax(mx, fabs(x));
      }
    red[0][tid] = acc;
    red[1][tid] = mx;
    __syncthreads();
    if (tid < 256) {
      red[0][tid] = (red[0][tid] + red[0][tid + 256]) +
                    (red[0][tid + 512] + red[0][tid + 768]);
      red[1][tid] = fmax(fmax(red[1][tid], red[1][tid + 256]),
                         fmax(red[1][tid + 512], red[1][tid + 768]));
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) {
        red[0][tid] += red[0][tid + w];
        red[1][tid] = fmax(red[1][tid], red[1][tid + w]);
      }
      __syncthreads();
    }
    if (tid == 0) {
      finish_norm(red[0][0], red[1][0], o, h, norm2, sigma, fallback);
      s_sig = scale_code(red[0][0], h);
    }
    __syncthreads();
    s = s_sig;
  }
  int8_t* dst = planes + (int64_t)o * kpad;
  const int nseg = (kpad + kSplitSeg - 1) / kSplitSeg;
  // The next segment's four values are loaded before this one is split.
  auto load4 = [&](int k0, double (&x)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      x[j] = (in_o && k0 + j < K) ? row[k0 + j] : 0.0;
  };
  double cur[4], nxt[4] = {0.0, 0.0, 0.0, 0.0};
  int k0 = (nseg - 1) * kSplitSeg + tid * 4;
  const double sc = scale_of(s);
  double d2 = 0.0;
  load4(k0, cur);
  for (int seg = nseg - 1; seg >= 0; --seg) {
    if (seg > 0) load4(k0 - kSplitSeg, nxt);
    if (k0 < kpad) {
      Split p[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        p[j] = split26<WIDE>(scaled_int(cur[j], sc, &d2));
      for (int l = 0; l < T.L; ++l) {
        const double m = T.md[l], inv = T.inv_m[l], c = T.c26[l],
                     c2 = T.c52[l];
        const uint32_t v = pack4(residue<WIDE>(p[0], m, inv, c, c2),
                                 residue<WIDE>(p[1], m, inv, c, c2),
                                 residue<WIDE>(p[2], m, inv, c, c2),
                                 residue<WIDE>(p[3], m, inv, c, c2));
        *reinterpret_cast<uint32_t*>(dst + l * plane + k0) = v;
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) cur[j] = nxt[j];
    k0 -= kSplitSeg;
  }
  const double tot = block_sum<1024>(d2, red[0]);
  if (tid == 0) dd2[o] = tot;
}

// Operand contiguous along o: workgroup of 64 o x 128 k, 512 threads.
// Thread (o, 16 consecutive k) loads along o (coalesced), then each plane's
// 64 x 128-byte block is transposed through LDS so that every 8 lanes store
// one whole 128-byte row segment.  LDS row pitch 144 bytes: the per-lane
// dword writes (lanes = 64 rows) and the 16-byte row reads are
// conflict-free.
constexpr int kTPitch = 144;

// dpart[kb][o]: the squared rounding error of row o over k block kb.
template <bool WIDE>
__global__ __launch_bounds__(512) void split_o_kernel(
    Operand op, int K, const int* sigma, Tables T, int8_t* planes,
    int64_t plane, int kpad, int opad, double* dpart) {
  __shared__ __align__(16) uint8_t tile[2][64 * kTPitch];
  __shared__ double red[8][64];
  const int kblocks = kpad / 128;
  const int ob = blockIdx.x / kblocks;
  const int kb = kblocks - 1 - (blockIdx.x - ob * kblocks);
  const int ol = threadIdx.x & 63, kq = threadIdx.x >> 6;
  const int o = ob * 64 + ol;
  const int k0 = kb * 128 + kq * 16;
  const bool in_o = o < op.extent;
  const double sc = scale_of(in_o ? sigma[o] : 0);
  Split p[16];
  double d2 = 0.0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int k = k0 + j;
    const double a = (in_o && k < K) ? op.ptr[op_addr(op, o, k)] : 0.0;
    p[j] = split26<WIDE>(scaled_int(a, sc, &d2));
  }
  red[kq][ol] = d2;
  __syncthreads();
  if (kq == 0) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) t += red[q][ol];
    dpart[(int64_t)kb * opad + o] = t;
  }
  // Store side: thread q is row q / 8, 16-byte chunk q % 8 of the
  // 64 x 128 block.
  const int r0 = threadIdx.x >> 3, c0 = threadIdx.x & 7;
  int8_t* dst0 = planes + (int64_t)(ob * 64 + r0) * kpad + kb * 128 + c0 * 16;
  for (int l = 0; l < T.L; ++l) {
    const double m = T.md[l], inv = T.inv_m[l], c = T.c26[l], c2 = T.c52[l];
    uint8_t* t = tile[l & 1];
    uint32_t* trow = reinterpret_cast<uint32_t*>(t + ol * kTPitch + kq * 16);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      trow[q] = pack4(residue<WIDE>(p[4 * q], m, inv, c, c2),
                      residue<WIDE>(p[4 * q + 1], m, inv, c, c2),
                      residue<WIDE>(p[4 * q + 2], m, inv, c, c2),
                      residue<WIDE>(p[4 * q + 3], m, inv, c, c2));
    __syncthreads();
    const uint4 v = *reinterpret_cast<const uint4*>(t + r0 * kTPitch + c0 * 16);
    *reinterpret_cast<uint4*>(dst0 + l * plane) = v;
    // tile[l & 1] is rewritten two planes later, after the next barrier.
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
  uint8_t* R;  // [L][N][mp] symmetric residues as signed bytes, column-major
  int64_t r_plane;
  int mp;      // column pitch of R: M rounded up to 16
  int grid;
  int dp;          // 1: each workgroup takes whole units (no fix-up)
  int kc, kcs;     // data-parallel: K chunks per tile, MAC steps per chunk
  int c0, nck;     // data-parallel: this launch's chunks c0 .. c0 + nck - 1
  int64_t units;   // L * kc * ntiles (data-parallel), L * ntiles (Stream-K)
  int64_t total;   // units * ipt (Stream-K mode)
  int* ws;
  int* flags;
  int* status;
  int64_t spin_limit;
};

// 16 bytes per lane from a buffer resource straight into LDS at
// lds_dst + 16 * lane (a non-template helper, so the host pass of the
// kernel templates never sees the target builtin).
__device__ __forceinline__ void lds_dma16(__amdgpu_buffer_rsrc_t r,
                                          int8_t* lds_dst, int voffset) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(
      r, (__attribute__((address_space(3))) void*)lds_dst, 16, voffset, 0, 0,
      0);
}

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
// chunk c ^ f(r).  f is chosen so that a ds_read_b128 of 32 rows at one
// chunk index and a row-wise ds_write_b128 are free of bank conflicts both
// when 8 lanes share a 32-bank cycle and when 16 lanes share a 64-bank
// cycle (checked exhaustively on the host for both models).
template <int BK>
__device__ __forceinline__ int lds_off(int r, int c) {
  if (BK == 128) return r * 128 + ((c ^ (r & 7) ^ ((r >> 3) & 1)) << 4);
  return r * 64 + ((c ^ (r & 3) ^ ((r >> 2) & 3)) << 4);
}

template <int BM, int BN, int WM, int WN, int BK, int PD>
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

  // PD == 0: operands go straight from memory into LDS (16-byte LDS DMA:
  // lane l of a wave writes LDS base + 16 l, measured on gfx950), one wave
  // instruction per 1 KB = 1024 / BK rows; the swizzle is applied on the
  // global side (lane l fetches logical chunk (l % CPR) ^ f(row)).
  constexpr int RPD = 1024 / BK;                       // rows per DMA
  constexpr int NDA = BM / RPD, NDB = BN / RPD;        // DMAs per stage
  constexpr int DPW = (NDA + NDB) / (WM * WN);         // DMAs per wave
  // PD 3, 4: the same with PD LDS stages (PD - 1 steps in flight).
  // PD 8..11: PD - 6 LDS stages (steps issued PD - 7 steps ahead), the
  // barrier that publishes step k + 1 placed before the last sub-step of
  // step k (fragments read across it).
  constexpr bool EB = PD >= 8 && PD <= 11;
  constexpr bool DMA = PD == 0 || PD == 3 || PD == 4 || EB;
  constexpr int NS = (PD == 3 || PD == 4) ? PD : EB ? PD - 6 : 2;
  static_assert(!DMA || (NDA + NDB) % (WM * WN) == 0, "DMA split");
  int dma_off[DMA ? DPW : 1];
  if (DMA) {
#pragma unroll
    for (int d = 0; d < DPW; ++d) {
      const int q = wave * DPW + d;
      const int r = (q < NDA ? q : q - NDA) * RPD + lane / CPR;
      const int pc = lane % CPR;
      const int c = BK == 128 ? (pc ^ (r & 7) ^ ((r >> 3) & 1))
                              : (pc ^ (r & 3) ^ ((r >> 2) & 3));
      dma_off[d] = r * G.kpad + c * 16;
    }
  }

  v16i acc[MB][NB];

  // MAC steps [kb, ke) of unit `unit` into acc.
  // MAC steps [kb, ke) of tile `tile` of plane l into acc.
  auto mac = [&](int l, int64_t tile, int kb, int ke, int& m0, int& n0) {
    int ti, tj;
    if (G.syrk) {
      tri_tile(tile, ti, tj);
    } else {
      ti = (int)(tile / G.tiles_n);
      tj = (int)(tile % G.tiles_n);
    }
    m0 = ti * BM;
    n0 = tj * BN;
    // SYRK diagonal tile: a wave whose sub-tile lies entirely above the
    // diagonal does no MFMA work (its residues are never read); the other
    // waves on its SIMD get the matrix core to themselves.
    const bool idle = G.syrk && ti == tj && wm * TM + TM <= wn * TN;
    const int8_t* baseA = G.A + l * G.a_plane + (int64_t)m0 * G.kpad;
    const int8_t* baseB = G.B + l * G.b_plane + (int64_t)n0 * G.kpad;
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[i][j][v] = 0;

    // Two register sets: while step kk is multiplied out of LDS[cur], step
    // kk + 1 waits in one set and step kk + 2 is in flight into the other.
    // Steps past the unit's end are clamped to its last step (in bounds;
    // never stored).
    auto load = [&](v4i(&xa)[CA], v4i(&xb)[CB], int kk) {
      kk = kk < ke ? kk : ke - 1;
      const __amdgpu_buffer_rsrc_t sa = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<int8_t*>(baseA + (int64_t)kk * BK), (short)0,
          kNumRecords, kRsrcWord3);
      const __amdgpu_buffer_rsrc_t sb = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<int8_t*>(baseB + (int64_t)kk * BK), (short)0,
          kNumRecords, kRsrcWord3);
#pragma unroll
      for (int p = 0; p < CA; ++p)
        xa[p] = __builtin_amdgcn_raw_buffer_load_b128(
            sa, rowA[p] * G.kpad + chA[p] * 16, 0, 0);
#pragma unroll
      for (int p = 0; p < CB; ++p)
        xb[p] = __builtin_amdgcn_raw_buffer_load_b128(
            sb, rowB[p] * G.kpad + chB[p] * 16, 0, 0);
    };
    auto store = [&](const v4i(&xa)[CA], const v4i(&xb)[CB], int buf) {
      int8_t* base = lds + buf * (BM + BN) * BK;
#pragma unroll
      for (int p = 0; p < CA; ++p)
        *reinterpret_cast<v4i*>(base + lds_off<BK>(rowA[p], chA[p])) = xa[p];
#pragma unroll
      for (int p = 0; p < CB; ++p)
        *reinterpret_cast<v4i*>(base + BM * BK + lds_off<BK>(rowB[p], chB[p])) =
            xb[p];
    };
    auto compute = [&](int buf) {
      const int8_t* la = lds + buf * (BM + BN) * BK;
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
    };
    int kk = kb;
    int cur = 0;
    if (PD == 0) {
      auto issue = [&](int buf, int step) {
        const __amdgpu_buffer_rsrc_t sa = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<int8_t*>(baseA + (int64_t)step * BK), (short)0,
            kNumRecords, kRsrcWord3);
        const __amdgpu_buffer_rsrc_t sb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<int8_t*>(baseB + (int64_t)step * BK), (short)0,
            kNumRecords, kRsrcWord3);
        int8_t* base = lds + buf * (BM + BN) * BK;
#pragma unroll
        for (int d = 0; d < DPW; ++d) {
          const int q = wave * DPW + d;  // wave-uniform
          if (q < NDA)
            lds_dma16(sa, base + q * 1024, dma_off[d]);
          else
            lds_dma16(sb, base + BM * BK + (q - NDA) * 1024, dma_off[d]);
        }
      };
      issue(0, kk);
      for (; kk < ke; ++kk) {
        // This wave's copies into LDS[cur] are done; after the barrier
        // everyone's are, and everyone has finished reading LDS[cur ^ 1].
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (kk + 1 < ke) issue(cur ^ 1, kk + 1);
        if (!idle) compute(cur);
        cur ^= 1;
      }
      __syncthreads();
      return;
    }
    if constexpr (EB) {
      // NS stages hold steps k .. k + NS - 1; step k + NS is issued into
      // step k's stage once every wave has read all of step k.  The
      // barrier that publishes step k + 1 (and frees step k's stage, whose
      // fragments are all in registers by then) comes before the last
      // sub-step of step k, so that sub-step's MFMAs keep the matrix core
      // busy across it while step k + 1's first fragments are read.
      // Issues past the unit's end are clamped to its last step (in
      // bounds, never read), so exactly (NS - 2) DPW copies are younger
      // than step k + 1's at its barrier.
      constexpr int S = BK / 32;  // MFMA sub-steps per step
      static_assert(S >= 2 && S % 2 == 0, "sub-steps");
      constexpr int kAhead = (NS - 2) * DPW;
      constexpr int kWait = (kAhead & 15) | ((kAhead >> 4) << 14) |
                            (7 << 4) | (0 << 8);  // vmcnt(kAhead) lgkmcnt(0)
      auto issue = [&](int buf, int step) {
        step = step < ke ? step : ke - 1;
        const __amdgpu_buffer_rsrc_t sa = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<int8_t*>(baseA + (int64_t)step * BK), (short)0,
            kNumRecords, kRsrcWord3);
        const __amdgpu_buffer_rsrc_t sb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<int8_t*>(baseB + (int64_t)step * BK), (short)0,
            kNumRecords, kRsrcWord3);
        int8_t* base = lds + buf * (BM + BN) * BK;
#pragma unroll
        for (int d = 0; d < DPW; ++d) {
          const int q = wave * DPW + d;  // wave-uniform
          if (q < NDA)
            lds_dma16(sa, base + q * 1024, dma_off[d]);
          else
            lds_dma16(sb, base + BM * BK + (q - NDA) * 1024, dma_off[d]);
        }
      };
      auto frag = [&](v4i(&a)[MB], v4i(&b)[NB], int buf, int s) {
        const int8_t* la = lds + buf * (BM + BN) * BK;
        const int8_t* lb = la + BM * BK;
        const int c = 2 * s + (lane >> 5);
#pragma unroll
        for (int i = 0; i < MB; ++i)
          a[i] = *reinterpret_cast<const v4i*>(
              la + lds_off<BK>(wm * TM + i * 32 + (lane & 31), c));
#pragma unroll
        for (int j = 0; j < NB; ++j)
          b[j] = *reinterpret_cast<const v4i*>(
              lb + lds_off<BK>(wn * TN + j * 32 + (lane & 31), c));
      };
      auto mma = [&](const v4i(&a)[MB], const v4i(&b)[NB]) {
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                a[i], b[j], acc[i][j], 0, 0, 0);
      };
      v4i fa0[MB], fb0[NB], fa1[MB], fb1[NB];
#pragma unroll
      for (int s0 = 0; s0 < NS; ++s0) issue(s0, kk + s0);
      // step kk landed: the NS - 1 younger steps' copies may still fly
      {
        constexpr int kFirst = (NS - 1) * DPW;
        constexpr int kW0 = (kFirst & 15) | ((kFirst >> 4) << 14) |
                            (7 << 4) | (15 << 8);
        __builtin_amdgcn_s_waitcnt(kW0);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
      }
      frag(fa0, fb0, 0, 0);
      int stage = 0;
      for (; kk < ke; ++kk) {
#pragma unroll
        for (int s = 0; s + 1 < S; ++s) {
          if (s % 2 == 0) {
            frag(fa1, fb1, stage, s + 1);
            if (!idle) mma(fa0, fb0);
          } else {
            frag(fa0, fb0, stage, s + 1);
            if (!idle) mma(fa1, fb1);
          }
        }
        // Step k + 1's copies and this wave's fragment reads are done;
        // after the barrier everyone's are.
        __builtin_amdgcn_s_waitcnt(kWait);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        const int next = stage + 1 == NS ? 0 : stage + 1;
        if (kk + 1 < ke) frag(fa0, fb0, next, 0);
        if (!idle) mma(fa1, fb1);  // sub-step S - 1 (S even)
        issue(stage, kk + NS);
        stage = next;
      }
      __builtin_amdgcn_s_waitcnt(0);  // clamped tail copies
      __syncthreads();
      return;
    }
    if constexpr (PD == 3 || PD == 4) {
      // NS stages: steps kk + 1 .. kk + NS - 1 are in flight while step kk
      // is multiplied out.  Steps past the unit's end are clamped to its
      // last step (in bounds, never read), so exactly (NS - 2) DPW copies
      // are younger than step kk's and vmcnt <= (NS - 2) DPW means step
      // kk's copies have landed.
      constexpr int kAhead = (NS > 2 ? NS - 2 : 0) * DPW;
      constexpr int kWait = (kAhead & 15) | ((kAhead >> 4) << 14) |
                            (7 << 4) | (15 << 8);  // vmcnt only
      auto issue = [&](int buf, int step) {
        step = step < ke ? step : ke - 1;
        const __amdgpu_buffer_rsrc_t sa = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<int8_t*>(baseA + (int64_t)step * BK), (short)0,
            kNumRecords, kRsrcWord3);
        const __amdgpu_buffer_rsrc_t sb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<int8_t*>(baseB + (int64_t)step * BK), (short)0,
            kNumRecords, kRsrcWord3);
        int8_t* base = lds + buf * (BM + BN) * BK;
#pragma unroll
        for (int d = 0; d < DPW; ++d) {
          const int q = wave * DPW + d;  // wave-uniform
          if (q < NDA)
            lds_dma16(sa, base + q * 1024, dma_off[d]);
          else
            lds_dma16(sb, base + BM * BK + (q - NDA) * 1024, dma_off[d]);
        }
      };
#pragma unroll
      for (int s0 = 0; s0 < NS - 1; ++s0) issue(s0, kk + s0);
      int stage = 0;
      for (; kk < ke; ++kk) {
        // A bare s_barrier: a workgroup fence would also wait for the
        // younger copies still in flight (they write LDS).  The empty asm
        // statements keep the compiler from moving LDS accesses across it.
        __builtin_amdgcn_s_waitcnt(kWait);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();  // all of `stage` landed; stage-1 free
        asm volatile("" ::: "memory");
        const int free_stage = stage == 0 ? NS - 1 : stage - 1;
        issue(free_stage, kk + NS - 1);
        if (!idle) compute(stage);
        stage = stage + 1 == NS ? 0 : stage + 1;
      }
      __builtin_amdgcn_s_waitcnt(0);  // clamped tail copies
      __syncthreads();
      return;
    }
    v4i ra0[CA], rb0[CB], ra1[CA], rb1[CB];
    if (PD == 1) {
      // One register set: step kk + 1 is in flight while step kk is
      // multiplied out (for tiles whose two sets would not fit).
      load(ra0, rb0, kk);
      store(ra0, rb0, 0);
      __syncthreads();
      for (;;) {
        load(ra0, rb0, kk + 1);
        __builtin_amdgcn_sched_barrier(0);
        if (!idle) compute(cur);
        if (kk + 1 >= ke) break;
        store(ra0, rb0, cur ^ 1);
        __syncthreads();
        cur ^= 1;
        ++kk;
      }
      __syncthreads();
      return;
    }
    load(ra0, rb0, kk);
    load(ra1, rb1, kk + 1);
    store(ra0, rb0, 0);
    __syncthreads();
    // sched_barrier: keep each prefetch issued ahead of the MFMAs it is
    // meant to overlap (the scheduler would otherwise sink it below them).
    for (;;) {
      load(ra0, rb0, kk + 2);
      __builtin_amdgcn_sched_barrier(0);
      if (!idle) compute(cur);
      if (kk + 1 >= ke) break;
      store(ra1, rb1, cur ^ 1);
      __syncthreads();
      cur ^= 1;
      ++kk;
      load(ra1, rb1, kk + 2);
      __builtin_amdgcn_sched_barrier(0);
      if (!idle) compute(cur);
      if (kk + 1 >= ke) break;
      store(ra0, rb0, cur ^ 1);
      __syncthreads();
      cur ^= 1;
      ++kk;
    }
    __syncthreads();  // LDS is refilled by the next unit's prologue
  };

  // acc mod m_l -> R[l], column-major with column pitch G.mp (a multiple
  // of 16).  The tile's residues are transposed through LDS (column-major,
  // pitch BM + 16: dword writes of 4 rows and 16-byte reads are free of bank
  // conflicts) and stored as 16-byte row runs of a column.  Padding rows and
  // the upper part of diagonal SYRK tiles are stored too; the CRT pass
  // never reads them.  The residue is the symmetric one, a - m rint(a / m),
  // stored as a signed byte: acc is exact in FP64 and, for odd m, a / m
  // lies at least 1 / (2m) from a half-integer while fl(a fl(1/m)) is
  // within 2^-21 of it, so rint gives rint(a / m) and r lies in
  // [-(m-1)/2, (m-1)/2]; for m = 256 any r in [-128, 128] is the byte of
  // the right class (128 and -128 share it).
  auto write_residues = [&](int l, int out_plane, int m0, int n0) {
    const int m = G.mods[l];
    const double md = (double)m, inv = G.inv_mods[l];
    constexpr int TP = BM + 16;
    // Accumulator element v of lane x is row 8(v/4) + 4(x/32) + v%4,
    // column x%32 of its 32x32 block (measured on gfx950).
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int col = wn * TN + j * 32 + (lane & 31);
#pragma unroll
        for (int v4 = 0; v4 < 4; ++v4) {
          uint32_t word = 0;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            // a - m rint(a fl(1/m)) as one exact FMA (|m q| < 2^40): no
            // quarter-rate 32-bit integer multiply, no correction.
            const double ad = (double)acc[i][j][v4 * 4 + u];
            const int r = (int)fma(-md, rint(ad * inv), ad);
            word |= ((uint32_t)r & 255u) << (8 * u);
          }
          const int row = wm * TM + i * 32 + v4 * 8 + (lane >> 5) * 4;
          *reinterpret_cast<uint32_t*>(lds + col * TP + row) = word;
        }
      }
    __syncthreads();
    uint8_t* rp = G.R + out_plane * G.r_plane;
    constexpr int CH = BM / 16;  // 16-byte runs per tile column
#pragma unroll
    for (int q0 = 0; q0 < BN * CH; q0 += NT) {
      const int q = q0 + tid;
      const int col = q / CH, rc = q % CH;
      const int gc = n0 + col, gr = m0 + rc * 16;
      if (gc < G.N && gr < G.mp) {
        const uint4 val =
            *reinterpret_cast<const uint4*>(lds + col * TP + rc * 16);
        *reinterpret_cast<uint4*>(rp + (int64_t)gc * G.mp + gr) = val;
      }
    }
    __syncthreads();  // LDS is refilled by the next unit's prologue
  };

  if (G.dp) {
    // Whole units (plane l, K chunk c, tile), tile fastest, round-robin: in
    // each round the workgroups of one XCD take consecutive units, i.e.
    // tiles of one plane and K chunk that share panels in that XCD's L2 and
    // advance through K together.  Chunk c writes its own residue plane
    // (c L + l); the CRT pass adds the chunks modulo m_l.
    const int64_t per_plane = (int64_t)G.nck * G.ntiles;
    for (int64_t u = w; u < G.units; u += Gn) {
      const int l = (int)(u / per_plane);
      const int64_t r = u - (int64_t)l * per_plane;
      const int cr = (int)(r / G.ntiles);  // chunk within this launch
      const int64_t tile = r - (int64_t)cr * G.ntiles;
      const int c = G.c0 + cr;
      const int kb = c * G.kcs;
      const int ke = min(G.ipt, kb + G.kcs);
      int m0, n0;
      mac(l, tile, kb, ke, m0, n0);
      write_residues(l, c * G.L + l, m0, n0);
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
    const int l = (int)(unit / G.ntiles);
    int m0, n0;
    mac(l, unit - (int64_t)l * G.ntiles, kb, ke, m0, n0);

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
        // One accumulator block's loads are all issued before any is
        // added (one round trip per block).
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j) {
            int x[16];
#pragma unroll
            for (int v = 0; v < 16; ++v)
              x[v] = __builtin_amdgcn_raw_buffer_load_b32(
                  slot, tid * 4, ((i * NB + j) * 16 + v) * NT * 4,
                  kAgentCoherent);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] += x[v];
          }
      }
    }
    if (write_out) write_residues(l, l, m0, n0);
    it = t0 + ke;
  }
}

// ---- 4. CRT reconstruction and epilogue ----------------------------------------

struct CrtLaunch {
  const uint8_t* R;  // [kc][L][N][mp]
  int64_t r_plane;
  int mp;
  int kc;            // K chunks whose residues are added modulo m_l
  Tables T;
  const int* sig_a;
  const int* sig_b;
  const int* fallback;  // set: the plain FP64 path writes this result
  Output d;
  int M, N, syrk;
  double alpha, beta, gamma;
  double* fro2;  // per workgroup: its part of ||C||_F^2 (SYRK: both halves)
};

// Chunk sums the CRT takes without a reduction modulo m_l.
constexpr int kCrtMaxChunks = 32;

// x = S - q P with S = sum_l r_l W_l, in NL limbs of B bits held in FP64
// accumulators: up to 14 moduli (P < 2^111) three 37-bit limbs, above that
// four 32-bit limbs, all of it exact FMA arithmetic (full-rate FP64
// instead of quarter-rate 64-bit integer multiplies).  q = floor(S / P +
// 1/2) comes from the FP64 estimate f, within 2^-36 of the true value,
// whose fractional part lies in [1/4, 3/4] because |x| <= P / 4.  Carries
// are propagated exactly with floor(b 2^-B).  r[l] is any integer of the
// residue class of x mod m_l with |r[l]| <= 128 kCrtMaxChunks (a residue
// byte, or the plain sum of K chunks' residue bytes: S changes by a
// multiple of P and q absorbs it); then every product stays below 2^50
// (2^45) and every limb sum below 2^53 (2^49), so the arithmetic is exact.
// Returns x = hi 2^64 + lo with lo unsigned, for the four rows of a group
// at once, modulus by modulus (each modulus's
// table entries are read once and used for all four rows, so they need not
// stay live in scalar registers).  Residue l of row e: CH false, the signed
// byte e of rw[l]; CH true, the 16-bit lane of ev[l] (rows 0, 2) or od[l]
// (rows 1, 3) less `bias` (sums of biased bytes r + 128 over K chunks).
template <int NL, int B, bool CH>
__device__ __forceinline__ void crt4(const Tables& T, const uint32_t* rw,
                                     const uint32_t* ev, const uint32_t* od,
                                     int bias, int64_t* hi, uint64_t* lo) {
  double a[4][NL], f[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    f[e] = 0.5;
#pragma unroll
    for (int j = 0; j < NL; ++j) a[e][j] = 0.0;
  }
#pragma unroll
  for (int l = 0; l < kMaxMod; ++l) {
    if (l < T.L) {
      double w[NL];
#pragma unroll
      for (int j = 0; j < NL; ++j) w[j] = T.W[l][j];
      const double wf = T.wfrac[l];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        int ri;
        if constexpr (CH) {
          const uint32_t v = (e & 1) ? od[l] : ev[l];
          ri = (int)((e & 2) ? (v >> 16) : (v & 0xFFFFu)) - bias;
        } else {
          ri = (int)(int8_t)(uint8_t)(rw[l] >> (8 * e));
        }
        const double rl = (double)ri;
#pragma unroll
        for (int j = 0; j < NL; ++j) a[e][j] = fma(rl, w[j], a[e][j]);
        f[e] = fma(rl, wf, f[e]);
      }
    }
  }
  const double two = (double)(1ull << B), inv = 1.0 / (double)(1ull << B);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const double q = floor(f[e]);
    double b[NL];
#pragma unroll
    for (int j = 0; j < NL; ++j) b[j] = fma(-q, T.P[j], a[e][j]);
#pragma unroll
    for (int j = 0; j + 1 < NL; ++j) {
      const double c = floor(b[j] * inv);
      b[j] = fma(-c, two, b[j]);
      b[j + 1] += c;
    }
    if constexpr (NL == 4) {
      lo[e] = (uint64_t)(uint32_t)b[0] | ((uint64_t)(uint32_t)b[1] << 32);
      hi[e] = (int64_t)(((uint64_t)(int64_t)b[3] << 32) |
                        (uint64_t)(uint32_t)b[2]);
    } else {
      // x = b2 2^74 + b1 2^37 + b0 with b0, b1 in [0, 2^37).
      const uint64_t l0 = (uint64_t)b[0], l1 = (uint64_t)b[1];
      lo[e] = l0 | (l1 << 37);
      hi[e] = (int64_t)b[2] * 1024 + (int64_t)(l1 >> 27);
    }
  }
}

// Blocks of 256 rows x 4 columns; a thread takes 4 consecutive rows of one
// column (one dword of every residue plane: all loads issued together).
// Grid-stride; SYRK skips blocks strictly above the diagonal.  Each
// workgroup writes its part of ||C||_F^2 to fro2[blockIdx.x] (summed on the
// host in a fixed order).  CHUNKS: the residue planes of p.kc > 1 K chunks
// are added first.
template <bool CHUNKS>
__global__ __launch_bounds__(256) void crt_kernel(CrtLaunch p) {
  __shared__ double red[256];
  if (*p.fallback != 0) return;
  const int nbr = (p.M + 255) / 256, nbc = (p.N + 3) / 4;
  const int64_t nblocks = (int64_t)nbc * nbr;
  double contrib = 0.0;
  for (int64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const int bc = (int)(b / nbr), br = (int)(b - (int64_t)bc * nbr);
    // First column whose entries this SYRK keeps at or below a row: the
    // column itself, or the start of its 16-block when whole diagonal
    // blocks are stored (syrk == 2).
    const int cfirst = p.syrk == 2 ? ((bc * 4) & ~15) : bc * 4;
    if (p.syrk && br * 256 + 255 < cfirst) continue;
    const int row0 = br * 256 + (threadIdx.x & 63) * 4;
    const int col = bc * 4 + (threadIdx.x >> 6);
    const int cmin = p.syrk == 2 ? (col & ~15) : col;
    if (row0 >= p.M || col >= p.N || (p.syrk && row0 + 3 < cmin)) continue;
    const uint8_t* src = p.R + (int64_t)col * p.mp + row0;
    uint32_t rw[kMaxMod];
#pragma unroll
    for (int l = 0; l < kMaxMod; ++l)
      rw[l] = l < p.T.L
                  ? *reinterpret_cast<const uint32_t*>(src + l * p.r_plane)
                  : 0u;
    // With K chunks: each chunk's residue is a signed byte; biased to
    // r + 128 and split into the 16-bit lanes of an even-row and an odd-row
    // word, kc <= 257 sums stay below 2^16 (the plain sum is another member
    // of the class, which the CRT takes as it is).
    uint32_t ev[CHUNKS ? kMaxMod : 1], od[CHUNKS ? kMaxMod : 1];
    if constexpr (CHUNKS) {
#pragma unroll
      for (int l = 0; l < kMaxMod; ++l) {
        const uint32_t v = rw[l] ^ 0x80808080u;
        ev[l] = v & 0x00FF00FFu;
        od[l] = (v >> 8) & 0x00FF00FFu;
      }
      for (int c = 1; c < p.kc; ++c) {
#pragma unroll
        for (int l = 0; l < kMaxMod; ++l) {
          if (l < p.T.L) {
            const uint32_t v =
                *reinterpret_cast<const uint32_t*>(
                    src + ((int64_t)c * p.T.L + l) * p.r_plane) ^
                0x80808080u;
            ev[l] += v & 0x00FF00FFu;
            od[l] += (v >> 8) & 0x00FF00FFu;
          }
        }
      }
    }
    const int bias = 128 * p.kc;
    const int cb = p.sig_b[col];
    int64_t xh4[4];
    uint64_t xl4[4];
    if (p.T.L <= kLimb37Max)
      crt4<3, 37, CHUNKS>(p.T, rw, ev, od, bias, xh4, xl4);
    else
      crt4<4, 32, CHUNKS>(p.T, rw, ev, od, bias, xh4, xl4);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = row0 + e;
      if (row < p.M && syrk_keep(p.syrk, row, col)) {
        const double d = (double)xh4[e] * 18446744073709551616.0 + (double)xl4[e];
        // Unscale by s_a s_b = 2^(ea + eb - 6) ma mb: one rounded division
        // by the integer ma mb (<= 225), then an exact power of two.
        const int ca = p.sig_a[row];
        const double c =
            ldexp(d / (double)(code_mant(ca) * code_mant(cb)),
                  -(code_exp(ca) + code_exp(cb) - 6));
        // ||C||_F^2 from the lower triangle (off-diagonal entries twice).
        if (!p.syrk || row >= col)
          contrib += (p.syrk && row != col) ? 2.0 * c * c : c * c;
        double* dst = p.d.ptr + out_addr(p.d, row, col);
        double val = p.alpha * c;
        if (p.beta != 0.0) val = fma(p.beta, *dst, val);
        *dst = val + p.gamma;
      }
    }
  }
  red[threadIdx.x] = contrib;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) p.fro2[blockIdx.x] = red[0];
}

constexpr int kCrtGrid = 4096;

// Inputs of the error bound reduced on the device, so that a call reads
// back eight doubles instead of every row's norms: out = {sum n^2, sum e^2,
// sum n e} of A's rows, the same of B's columns, ||C||_F^2 (the CRT
// workgroups' parts in order), and the fallback flag.  n = sqrt(norm2 g),
// e = sqrt(delta2 g) / s with g = 1 + (K + 2) 2^-52 and a factor 1 + 2^-51
// per rounded operation (the bound's rigour does not rest on sqrt or the
// division being correctly rounded); the host applies the summation factor.
// One workgroup, fixed order.
__global__ __launch_bounds__(1024) void bound_sums_kernel(
    const double* n2a, const int* ca, const double* dda, int M,
    const double* n2b, const int* cb, const double* ddb, int N, int K,
    const double* fro2, int nparts, const int* fallback, const int* status,
    double* out) {
  __shared__ double red[1024];
  const double u2 = 0x1p-51, g = 1.0 + (K + 2.0) * 0x1p-52;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < M; i += 1024) {
    const double n = sqrt(n2a[i] * g) * (1.0 + u2);
    const double e = sqrt(dda[i] * g) * (1.0 + u2) / scale_of(ca[i]) *
                     (1.0 + u2);
    v[0] += n * n;
    v[1] += e * e;
    v[2] += n * e;
  }
  for (int j = threadIdx.x; j < N; j += 1024) {
    const double n = sqrt(n2b[j] * g) * (1.0 + u2);
    const double e = sqrt(ddb[j] * g) * (1.0 + u2) / scale_of(cb[j]) *
                     (1.0 + u2);
    v[3] += n * n;
    v[4] += e * e;
    v[5] += n * e;
  }
  for (int q = threadIdx.x; q < nparts; q += 1024) v[6] += fro2[q];
  for (int t = 0; t < 7; ++t) {
    red[threadIdx.x] = v[t];
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[t] = red[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[7] = (double)*fallback;
    out[8] = (double)*status;  // the context's fix-up status, same readback
  }
}

// ---- host ----------------------------------------------------------------------

struct GemmCfg {
  int bm, bn, wm, wn, bk, pd;
};
// 0: 256x256, 8 waves of 64x128, BK 64,  2-deep prefetch (64 KB LDS)
// 1: 128x128, 4 waves of 64x64,  BK 128, 2-deep (64 KB LDS)
// 2: 128x128, 2 waves of 64x128, BK 128, 2-deep (64 KB LDS)
// 3: 128x128, 4 waves of 64x64,  BK 64,  2-deep (32 KB LDS)
// 4: 256x256, 8 waves of 64x128, BK 128, 1-deep (128 KB LDS)
// 5: 256x256, 8 waves of 64x128, BK 128, LDS DMA, 2 stages (128 KB LDS)
// 6: 128x128, 4 waves of 64x64,  BK 128, LDS DMA, 2 stages (64 KB LDS)
// 7: 128x128, 4 waves of 64x64,  BK 64,  LDS DMA, 4 stages (64 KB LDS)
// 8: 128x128, 4 waves of 64x64,  BK 128, LDS DMA, 3 stages (96 KB LDS)
// 9: 256x256, 8 waves of 64x128, BK 64,  LDS DMA, 3 stages (96 KB LDS)
// Early barrier (the barrier publishing step k + 1 sits before the last
// sub-step of step k; measured beside the rest, selectable plans):
// 10: 256x256, 8 waves of 64x128, BK 128, LDS DMA, 2 stages (128 KB LDS)
// 11: 128x128, 4 waves of 64x64,  BK 128, LDS DMA, 2 stages (64 KB LDS)
// 12: 256x256, 8 waves of 64x128, BK 64,  LDS DMA, 5 stages (160 KB LDS)
static const GemmCfg kGemmCfgs[] = {{256, 256, 4, 2, 64, 2},
                                    {128, 128, 2, 2, 128, 2},
                                    {128, 128, 2, 1, 128, 2},
                                    {128, 128, 2, 2, 64, 2},
                                    {256, 256, 4, 2, 128, 1},
                                    {256, 256, 4, 2, 128, 0},
                                    {128, 128, 2, 2, 128, 0},
                                    {128, 128, 2, 2, 64, 4},
                                    {128, 128, 2, 2, 128, 3},
                                    {256, 256, 4, 2, 64, 3},
                                    {256, 256, 4, 2, 128, 8},
                                    {128, 128, 2, 2, 128, 8},
                                    {256, 256, 4, 2, 64, 11}};
constexpr int kNumGemmCfgs = 13;

#define SKG_OZ_CFGS(X)          \
  X(0, 256, 256, 4, 2, 64, 2)   \
  X(1, 128, 128, 2, 2, 128, 2)  \
  X(2, 128, 128, 2, 1, 128, 2)  \
  X(3, 128, 128, 2, 2, 64, 2)   \
  X(4, 256, 256, 4, 2, 128, 1)  \
  X(5, 256, 256, 4, 2, 128, 0)  \
  X(6, 128, 128, 2, 2, 128, 0)  \
  X(7, 128, 128, 2, 2, 64, 4)   \
  X(8, 128, 128, 2, 2, 128, 3)  \
  X(9, 256, 256, 4, 2, 64, 3)   \
  X(10, 256, 256, 4, 2, 128, 8) \
  X(11, 128, 128, 2, 2, 128, 8) \
  X(12, 256, 256, 4, 2, 64, 11)

static const void* gemm_fn(int cfg) {
#define SKG_FN_CASE(i, bm, bn, wm, wn, bk, pd) \
  case i:                                      \
    return reinterpret_cast<const void*>(      \
        gemm_i8_kernel<bm, bn, wm, wn, bk, pd>);
  switch (cfg) {
    SKG_OZ_CFGS(SKG_FN_CASE)
    default: return nullptr;
  }
#undef SKG_FN_CASE
}

static size_t gemm_lds(int cfg) {
  const GemmCfg& g = kGemmCfgs[cfg];
  const size_t stages = (g.pd == 3 || g.pd == 4)    ? g.pd
                        : (g.pd >= 8 && g.pd <= 11) ? g.pd - 6
                                                    : 2;
  const size_t pipe = stages * (g.bm + g.bn) * g.bk;
  const size_t tile = (size_t)g.bn * (g.bm + 16);  // residue transpose
  return std::max(pipe, tile);
}

static hipError_t gemm_attr(int cfg) {
  static bool done[kNumGemmCfgs] = {};
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
#define SKG_LAUNCH_CASE(i, bm, bn, wm, wn, bk, pd)                        \
  case i:                                                                 \
    hipLaunchKernelGGL((gemm_i8_kernel<bm, bn, wm, wn, bk, pd>), grid,    \
                       block, lds, s, G);                                 \
    break;
  switch (cfg) {
    SKG_OZ_CFGS(SKG_LAUNCH_CASE)
    default: return hipErrorInvalidValue;
  }
#undef SKG_LAUNCH_CASE
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
  // Scaled rows have ||x'|| < 2^(h-1) (plus the rounding, sqrt(K)/2), so
  // |C'_ij| < 2^(2h-2) (1 + 2^-20) by Cauchy-Schwarz, which must stay below
  // P / 4 = 2^(bits-2) for the CRT: h <= bits / 2 less a margin of 0.05
  // bits (2^0.1 > 1 + 2^-20).  L = 14: bits = 110.16, h = 55.
  T.h = (int)std::floor(bits / 2.0 - 0.05);
  typedef unsigned __int128 u128;
  u128 P = 1;
  for (int l = 0; l < L; ++l) P *= (u128)kModuli[l];
  const int nl = L <= kLimb37Max ? 3 : 4, lb = L <= kLimb37Max ? 37 : 32;
  const u128 lmask = ((u128)1 << lb) - 1;
  for (int j = 0; j < nl; ++j) T.P[j] = (double)(uint64_t)((P >> (lb * j)) & lmask);
  for (int l = 0; l < L; ++l) {
    const u128 Pl = P / (u128)kModuli[l];
    const int inv = modinv((int)(Pl % (u128)kModuli[l]), kModuli[l]);
    const u128 W = Pl * (u128)inv;  // < P
    for (int j = 0; j < nl; ++j)
      T.W[l][j] = (double)(uint64_t)((W >> (lb * j)) & lmask);
    // W / P = inv / m_l exactly; rounded once to FP64.
    T.wfrac[l] = (double)inv / (double)kModuli[l];
  }
  return T;
}

}  // namespace oz
}  // namespace skg

// ---- context extras ------------------------------------------------------------

struct skgemm_ctx::Extra {
  static constexpr int kBufs = 17;
  void* buf[kBufs] = {nullptr};
  size_t n[kBufs] = {0};
  uint64_t* gen = nullptr;  // the context's mem_gen
  // Pipelined split SYRK: a second stream for the INT8 products and the
  // events that order them after each K chunk's residue pass.
  static constexpr int kMaxPipe = 8;
  hipStream_t aux = nullptr;
  hipEvent_t ev[kMaxPipe + 1] = {nullptr};
  bool pipe_ready() {
    if (aux) return true;
    if (hipStreamCreateWithFlags(&aux, hipStreamNonBlocking) != hipSuccess) {
      aux = nullptr;
      return false;
    }
    for (auto& e : ev)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        e = nullptr;
        return false;
      }
    return true;
  }
  // Pinned host buffer for the error bound's eight sums.
  double* pinned = nullptr;
  double* pinned_sums() {
    if (pinned == nullptr &&
        hipHostMalloc((void**)&pinned, 16 * sizeof(double), 0) != hipSuccess)
      pinned = nullptr;
    return pinned;
  }
  // Marks the bound readback in the stream; polled instead of a blocking
  // stream synchronisation (no wake-up latency between calls).
  hipEvent_t read_ev = nullptr;
  hipEvent_t readback_event() {
    if (read_ev == nullptr &&
        hipEventCreateWithFlags(&read_ev, hipEventDisableTiming) != hipSuccess)
      read_ev = nullptr;
    return read_ev;
  }
  void* get(int i, size_t bytes) {
    if (bytes > n[i]) {
      if (gen) ++*gen;
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
  if (extra->aux) hipStreamSynchronize(extra->aux);
  for (auto& e : extra->ev)
    if (e) hipEventDestroy(e);
  if (extra->read_ev) hipEventDestroy(extra->read_ev);
  if (extra->aux) hipStreamDestroy(extra->aux);
  for (int i = 0; i < Extra::kBufs; ++i)
    if (extra->buf[i]) hipFree(extra->buf[i]);
  if (extra->pinned) hipHostFree(extra->pinned);
  delete extra;
  extra = nullptr;
}

using namespace skg;
using namespace skg::oz;

namespace {

enum {
  kPlanesA = 0, kPlanesB, kRes, kNormA, kNormB, kSigA, kSigB, kScalars,
  kSavedD, kDeltaA, kDeltaB, kDpartA, kDpartB, kBoundSums,
  kPartA, kPartB
};

int64_t round_up(int64_t x, int64_t q) { return (x + q - 1) / q * q; }

// K segments of the o-contiguous norm pass: enough (o block, k segment)
// workgroups to fill the chip.
int norm_segments(int O, int K) {
  const int oblocks = (O + 255) / 256;
  return std::max(1, std::min({(2048 + oblocks - 1) / oblocks, (K + 15) / 16,
                               512}));
}

// Scale + split one operand (free index o, contraction k) into planes.
// norms_ready: norm2, sigma and the fallback flag are already on the device
// (the certifying mode computed them first); only the residues are formed.
int prepare_operand(skgemm_ctx* c, const Operand& op, int K, const Tables& T,
                    int8_t* planes, int64_t plane, int kpad, int opad,
                    double* norm2, int* sigma, int* fallback, int part_buf,
                    bool norms_ready, double* dd2, int dpart_buf) {
  const int O = op.extent;
  const bool wide = T.h > 55;  // |x'| < 2^54: two 26-bit limbs suffice
  const char* fuse = std::getenv("SKGEMM_OZ_FUSE");
  if (op.kcontig && !norms_ready && !(fuse != nullptr && fuse[0] == '0')) {
    // One pass: norm, scale, residues and rounding-error norm per row.
    if (wide)
      hipLaunchKernelGGL(norm_split_k_kernel<true>, dim3(opad), dim3(1024), 0,
                         c->stream, op, K, T.h, T, planes, plane, kpad, norm2,
                         sigma, fallback, dd2);
    else
      hipLaunchKernelGGL(norm_split_k_kernel<false>, dim3(opad), dim3(1024),
                         0, c->stream, op, K, T.h, T, planes, plane, kpad,
                         norm2, sigma, fallback, dd2);
  } else if (op.kcontig) {
    if (O > 0 && !norms_ready)
      hipLaunchKernelGGL(norm_k_kernel, dim3(O), dim3(256), 0, c->stream, op,
                         K, T.h, norm2, sigma, fallback);
    const int nseg = (kpad + kSplitSeg - 1) / kSplitSeg;
    double* dpart =
        (double*)c->extra->get(dpart_buf, sizeof(double) * nseg * opad);
    if (dpart == nullptr) return set_error(SKGEMM_ERR_NOMEM, "error norms");
    const dim3 grid((unsigned)((int64_t)opad * nseg));
    if (wide)
      hipLaunchKernelGGL(split_k_kernel<true>, grid, dim3(256), 0, c->stream,
                         op, K, sigma, T, planes, plane, kpad, 0, nseg, dpart,
                         opad);
    else
      hipLaunchKernelGGL(split_k_kernel<false>, grid, dim3(256), 0, c->stream,
                         op, K, sigma, T, planes, plane, kpad, 0, nseg, dpart,
                         opad);
    hipLaunchKernelGGL(delta_finish_kernel, dim3((opad + 255) / 256),
                       dim3(256), 0, c->stream, opad, nseg, dpart, opad, dd2);
  } else {
    if (O > 0 && !norms_ready) {
      const int nseg = norm_segments(O, K);
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
    const int kblocks = kpad / 128;
    double* dpart =
        (double*)c->extra->get(dpart_buf, sizeof(double) * kblocks * opad);
    if (dpart == nullptr) return set_error(SKGEMM_ERR_NOMEM, "error norms");
    const dim3 grid((unsigned)((int64_t)(opad / 64) * kblocks));
    if (wide)
      hipLaunchKernelGGL(split_o_kernel<true>, grid, dim3(512), 0, c->stream,
                         op, K, sigma, T, planes, plane, kpad, opad, dpart);
    else
      hipLaunchKernelGGL(split_o_kernel<false>, grid, dim3(512), 0, c->stream,
                         op, K, sigma, T, planes, plane, kpad, opad, dpart);
    hipLaunchKernelGGL(delta_finish_kernel, dim3((opad + 255) / 256),
                       dim3(256), 0, c->stream, opad, kblocks, dpart, opad,
                       dd2);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(e, "ozaki split");
  return SKGEMM_OK;
}

struct OzPlan {
  int cfg, dp, grid, kc;
};

// Tile configuration and schedule of the INT8 product.  Data-parallel whole
// units (plane, K chunk, tile) when they fill the chip in near-even rounds
// (L2 reuse of panels between the tiles of one plane on one XCD, no
// inter-workgroup waits); Stream-K otherwise.  SKGEMM_OZ_PLAN=
// "cfg,dp[,grid[,kc]]" overrides (tuning sweeps); a grid forced with
// skgemm_set_tuning selects Stream-K on that grid, except on contexts that
// must not wait (oz_no_wait), where it caps the data-parallel grid.
// K chunks a data-parallel plan needs so that no unit sums more than
// kMaxKpad products (each INT32 sum stays exact); above kMaxKpad only
// data-parallel plans are possible (a Stream-K unit spans all of K).
static int min_chunks(int kpad, int bk) {
  const int ipt = kpad / bk, per = kMaxKpad / bk;
  return (ipt + per - 1) / per;
}

OzPlan choose_plan_any(skgemm_ctx* c, int kind, int M, int N, int L,
                       int kpad) {
  auto tiles_for = [&](int cfg) {
    const GemmCfg& g = kGemmCfgs[cfg];
    const int64_t tm = (M + g.bm - 1) / g.bm, tn = (N + g.bn - 1) / g.bn;
    return kind == 1 ? tm * (tm + 1) / 2 : tm * tn;
  };
  if (const char* env = std::getenv("SKGEMM_OZ_PLAN")) {
    int cfg = 1, dp = 0, grid = 0, kc = 1;
    if (std::sscanf(env, "%d,%d,%d,%d", &cfg, &dp, &grid, &kc) >= 2 &&
        cfg >= 0 && cfg < kNumGemmCfgs) {
      const int ipt = kpad / kGemmCfgs[cfg].bk;
      if (kpad > kMaxKpad) dp = 1;
      kc = dp ? std::max(min_chunks(kpad, kGemmCfgs[cfg].bk),
                         std::max(1, std::min(kc, ipt)))
              : 1;
      const int cap = gemm_occupancy(cfg) * c->num_cus;
      const int64_t U = (int64_t)L * tiles_for(cfg) * kc;
      int g = grid > 0 ? std::min(grid, cap) : cap;
      if (dp) g = (int)std::min<int64_t>(g, U);
      return OzPlan{cfg, dp ? 1 : 0, std::max(1, g), kc};
    }
  }
  if (c->forced_grid > 0 && !c->oz_no_wait && kpad <= kMaxKpad)
    return OzPlan{1, 0, std::min(c->forced_grid,   // Stream-K on that many
                                 gemm_occupancy(1) * c->num_cus), 1};
  OzPlan best{1, 0, 0, 1};
  double best_score = 0.0;
  for (int cfg : {5, 6}) {
    const int cap = gemm_occupancy(cfg) * c->num_cus;
    const int ipt = kpad / kGemmCfgs[cfg].bk;
    const int k0 = min_chunks(kpad, kGemmCfgs[cfg].bk);
    for (int kc = k0; kc <= std::max(16, k0) && (kc == k0 || ipt / kc >= 32);
         ++kc) {
      const int64_t U = (int64_t)L * tiles_for(cfg) * kc;
      const int64_t rounds = (U + cap - 1) / cap;
      const double eff = (double)U / (double)(rounds * cap);
      // Bigger tiles halve the L2 traffic per MFMA; each extra chunk costs
      // one more residue plane to write and read.
      const double score =
          eff * (kGemmCfgs[cfg].bm == 256 ? 1.0 : 0.97) *
          (1.0 - 0.01 * (kc - 1));
      if (eff >= 0.75 && score > best_score) {
        best_score = score;
        best = OzPlan{cfg, 1, (int)std::min<int64_t>(cap, U), kc};
      }
    }
  }
  if (!best.dp && (c->oz_no_wait || kpad > kMaxKpad)) {
    const int k0 = min_chunks(kpad, kGemmCfgs[6].bk);
    const int64_t U = (int64_t)L * tiles_for(6) * k0;
    best = OzPlan{6, 1, (int)std::min<int64_t>(gemm_occupancy(6) * c->num_cus, U),
                  k0};
  }
  if (best.dp) {
    if (c->forced_grid > 0) best.grid = std::min(best.grid, c->forced_grid);
    return best;
  }
  return OzPlan{1, 0, gemm_occupancy(1) * c->num_cus, 1};
}

OzPlan choose_plan(skgemm_ctx* c, int kind, int M, int N, int L, int kpad) {
  OzPlan op = choose_plan_any(c, kind, M, N, L, kpad);
  // Whatever chose it, no unit may sum more than kMaxKpad products, and
  // the CRT takes at most kCrtMaxChunks chunk residues.
  op.kc = std::min(op.kc, kCrtMaxChunks);
  if (kpad > kMaxKpad) {
    op.dp = 1;
    op.kc = std::max(op.kc, min_chunks(kpad, kGemmCfgs[op.cfg].bk));
  }
  return op;
}

// K the split form accepts: kCrtMaxChunks chunks of kMaxKpad products.
constexpr int64_t kMaxSplitK = (int64_t)kCrtMaxChunks * kMaxKpad;

// A-posteriori bound on ||C - AB||_F / ||AB||_F.  Row i of A is scaled by
// s_i and rounded, x'_i = s_i a_i - d_i, column j of B by t_j, y'_j =
// t_j b_j - g_j.  Then x'_i . y'_j / (s_i t_j) - a_i . b_j =
// -(a_i . g_j) / t_j - (d_i . b_j) / s_i + (d_i . g_j) / (s_i t_j), so with
// n_i >= ||a_i||, e_i >= ||d_i|| / s_i (and m_j, f_j for B)
//   |C_ij - (AB)_ij| <= n_i f_j + e_i m_j + e_i f_j,
// whose squared Frobenius norm expands into products of row sums and
// column sums.  ||d_i||^2 is the residue pass's sum of the rounding errors
// (each the exact error rounded once by an FMA), the norms the norm pass's
// sums of squares; both lie within (K + 2) u of their values, which the
// factor g covers.  Converting C' to FP64 and unscaling add two roundings.
struct RowSums {
  double nn = 0, ee = 0, ne = 0;  // sum n^2, sum e^2, sum n e
};

double bound_numerator(const RowSums& A, const RowSums& B) {
  const double s = A.nn * B.ee + A.ee * B.nn + A.ee * B.ee +
                   2.0 * (A.ne * B.ne + A.ne * B.ee + A.ee * B.ne);
  return std::sqrt(std::max(s, 0.0)) * (1.0 + 0x1p-40);
}

// From the device's sums (bound_sums_kernel): the summation of up to R
// positive terms in FP64 is covered by the factor 1 + (R + 4) 2^-52.
double error_bound(const double* sums, int M, int N, int nparts) {
  const double u = 0x1p-52;
  const double fa = 1.0 + (M + 4.0) * u, fb = 1.0 + (N + 4.0) * u;
  RowSums A, B;
  A.nn = sums[0] * fa;
  A.ee = sums[1] * fa;
  A.ne = sums[2] * fa;
  B.nn = sums[3] * fb;
  B.ee = sums[4] * fb;
  B.ne = sums[5] * fb;
  const double num = bound_numerator(A, B);
  const double fro = std::sqrt(sums[6]) * (1.0 - (nparts + 4.0) * u) *
                     (1.0 - 0x1p-30);
  const double den = fro - num;
  if (!(den > 0)) return INFINITY;
  return num / den * (1.0 + 0x1p-40) + 9e-16;  // + the final roundings
}

}  // namespace

namespace skg {

int ozaki_reserve(skgemm_ctx* c, int kind, int la, int lb, const Problem& pr,
                  int slices) {
  const int M = pr.M, N = pr.N, K = pr.K;
  if ((int64_t)K > kMaxSplitK)
    return set_error(SKGEMM_ERR_NOTSUP, "split form: K = %d exceeds %lld", K,
                     (long long)kMaxSplitK);
  const int L = slices == 0 ? 14 : slices;
  if (L < 8 || L > kMaxMod)
    return set_error(SKGEMM_ERR_ARG, "slices (moduli) must be 0 or 8..%d",
                     kMaxMod);
  const int kpad = (int)round_up(std::max(K, 1), 128);
  if (c->extra == nullptr) {
    c->extra = new skgemm_ctx::Extra();
    c->extra->gen = &c->mem_gen;
  }
  auto* X = c->extra;
  const OzPlan op = choose_plan(c, kind, M, N, L, kpad);
  const GemmCfg& gc = kGemmCfgs[op.cfg];
  const int opa = (int)round_up(M, 256), opb = (int)round_up(N, 256);
  const int mp = (int)round_up(M, 16);
  bool ok = X->get(kPlanesA, (size_t)L * opa * kpad) &&
            X->get(kRes, (size_t)op.kc * L * N * mp) &&
            X->get(kNormA, sizeof(double) * opa) &&
            X->get(kSigA, sizeof(int) * opa) &&
            X->get(kScalars, 8 + sizeof(double) * kCrtGrid);
  if (kind == 0)
    ok = ok && X->get(kPlanesB, (size_t)L * opb * kpad) &&
         X->get(kNormB, sizeof(double) * opb) &&
         X->get(kSigB, sizeof(int) * opb);
  if (!pr.a.kcontig && M > 0)
    ok = ok && X->get(kPartA, sizeof(double) * 2 * norm_segments(M, K) * M);
  if (kind == 0 && !pr.b.kcontig && N > 0)
    ok = ok && X->get(kPartB, sizeof(double) * 2 * norm_segments(N, K) * N);
  // Rounding-error norms and their per-block parts (kpad / 128 blocks, at
  // least the 4096-wide segments of the k-contiguous pass).
  const size_t kbl = (size_t)kpad / 128;
  ok = ok && X->get(kDeltaA, sizeof(double) * opa) &&
       X->get(kDpartA, sizeof(double) * kbl * opa);
  if (kind == 0)
    ok = ok && X->get(kDeltaB, sizeof(double) * opb) &&
         X->get(kDpartB, sizeof(double) * kbl * opb);
  if (!ok) return set_error(SKGEMM_ERR_NOMEM, "split-form workspace");
  if (!op.dp) {
    const int rc = c->ensure_ws((size_t)op.grid * gc.bm * gc.bn / 2 + 1, op.grid);
    if (rc != SKGEMM_OK) return rc;
  }
  Plan pl = plan(c, kind, la, lb, M, N, K);
  return c->ensure_ws(c->ws_doubles(pl.cfg, pl.grid), pl.grid);
}

namespace {

// The split-form sequence with L moduli.  kind 0: C = A B; kind 1:
// C = Psi^T Psi (B == A).  norms_ready: the norms, scales and fallback flag
// of A (and B) are already in the context's buffers.
int run_moduli(skgemm_ctx* c, int kind, int la, int lb, const Problem& pr,
               int L, double* err_bound, bool norms_ready,
               bool timing_begun = false) {
  const Operand& A = pr.a;
  const Operand& B = pr.b;
  const int M = pr.M, N = pr.N, K = pr.K;
  const int kpad = (int)round_up(std::max(K, 1), 128);
  if (hipSetDevice(c->device) != hipSuccess)
    return set_error(SKGEMM_ERR_HIP, "hipSetDevice");
  if (c->extra == nullptr) {
    c->extra = new skgemm_ctx::Extra();
    c->extra->gen = &c->mem_gen;
  }
  const Tables T = make_tables(L);
  const OzPlan op = choose_plan(c, kind, M, N, L, kpad);
  const GemmCfg& gc = kGemmCfgs[op.cfg];
  const int opa = (int)round_up(M, 256), opb = (int)round_up(N, 256);
  const int64_t pa = (int64_t)opa * kpad, pb = (int64_t)opb * kpad;
  auto* X = c->extra;
  int8_t* planesA = (int8_t*)X->get(kPlanesA, (size_t)L * pa);
  int8_t* planesB = kind == 1 ? planesA : (int8_t*)X->get(kPlanesB, (size_t)L * pb);
  const int mp = (int)round_up(M, 16);
  // Pipelined SYRK, opt-in with SKGEMM_OZ_PIPE=1 (k-contiguous operand,
  // data-parallel plan, not inside a capture, not on a serving slot): K is
  // cut into `pipe` chunks of
  // `chunk_k` (a multiple of the residue pass's 4096-wide segments); the
  // residue pass of chunk c + 1 runs on the context's stream while the
  // INT8 products of chunk c run on a second stream, and the CRT pass adds
  // the chunks' residues modulo m.  Integer results are identical to every
  // other plan's.  Off by default: on a 65536 x 1024 operand the overlap
  // costs more than it hides (eight narrower INT8 launches sharing the
  // chip with an HBM-bound pass, eight residue planes per modulus to write
  // and sum), 1.06-1.08 ms against 1.00-1.02 ms for one launch.
  int pipe = 1, chunk_k = 0;
  const char* pipe_env = std::getenv("SKGEMM_OZ_PIPE");
  if (kind == 1 && A.kcontig && op.dp && !c->capturing && !c->oz_no_wait &&
      pipe_env != nullptr && pipe_env[0] == '1') {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &cs) == hipSuccess &&
        cs == hipStreamCaptureStatusNone) {
      chunk_k = (int)round_up((kpad + skgemm_ctx::Extra::kMaxPipe - 1) /
                                  skgemm_ctx::Extra::kMaxPipe,
                              kSplitSeg);
      pipe = (kpad + chunk_k - 1) / chunk_k;
      if (pipe < 2 || chunk_k > kMaxKpad || !X->pipe_ready()) pipe = 1;
    }
  }
  uint8_t* R = (uint8_t*)X->get(
      kRes, (size_t)std::max(op.kc, pipe) * L * N * mp);
  double* n2a = (double*)X->get(kNormA, sizeof(double) * opa);
  double* n2b = kind == 1 ? n2a : (double*)X->get(kNormB, sizeof(double) * opb);
  int* sga = (int*)X->get(kSigA, sizeof(int) * opa);
  int* sgb = kind == 1 ? sga : (int*)X->get(kSigB, sizeof(int) * opb);
  // Squared rounding errors of the scaled rows (for the error bound).
  double* dda = (double*)X->get(kDeltaA, sizeof(double) * opa);
  double* ddb = kind == 1 ? dda : (double*)X->get(kDeltaB, sizeof(double) * opb);
  // [0, 4): fallback flag; [8, 8 + 8 kCrtGrid): per-workgroup parts of
  // ||C||_F^2 written by the CRT kernel.
  char* scal = (char*)X->get(kScalars, 8 + sizeof(double) * kCrtGrid);
  if (!planesA || !planesB || !R || !n2a || !n2b || !sga || !sgb || !scal ||
      !dda || !ddb)
    return set_error(SKGEMM_ERR_NOMEM, "split-form workspace");
  int* fallback = (int*)scal;
  double* fro2 = (double*)(scal + 8);
  const int64_t crt_blocks = (int64_t)((M + 255) / 256) * ((N + 3) / 4);
  const int crt_grid = (int)std::max<int64_t>(
      1, std::min<int64_t>(kCrtGrid, crt_blocks));

  if (norms_ready) pipe = 1;
  // Stage timing (skgemm_timing_stages): one event pair per stage.
  const bool stages = c->stage_timing && pipe == 1;
  // timing_begun: the certifying mode's norm pass opened the event pair.
  if (!timing_begun) c->record_begin();
  hipError_t e = hipSuccess;
  if (!norms_ready) {
    // Cleared by a kernel, not a memset node, so the sequence stays valid
    // inside a captured graph.
    e = launch_zero_ints(fallback, 1, c->stream);
    if (e != hipSuccess) return hip_error(e, "clear split-form flag");
  }
  int rc = SKGEMM_OK;
  if (pipe > 1) {
    // Norms of the whole operand first (the scale of a row needs all of K).
    hipLaunchKernelGGL(norm_k_kernel, dim3(A.extent), dim3(256), 0, c->stream,
                       A, K, T.h, n2a, sga, fallback);
  } else {
    rc = prepare_operand(c, A, K, T, planesA, pa, kpad, opa, n2a, sga,
                         fallback, kPartA, norms_ready, dda, kDpartA);
  }
  if (rc != SKGEMM_OK) return rc;
  if (kind == 0) {
    rc = prepare_operand(c, B, K, T, planesB, pb, kpad, opb, n2b, sgb,
                         fallback, kPartB, norms_ready, ddb, kDpartB);
    if (rc != SKGEMM_OK) return rc;
  }
  if (stages) {
    c->record_end();
    c->record_begin();
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
  G.mp = mp;
  G.r_plane = (int64_t)N * mp;
  G.dp = op.dp;
  // No empty chunk: kcs = ceil(ipt / kc), then kc = ceil(ipt / kcs) <= kc.
  G.kcs = (G.ipt + op.kc - 1) / op.kc;
  G.kc = (G.ipt + G.kcs - 1) / G.kcs;
  G.c0 = 0;
  G.nck = G.kc;
  G.units = G.ntiles * L * (G.dp ? G.kc : 1);
  G.total = G.ntiles * L * G.ipt;
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
  if (pipe > 1) {
    G.kcs = chunk_k / gc.bk;
    G.kc = pipe;
    G.nck = 1;
    G.units = G.ntiles * L;
    G.grid = (int)std::max<int64_t>(1, std::min<int64_t>(op.grid, G.units));
    const int nseg = (kpad + kSplitSeg - 1) / kSplitSeg;
    const int spc = chunk_k / kSplitSeg;  // segments per chunk
    const bool wide = T.h > 55;  // |x'| < 2^54: two 26-bit limbs suffice
    double* dpart =
        (double*)X->get(kDpartA, sizeof(double) * nseg * opa);
    if (dpart == nullptr) return set_error(SKGEMM_ERR_NOMEM, "error norms");
    for (int ch = 0; ch < pipe; ++ch) {
      const int seg0 = ch * spc, nsegc = std::min(spc, nseg - seg0);
      const dim3 grid((unsigned)((int64_t)opa * nsegc));
      if (wide)
        hipLaunchKernelGGL(split_k_kernel<true>, grid, dim3(256), 0,
                           c->stream, A, K, sga, T, planesA, pa, kpad, seg0,
                           nsegc, dpart, opa);
      else
        hipLaunchKernelGGL(split_k_kernel<false>, grid, dim3(256), 0,
                           c->stream, A, K, sga, T, planesA, pa, kpad, seg0,
                           nsegc, dpart, opa);
      if ((e = hipGetLastError()) != hipSuccess)
        return hip_error(e, "ozaki split");
      if ((e = hipEventRecord(X->ev[ch], c->stream)) != hipSuccess ||
          (e = hipStreamWaitEvent(X->aux, X->ev[ch], 0)) != hipSuccess)
        return hip_error(e, "ozaki pipeline");
      G.c0 = ch;
      e = launch_gemm(op.cfg, G, X->aux);
      if (e != hipSuccess) return hip_error(e, "ozaki gemm launch");
    }
    hipLaunchKernelGGL(delta_finish_kernel, dim3((opa + 255) / 256),
                       dim3(256), 0, c->stream, opa, nseg, dpart, opa, dda);
    if ((e = hipEventRecord(X->ev[pipe], X->aux)) != hipSuccess ||
        (e = hipStreamWaitEvent(c->stream, X->ev[pipe], 0)) != hipSuccess)
      return hip_error(e, "ozaki pipeline");
  } else {
    e = launch_gemm(op.cfg, G, c->stream);
    if (e != hipSuccess) return hip_error(e, "ozaki gemm launch");
  }
  if (stages) {
    c->record_end();
    c->record_begin();
  }

  CrtLaunch P{};
  P.R = R;
  P.mp = mp;
  P.kc = pipe > 1 ? pipe : (G.dp ? G.kc : 1);
  P.r_plane = (int64_t)N * mp;
  P.T = T;
  P.sig_a = sga;
  P.sig_b = sgb;
  P.fallback = fallback;
  P.d = pr.d;
  P.M = M;
  P.N = N;
  P.syrk = kind == 1 ? (c->syrk_store ? 2 : 1) : 0;
  P.alpha = pr.alpha;
  P.beta = pr.beta;
  P.gamma = pr.gamma;
  P.fro2 = fro2;
  if (P.kc > 1)
    hipLaunchKernelGGL(crt_kernel<true>, dim3(crt_grid), dim3(256), 0,
                       c->stream, P);
  else
    hipLaunchKernelGGL(crt_kernel<false>, dim3(crt_grid), dim3(256), 0,
                       c->stream, P);
  if ((e = hipGetLastError()) != hipSuccess) return hip_error(e, "crt launch");

  // Plain FP64 fallback, decided on the device: the launch does its work
  // only when a row of A or column of B could not be scaled (fallback set).
  {
    Problem fp = pr;
    Plan pl = plan(c, kind, la, lb, M, N, K);
    pl.kchunks = 1;  // one launch, no chunk buffers
    rc = fill_problem(c, &fp, kind, pl, M, N, K);
    if (rc != SKGEMM_OK) return rc;
    rc = run_single(c, fp, pl, fallback, false);
    if (rc != SKGEMM_OK) return rc;
  }
  c->record_end();

  if (err_bound != nullptr) {
    double* sums = (double*)X->get(kBoundSums, sizeof(double) * 9);
    double* host = X->pinned_sums();
    hipEvent_t rev = X->readback_event();
    if (sums == nullptr || host == nullptr || rev == nullptr)
      return set_error(SKGEMM_ERR_NOMEM, "error-bound sums");
    hipLaunchKernelGGL(bound_sums_kernel, dim3(1), dim3(1024), 0, c->stream,
                       n2a, sga, dda, M, n2b, sgb, ddb, N, K, fro2, crt_grid,
                       fallback, c->status, sums);
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = hipMemcpyAsync(host, sums, sizeof(double) * 9,
                            hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
        (e = hipEventRecord(rev, c->stream)) != hipSuccess)
      return hip_error(e, "error-bound readback");
    while ((e = hipEventQuery(rev)) == hipErrorNotReady) {
    }
    if (e != hipSuccess) return hip_error(e, "error-bound readback");
    // A fix-up wait that gave up (status != 0): the context's handler
    // clears the flags and reports it.
    if (host[8] != 0.0) {
      rc = c->check_status();
      if (rc != SKGEMM_OK) return rc;
    }
    *err_bound = host[7] != 0.0 ? -1.0 : error_bound(host, M, N, crt_grid);
  }
  c->oz_last_moduli = L;
  return SKGEMM_OK;
}

// Plain FP64 in place of the split form (the certifying mode's last resort).
int run_plain(skgemm_ctx* c, int kind, int la, int lb, const Problem& pr,
              double* err_bound) {
  Problem fp = pr;
  const Plan pl = plan(c, kind, la, lb, pr.M, pr.N, pr.K);
  SKG_TRY(fill_problem(c, &fp, kind, pl, pr.M, pr.N, pr.K));
  fp.alpha = pr.alpha;
  fp.beta = pr.beta;
  fp.gamma = pr.gamma;
  SKG_TRY(run_single(c, fp, pl));
  c->oz_last_moduli = 0;
  if (err_bound) *err_bound = -1.0;
  return SKGEMM_OK;
}

// The tolerance the certifying mode (SKGEMM_SLICES_AUTO) guarantees.
constexpr double kCertifyTarget = 1e-13;

// Certifying mode: 14 moduli first (the default count, whose fused norm and
// residue pass reads the operand once more than the residues need); while
// the rigorous a-posteriori bound misses kCertifyTarget, the count the
// bound's scaling predicts (the bound falls by 2^(h(L) - h(14)) with the
// headroom h), then one more at a time, with D restored from a copy first
// when beta != 0; plain FP64 when 16 moduli cannot certify.
int run_certify(skgemm_ctx* c, int kind, int la, int lb, const Problem& pr,
                double* err_bound) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (c->capturing || hipStreamIsCapturing(c->stream, &cs) != hipSuccess ||
      cs != hipStreamCaptureStatusNone)
    return set_error(SKGEMM_ERR_NOTSUP,
                     "the certifying split form reads its bound back and "
                     "cannot be captured; give a modulus count");
  const int M = pr.M;
  hipError_t e;
  if (c->extra == nullptr) {
    c->extra = new skgemm_ctx::Extra();
    c->extra->gen = &c->mem_gen;
  }
  auto* X = c->extra;
  // D is kept when beta != 0, so that a further attempt starts from it.
  void* saved = nullptr;
  int64_t nd = 0;
  if (pr.beta != 0.0) {
    // Elements the output spans: its largest offset is that of (M-1, N-1)
    // in every layout (a panel stride exceeds the panel's contents).
    const Output& d = pr.d;
    const int n = pr.N - 1;
    nd = (int64_t)(n >> d.pshift) * d.pstride +
         (int64_t)(n & ((1 << d.pshift) - 1)) * d.ns +
         (int64_t)(M - 1) * d.ms + 1;
    saved = X->get(kSavedD, sizeof(double) * nd);
    if (saved == nullptr) return set_error(SKGEMM_ERR_NOMEM, "saved D");
    if ((e = hipMemcpyAsync(saved, pr.d.ptr, sizeof(double) * nd,
                            hipMemcpyDeviceToDevice, c->stream)) != hipSuccess)
      return hip_error(e, "save D");
  }
  auto restore = [&]() -> int {
    if (saved != nullptr &&
        (e = hipMemcpyAsync(pr.d.ptr, saved, sizeof(double) * nd,
                            hipMemcpyDeviceToDevice, c->stream)) != hipSuccess)
      return hip_error(e, "restore D");
    return SKGEMM_OK;
  };
  const int h14 = make_tables(14).h;
  double b = 0.0;
  SKG_TRY(run_moduli(c, kind, la, lb, pr, 14, &b, false));
  if (b < 0.0 || b <= kCertifyTarget) {  // -1: computed by plain FP64
    if (err_bound) *err_bound = b;
    return SKGEMM_OK;
  }
  // Predicted count: the smallest whose extra headroom brings b down.
  int L = 15;
  while (L < kMaxMod &&
         b * std::ldexp(1.0, -(make_tables(L).h - h14)) > 0.9 * kCertifyTarget)
    ++L;
  for (; L <= kMaxMod; ++L) {
    SKG_TRY(restore());
    SKG_TRY(run_moduli(c, kind, la, lb, pr, L, &b, false));
    if (b <= kCertifyTarget) {
      if (err_bound) *err_bound = b;
      return SKGEMM_OK;
    }
  }
  SKG_TRY(restore());
  return run_plain(c, kind, la, lb, pr, err_bound);
}

}  // namespace

int ozaki_run(skgemm_ctx* c, int kind, int la, int lb, const Problem& pr,
              int slices, double* err_bound) {
  if ((int64_t)pr.K > kMaxSplitK)
    return set_error(SKGEMM_ERR_NOTSUP,
                     "split form: K = %d exceeds %lld (%d chunks of %d)",
                     pr.K, (long long)kMaxSplitK, kCrtMaxChunks, kMaxKpad);
  if (slices == SKGEMM_SLICES_AUTO)
    return run_certify(c, kind, la, lb, pr, err_bound);
  const int L = slices == 0 ? 14 : slices;
  if (L < 8 || L > kMaxMod)
    return set_error(SKGEMM_ERR_ARG,
                     "slices (moduli) must be %d (certifying), 0 or 8..%d",
                     SKGEMM_SLICES_AUTO, kMaxMod);
  return run_moduli(c, kind, la, lb, pr, L, err_bound, false);
}

}  // namespace skg

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
