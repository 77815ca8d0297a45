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
constexpr int kBK = 64;        // bytes of K per stage
constexpr int kMaxKpad = 131072;  // Kpad * 128^2 <= 2^31

struct Tables {
  int L;
  int h;
  int m[kMaxMod];
  double inv_m[kMaxMod];
  uint8_t cinv[kMaxMod][kMaxMod];  // m_j^{-1} mod m_l (j < l)
};

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(1))) int gint;

constexpr int kRsrcWord3 = 0x00020000;
constexpr int kNumRecords = 0x7FFFFF00;

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

// ---- 1. norms and scale exponents -------------------------------------------

// Operand contiguous along k: one workgroup per free index o.
__global__ void norm_k_kernel(Operand op, int K, int h, double* norm2,
                              int* sigma) {
  __shared__ double red[256];
  const int o = blockIdx.x;
  double s = 0.0;
  if (o < op.extent)
    for (int k = threadIdx.x; k < K; k += 256) {
      const double x = op.ptr[op_addr(op, o, k)];
      s = fma(x, x, s);
    }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    norm2[o] = red[0];
    sigma[o] = sigma_of(red[0], h);
  }
}

// Operand contiguous along o: one thread per o.
__global__ void norm_o_kernel(Operand op, int K, int h, double* norm2,
                              int* sigma) {
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o >= op.extent) return;
  double s = 0.0;
  for (int k = 0; k < K; ++k) {
    const double x = op.ptr[op_addr(op, o, k)];
    s = fma(x, x, s);
  }
  norm2[o] = s;
  sigma[o] = sigma_of(s, h);
}

// ---- 2. residue planes -----------------------------------------------------------

__device__ __forceinline__ int sym_residue(double x, int m, double inv_m) {
  const double q = rint(x * inv_m);
  double r = fma(-q, (double)m, x);  // exact: |r| is a small integer
  const int hi = (m - 1) >> 1, lo = hi - m + 1;
  if (r > hi) r -= m;
  if (r < lo) r += m;
  return (int)r;
}

// planes[l][o][k] (row pitch kpad bytes) for o < opad, k < kpad; zeros past
// the matrix.  Each thread handles one o and 16 consecutive k.
template <int KC>
__global__ void split_kernel(Operand op, int K, const int* sigma, Tables T,
                             int8_t* planes, int64_t plane, int kpad,
                             int opad) {
  int o, kg;
  if (KC) {
    const int kblocks = (kpad / 16 + 255) / 256;
    o = blockIdx.x / kblocks;
    kg = (blockIdx.x % kblocks) * 256 + threadIdx.x;
  } else {
    o = blockIdx.x * 256 + threadIdx.x;
    kg = blockIdx.y;
  }
  if (o >= opad || kg * 16 >= kpad) return;
  const int k0 = kg * 16;
  double x[16];
  const bool in_o = o < op.extent;
  const int s = in_o ? sigma[o] : 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int k = k0 + j;
    const double a = (in_o && k < K) ? op.ptr[op_addr(op, o, k)] : 0.0;
    x[j] = rint(ldexp(a, s));
  }
  for (int l = 0; l < T.L; ++l) {
    const int m = T.m[l];
    const double im = T.inv_m[l];
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t v = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        v |= (uint32_t)(uint8_t)(int8_t)sym_residue(x[4 * q + b], m, im)
             << (8 * b);
      w[q] = v;
    }
    *reinterpret_cast<uint4*>(planes + l * plane + (int64_t)o * kpad + k0) =
        make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// ---- 3. INT8 Stream-K GEMM over (modulus, tile) units ----------------------

struct GemmLaunch {
  const int8_t* A;
  const int8_t* B;
  int64_t a_plane, b_plane;  // bytes per plane
  int kpad;                  // bytes per row
  int M, N, syrk;
  int tiles_m, tiles_n;
  int64_t ntiles;  // per plane
  int ipt;         // kpad / kBK
  int L;
  int mods[kMaxMod];
  uint8_t* R;  // [L][M][N] residues in [0, m)
  int64_t r_plane;
  int grid;
  int64_t total;
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

// LDS row of kBK = 64 bytes = 4 chunks of 16 B; chunk c of row r is stored
// at chunk c ^ ((r >> 2) & 3), so that 16 consecutive rows read at one chunk
// index cover all 16 bank groups.
__device__ __forceinline__ int lds_off(int r, int c) {
  return r * kBK + ((c ^ ((r >> 2) & 3)) << 4);
}

template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(WM* WN * 64) void gemm_i8_kernel(GemmLaunch G) {
  constexpr int NT = WM * WN * 64;
  constexpr int TM = BM / WM, TN = BN / WN;
  constexpr int MB = TM / 32, NB = TN / 32;
  constexpr int CA = BM * 4 / NT, CB = BN * 4 / NT;  // 16-byte chunks/thread
  static_assert(BM * 4 % NT == 0 && BN * 4 % NT == 0, "tile vs threads");
  __shared__ __align__(16) int8_t lds[2][(BM + BN) * kBK];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave % WM, wn = wave / WM;
  const int Gn = G.grid, g = blockIdx.x;
  const int per = Gn >> 3, rem = Gn & 7, xcd = g & 7;
  const int w = xcd * per + (xcd < rem ? xcd : rem) + (g >> 3);
  int64_t it = wg_start(w, G.total, Gn);
  const int64_t end = wg_start(w + 1, G.total, Gn);

  // Per-thread chunk coordinates (row, chunk) and their byte offsets.
  int rowA[CA], chA[CA], rowB[CB], chB[CB];
#pragma unroll
  for (int p = 0; p < CA; ++p) {
    const int q = p * NT + tid;
    rowA[p] = q >> 2;
    chA[p] = q & 3;
  }
#pragma unroll
  for (int p = 0; p < CB; ++p) {
    const int q = p * NT + tid;
    rowB[p] = q >> 2;
    chB[p] = q & 3;
  }

  while (it < end) {
    const int64_t unit = it / G.ipt;
    const int64_t t0 = unit * G.ipt;
    const int l = (int)(unit / G.ntiles);
    const int64_t tile = unit - (int64_t)l * G.ntiles;
    const int kb = (int)(it - t0);
    const int ke = (int)((end - t0) < G.ipt ? (end - t0) : G.ipt);
    int ti, tj;
    if (G.syrk) {
      tri_tile(tile, ti, tj);
    } else {
      ti = (int)(tile / G.tiles_n);
      tj = (int)(tile % G.tiles_n);
    }
    const int m0 = ti * BM, n0 = tj * BN;
    const int8_t* baseA = G.A + l * G.a_plane + (int64_t)m0 * G.kpad;
    const int8_t* baseB = G.B + l * G.b_plane + (int64_t)n0 * G.kpad;

    v16i acc[MB][NB];
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[i][j][v] = 0;

    v4i ra[CA], rb[CB];
    auto load = [&](int kk) {
      const __amdgpu_buffer_rsrc_t sa = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<int8_t*>(baseA + (int64_t)kk * kBK), (short)0,
          kNumRecords, kRsrcWord3);
      const __amdgpu_buffer_rsrc_t sb = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<int8_t*>(baseB + (int64_t)kk * kBK), (short)0,
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
#pragma unroll
      for (int p = 0; p < CA; ++p)
        *reinterpret_cast<v4i*>(&lds[buf][lds_off(rowA[p], chA[p])]) = ra[p];
#pragma unroll
      for (int p = 0; p < CB; ++p)
        *reinterpret_cast<v4i*>(&lds[buf][BM * kBK + lds_off(rowB[p], chB[p])]) =
            rb[p];
    };
    load(kb);
    store(0);
    __syncthreads();
    int cur = 0;
    for (int kk = kb; kk < ke; ++kk) {
      const bool more = kk + 1 < ke;
      if (more) load(kk + 1);
      const int8_t* la = &lds[cur][0];
      const int8_t* lb = &lds[cur][BM * kBK];
#pragma unroll
      for (int s = 0; s < kBK / 32; ++s) {
        const int c = 2 * s + (lane >> 5);
        v4i a[MB], b[NB];
#pragma unroll
        for (int i = 0; i < MB; ++i)
          a[i] = *reinterpret_cast<const v4i*>(
              la + lds_off(wm * TM + i * 32 + (lane & 31), c));
#pragma unroll
        for (int j = 0; j < NB; ++j)
          b[j] = *reinterpret_cast<const v4i*>(
              lb + lds_off(wn * TN + j * 32 + (lane & 31), c));
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

    // Stream-K fix-up on exact INT32 partials (device-coherent sc1 traffic,
    // no L2-wide fences; same protocol as the FP64 kernel).
    bool write_out = true;
    if (kb != 0) {
      gint* slot = (gint*)(G.ws) + (int64_t)w * (BM * BN);
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int v = 0; v < 16; ++v)
            __hip_atomic_store(slot + ((i * NB + j) * 16 + v) * NT + tid,
                               acc[i][j][v], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
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
        gint* slot = (gint*)(G.ws) + (int64_t)w2 * (BM * BN);
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v)
              acc[i][j][v] += __hip_atomic_load(
                  slot + ((i * NB + j) * 16 + v) * NT + tid, __ATOMIC_RELAXED,
                  __HIP_MEMORY_SCOPE_AGENT);
      }
    }

    if (write_out) {
      const int m = G.mods[l];
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
              int r = acc[i][j][v] % m;
              if (r < 0) r += m;
              rp[(int64_t)row * G.N + col] = (uint8_t)r;
            }
          }
        }
    }
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
  Output d;
  int M, N, syrk;
  double alpha, beta, gamma;
  double* fro2;  // += ||C||_F^2 (SYRK: of the full symmetric matrix)
};

__global__ void crt_kernel(CrtLaunch p) {
  __shared__ double red[256];
  const int col = blockIdx.x * 64 + (threadIdx.x & 63);
  const int row = blockIdx.y * 4 + (threadIdx.x >> 6);
  double contrib = 0.0;
  if (row < p.M && col < p.N && (!p.syrk || row >= col)) {
    const int64_t idx = (int64_t)row * p.N + col;
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
  int bm, bn, wm, wn;
};
static const GemmCfg kGemmCfgs[] = {{256, 256, 4, 2}, {128, 128, 2, 2}};

template <int BM, int BN, int WM, int WN>
static int occ_t() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &nb, reinterpret_cast<const void*>(gemm_i8_kernel<BM, BN, WM, WN>),
          WM * WN * 64, 0) != hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

static int gemm_occupancy(int cfg) {
  return cfg == 0 ? occ_t<256, 256, 4, 2>() : occ_t<128, 128, 2, 2>();
}

static hipError_t launch_gemm(int cfg, const GemmLaunch& G, hipStream_t s) {
  if (cfg == 0)
    hipLaunchKernelGGL((gemm_i8_kernel<256, 256, 4, 2>), dim3(G.grid),
                       dim3(512), 0, s, G);
  else
    hipLaunchKernelGGL((gemm_i8_kernel<128, 128, 2, 2>), dim3(G.grid),
                       dim3(256), 0, s, G);
  return hipGetLastError();
}

static int modinv(int a, int m) {
  a %= m;
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 0;
}

Tables make_tables(int L) {
  Tables T{};
  T.L = L;
  double bits = 0.0;
  for (int l = 0; l < L; ++l) {
    T.m[l] = kModuli[l];
    T.inv_m[l] = 1.0 / kModuli[l];
    bits += std::log2((double)kModuli[l]);
  }
  for (int l = 0; l < L; ++l)
    for (int j = 0; j < l; ++j) T.cinv[j][l] = (uint8_t)modinv(kModuli[j], kModuli[l]);
  T.h = (int)std::floor(bits / 2.0) - 1;
  return T;
}

}  // namespace oz
}  // namespace skg

// ---- context extras ------------------------------------------------------------

struct skgemm_ctx::Extra {
  void* buf[8] = {nullptr};
  size_t n[8] = {0};
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
  for (int i = 0; i < 8; ++i)
    if (extra->buf[i]) hipFree(extra->buf[i]);
  delete extra;
  extra = nullptr;
}

using namespace skg;
using namespace skg::oz;

namespace {

enum { kPlanesA = 0, kPlanesB, kRes, kNormA, kNormB, kSigA, kSigB, kFro };

int64_t round_up(int64_t x, int64_t q) { return (x + q - 1) / q * q; }

// Scale + split one operand (free index o, contraction k) into planes.
int prepare_operand(skgemm_ctx* c, const Operand& op, int K, const Tables& T,
                    int8_t* planes, int64_t plane, int kpad, int opad,
                    double* norm2, int* sigma) {
  const int O = op.extent;
  if (O > 0) {
    if (op.kcontig)
      hipLaunchKernelGGL(norm_k_kernel, dim3(O), dim3(256), 0, c->stream, op, K,
                         T.h, norm2, sigma);
    else
      hipLaunchKernelGGL(norm_o_kernel, dim3((O + 255) / 256), dim3(256), 0,
                         c->stream, op, K, T.h, norm2, sigma);
  }
  const int kg = kpad / 16;
  if (op.kcontig)
    hipLaunchKernelGGL(split_kernel<1>, dim3(((kg + 255) / 256) * opad),
                       dim3(256), 0, c->stream, op, K, sigma, T, planes, plane,
                       kpad, opad);
  else
    hipLaunchKernelGGL(split_kernel<0>, dim3((opad + 255) / 256, kg), dim3(256),
                       0, c->stream, op, K, sigma, T, planes, plane, kpad, opad);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(e, "ozaki split");
  return SKGEMM_OK;
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
int ozaki_run(skgemm_ctx* c, int kind, const Operand& A, const Operand& B,
              const Output& D, int M, int N, int K, double alpha, double beta,
              double gamma, int slices, double* err_bound) {
  int L = slices == 0 ? 14 : slices;
  if (L < 8 || L > kMaxMod)
    return set_error(SKGEMM_ERR_ARG, "slices (moduli) must be 0 or 8..%d",
                     kMaxMod);
  const int kpad = (int)round_up(std::max(K, 1), kBK);
  if (kpad > kMaxKpad)
    return set_error(SKGEMM_ERR_NOTSUP,
                     "split form needs K <= %d for exact INT32 sums", kMaxKpad);
  if (hipSetDevice(c->device) != hipSuccess)
    return set_error(SKGEMM_ERR_HIP, "hipSetDevice");
  if (c->extra == nullptr) c->extra = new skgemm_ctx::Extra();
  const Tables T = make_tables(L);
  const int cfg = (M >= 1024 && N >= 1024) ? 0 : 1;
  const GemmCfg& gc = kGemmCfgs[cfg];
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
  double* fro2 = (double*)X->get(kFro, sizeof(double));
  if (!planesA || !planesB || !R || !n2a || !n2b || !sga || !sgb || !fro2)
    return set_error(SKGEMM_ERR_NOMEM, "split-form workspace");

  c->record_begin();
  int rc = prepare_operand(c, A, K, T, planesA, pa, kpad, opa, n2a, sga);
  if (rc != SKGEMM_OK) return rc;
  if (kind == 0) {
    rc = prepare_operand(c, B, K, T, planesB, pb, kpad, opb, n2b, sgb);
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
  G.ipt = kpad / kBK;
  G.L = L;
  for (int l = 0; l < L; ++l) G.mods[l] = T.m[l];
  G.R = R;
  G.r_plane = (int64_t)M * N;
  G.total = G.ntiles * L * G.ipt;
  const int cap = gemm_occupancy(cfg) * c->num_cus;
  int grid = c->forced_grid > 0 ? std::min(c->forced_grid, cap) : cap;
  grid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, G.total));
  G.grid = grid;
  rc = c->ensure_ws((size_t)grid * gc.bm * gc.bn / 2 + 1, grid);
  if (rc != SKGEMM_OK) return rc;
  G.ws = reinterpret_cast<int*>(c->ws);
  G.flags = c->flags;
  G.status = c->status;
  G.spin_limit = c->spin_limit;
  hipError_t e = launch_gemm(cfg, G, c->stream);
  if (e != hipSuccess) return hip_error(e, "ozaki gemm launch");

  CrtLaunch P{};
  P.R = R;
  P.r_plane = (int64_t)M * N;
  P.T = T;
  P.sig_a = sga;
  P.sig_b = sgb;
  P.d = D;
  P.M = M;
  P.N = N;
  P.syrk = kind == 1;
  P.alpha = alpha;
  P.beta = beta;
  P.gamma = gamma;
  P.fro2 = fro2;
  // Cleared by a kernel, not a memset node, so the sequence stays valid
  // inside a captured graph.
  if ((e = launch_zero_ints(reinterpret_cast<int*>(fro2), 2, c->stream)) !=
      hipSuccess)
    return hip_error(e, "clear norm accumulator");
  hipLaunchKernelGGL(crt_kernel, dim3((N + 63) / 64, (M + 3) / 4), dim3(256),
                     0, c->stream, P);
  if ((e = hipGetLastError()) != hipSuccess) return hip_error(e, "crt launch");
  c->record_end();

  if (err_bound != nullptr) {
    std::vector<double> na(M), nb(N), f(1);
    std::vector<int> sa(M), sb(N);
    if ((e = hipMemcpyAsync(na.data(), n2a, sizeof(double) * M,
                            hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(sa.data(), sga, sizeof(int) * M,
                            hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(nb.data(), n2b, sizeof(double) * N,
                            hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(sb.data(), sgb, sizeof(int) * N,
                            hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(f.data(), fro2, sizeof(double),
                            hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(c->stream)) != hipSuccess)
      return hip_error(e, "error-bound readback");
    rc = c->check_status();
    if (rc != SKGEMM_OK) return rc;
    *err_bound = error_bound(na, sa, nb, sb, K, f[0]);
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
  return ozaki_run(c, 0, p.a, p.b, p.d, M, N, K, alpha, beta, gamma, slices,
                   err_bound);
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
  return ozaki_run(c, 1, p.a, p.a, p.d, N, N, K, alpha, beta, 0.0, slices,
                   err_bound);
}
