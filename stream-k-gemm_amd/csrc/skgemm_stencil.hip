// skgemm_stencil.hip - real-space finite-difference stencil H psi on the
// 16-band-packed extended grid (include/stencil_extended_packed.hpp) and the
// single-domain halo fill (include/halo_packed.hpp).
//
// Layouts: local grid [ceil(nb/16)][K][16], K = ni nj nk, g = i + j ni +
// k ni nj; extended grid [ceil(nb/16)][K_ex][16] with the interior at offset
// (F, F, F), K_ex = (ni+2F)(nj+2F)(nk+2F); output [ceil(nb/T)][K][T], T = 32
// or 16.  A grid point's 16 bands are 128 contiguous bytes, so a wave (16
// band lanes x 4 consecutive grid points along i) reads and writes 512-byte
// runs.  Both kernels are HBM-bound.  The stencil takes its x / y
// neighbours from an LDS copy of each plane (the block's points and their
// cross-shaped halo) and its z neighbours from registers, so the extended
// input is read about once from HBM (the halo re-reads hit L2); its
// algorithmic traffic is the extended input once, the diagonal once and the
// output once.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/skgemm.h"
#include "skgemm_ctx.h"

namespace skg {
namespace {

constexpr int kMaxFDn = 8;

struct StencilCoef {
  double x[kMaxFDn + 1], y[kMaxFDn + 1], z[kMaxFDn + 1];
  double c0;
};

// One workgroup: 16 band lanes x an 8 x 8 block of (i, j) points of one
// 16-band tile (1024 threads), marching through kStencilPlanes planes of k;
// ceil(ni/8) ceil(nj/8) ceil(nk/kStencilPlanes) tiles16 workgroups.  A thread
// keeps its column's 2F + 1 values along k in registers (one new z load per
// plane).  Per plane the block's centres and the cross-shaped x / y halo
// (F points either side of each block row and column, 32 F points) go into
// LDS, from which the 4F x / y neighbours are read: about 1 + 32F/64 global
// loads per output instead of 4F + 1.  The next plane's halo is loaded into
// registers while the current plane is computed; the barriers wait on LDS
// only, so those loads stay in flight across them.
constexpr int kStencilPlanes = 8;
constexpr int kSB = 8;  // block edge (points along i and along j)

// Two workgroups per CU (8 waves per SIMD, at most 64 VGPRs) up to FDn 6:
// 157 us against 170 us with one on the c2 grid (32^3, 520 bands, FDn 6);
// FDn 7 and 8 would spill there and keep one.
template <int F>
__global__ __launch_bounds__(1024)
__attribute__((amdgpu_waves_per_eu(F <= 6 ? 8 : 4, 8))) void stencil_kernel(
    const double* __restrict__ ext, const double* __restrict__ A_diag, int ni,
    int nj, int nk, int nb, StencilCoef c, double* __restrict__ out,
    int tshift) {
  constexpr int E = kSB + 2 * F;             // LDS plane edge
  constexpr int kHalo = 4 * kSB * F;         // halo points per plane
  constexpr int kHs = (kHalo + 63) / 64;     // halo points per thread
  __shared__ double sl[E * E * 16];          // [j + F][i + F][band]
  const int tid = threadIdx.x;
  const int bl = tid & 15;
  const int p = tid >> 4;                    // 0..63
  // Work unit u = (tile, k chunk, i-j block), block fastest; hardware
  // workgroup g runs on XCD g % 8 and each XCD takes a contiguous range of
  // units, so the blocks and k chunks of one tile (whose x / y halos and z
  // prologues overlap) meet in one L2.
  const int nbi = (ni + kSB - 1) / kSB;
  const int nblk = nbi * ((nj + kSB - 1) / kSB);
  const int nkc = (nk + kStencilPlanes - 1) / kStencilPlanes;
  const int units = gridDim.x;
  const int xcd = blockIdx.x & 7, per = units >> 3, xr = units & 7;
  const int u = xcd * per + (xcd < xr ? xcd : xr) + (blockIdx.x >> 3);
  const int blkid = u % nblk;
  const int i0 = (blkid % nbi) * kSB, j0 = (blkid / nbi) * kSB;
  const int pi = p % kSB, pj = p / kSB;
  const int i = i0 + pi, j = j0 + pj;
  const int k0 = (u / nblk) % nkc * kStencilPlanes;
  const int k1 = k0 + kStencilPlanes < nk ? k0 + kStencilPlanes : nk;
  const int t16 = u / (nblk * nkc);
  const int b = t16 * 16 + bl;
  const bool band_ok = b < nb;
  const bool own = band_ok && i < ni && j < nj;
  // Block points past the domain edge but inside the extended grid hold
  // ghost values that the edge points read as x / y neighbours.
  const bool inext = band_ok && i < ni + F && j < nj + F;
  const int64_t K = (int64_t)ni * nj * nk;
  const int nie = ni + 2 * F;
  const int64_t sj = nie;
  const int64_t sk = sj * (nj + 2 * F);
  const int64_t Kex = sk * (nk + 2 * F);
  const double* s = ext + (int64_t)t16 * Kex * 16 + bl;
  const int T = 1 << tshift;
  double* o = out + ((int64_t)(b >> tshift) * K) * T + (b & (T - 1));

  // Halo points of this thread: index h = p + 64 m.  h < 2 kSB F: x halo of
  // block row h / (2F), offset h % (2F) -> i = -F .. -1, kSB .. kSB + F - 1;
  // else the y halo of block column (h') / (2F) likewise along j.
  int hoff[kHs];     // LDS offset (points) of halo point m, -1: none
  int hcol[kHs];     // extended column index (i + F) + (j + F) sj
#pragma unroll
  for (int m = 0; m < kHs; ++m) {
    const int h = p + 64 * m;
    hoff[m] = -1;
    hcol[m] = 0;
    if (h < kHalo && band_ok) {
      int hi, hj;
      const int x = h < 2 * kSB * F ? h : h - 2 * kSB * F;
      const int line = x / (2 * F), d = x % (2 * F);
      const int a = d < F ? d - F : kSB + d - F;
      if (h < 2 * kSB * F) {
        hi = a;
        hj = line;
      } else {
        hi = line;
        hj = a;
      }
      // Used only by block points inside the grid; those halo points lie
      // inside the extended grid.
      const int gi = i0 + hi, gj = j0 + hj;
      const bool need = (h < 2 * kSB * F) ? (j0 + line < nj && gi < ni + F)
                                           : (i0 + line < ni && gj < nj + F);
      if (need) {
        hoff[m] = (hj + F) * E + (hi + F);
        hcol[m] = (gi + F) + (gj + F) * nie;
      }
    }
  }
  // Column of this thread's own point.
  // (32-bit: the host keeps an extended plane below 2^31 points)
  const int col = (i + F) + (j + F) * nie;
  double q[2 * F + 1];  // q[F + d] = psi at plane k + d
  double hv[kHs];
#pragma unroll
  for (int d = 0; d < 2 * F; ++d)
    q[d] = inext ? s[(col + (int64_t)(k0 + d) * sk) * 16] : 0.0;
  auto load_halo = [&](int k) {
#pragma unroll
    for (int m = 0; m < kHs; ++m)
      hv[m] = hoff[m] >= 0 ? s[(hcol[m] + (int64_t)(k + F) * sk) * 16] : 0.0;
  };
  load_halo(k0);
  // The column's value at extended plane k + 2F is loaded one plane ahead
  // (qn), and the diagonal before the plane's LDS stores and barrier, so
  // neither load is waited for right after it is issued.
  const int64_t gcol = i + (int64_t)j * ni;
  double qn = inext ? s[(col + (int64_t)(k0 + 2 * F) * sk) * 16] : 0.0;
  const int me = ((pj + F) * E + (pi + F)) * 16 + bl;
  for (int k = k0; k < k1; ++k) {
    q[2 * F] = qn;
    const double ak = own ? A_diag[gcol + (int64_t)k * ni * nj] : 0.0;
    if (k + 1 < k1)
      qn = inext ? s[(col + (int64_t)(k + 1 + 2 * F) * sk) * 16] : 0.0;
    sl[me] = q[F];
#pragma unroll
    for (int m = 0; m < kHs; ++m)
      if (hoff[m] >= 0) sl[hoff[m] * 16 + bl] = hv[m];
    if (k + 1 < k1) load_halo(k + 1);
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): LDS stores done
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (own) {
      const int64_t g = gcol + (int64_t)k * ni * nj;
      // Same order as the reference's scalar version: centre, x terms, y
      // terms, z terms, r ascending in each (each term one fused
      // multiply-add).
      double res = q[F] * (ak + c.c0);
#pragma unroll
      for (int r = 1; r <= F; ++r)
        res = fma(c.x[r], sl[me + r * 16] + sl[me - r * 16], res);
#pragma unroll
      for (int r = 1; r <= F; ++r)
        res = fma(c.y[r], sl[me + r * E * 16] + sl[me - r * E * 16], res);
#pragma unroll
      for (int r = 1; r <= F; ++r) res = fma(c.z[r], q[F + r] + q[F - r], res);
      o[g * T] = res;
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);  // this thread's LDS reads done
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
#pragma unroll
    for (int d = 0; d < 2 * F; ++d) q[d] = q[d + 1];
  }
}

// Every extended point of every tile: the interior copies the local point,
// the halo the wrapped one (mode 1), zero (mode 0) or is left alone (mode
// 2).  A workgroup (16 band lanes x 16 points along i) fills whole rows of
// kHaloPlanes planes of one j; grid (nj_ex, ceil(nk_ex / kHaloPlanes) x
// tiles16).  The j and k wraps are formed once per row.
constexpr int kHaloPlanes = 4;

__device__ __forceinline__ int wrap(int x, int n) { return ((x % n) + n) % n; }

__global__ __launch_bounds__(256) void halo_extend_kernel(
    const double* __restrict__ local, int ni, int nj, int nk, int F, int mode,
    double* __restrict__ ext) {
  const int bl = threadIdx.x & 15, li = threadIdx.x >> 4;
  const int nie = ni + 2 * F, nje = nj + 2 * F, nke = nk + 2 * F;
  const int kchunks = (nke + kHaloPlanes - 1) / kHaloPlanes;
  const int je = blockIdx.x;
  const int t16 = blockIdx.y / kchunks;
  const int ke0 = (blockIdx.y % kchunks) * kHaloPlanes;
  const int ke1 = ke0 + kHaloPlanes < nke ? ke0 + kHaloPlanes : nke;
  const int64_t K = (int64_t)ni * nj * nk;
  const int64_t Kex = (int64_t)nie * nje * nke;
  const double* src = local + (int64_t)t16 * K * 16 + bl;
  double* dst = ext + (int64_t)t16 * Kex * 16 + bl;
  const int j = je - F;
  const bool jin = j >= 0 && j < nj;
  const int jj = wrap(j, nj);
  for (int ke = ke0; ke < ke1; ++ke) {
    const int k = ke - F;
    const bool row_in = jin && k >= 0 && k < nk;
    if (!row_in && mode == 2) continue;  // uniform: no interior point here
    const int64_t srow = ((int64_t)wrap(k, nk) * nj + jj) * ni;
    const int64_t drow = ((int64_t)ke * nje + je) * nie;
    for (int ie = li; ie < nie; ie += 16) {
      const int i = ie - F;
      const bool inside = row_in && i >= 0 && i < ni;
      if (!inside && mode == 2) continue;
      double v = 0.0;
      if (inside || mode == 1) v = src[(srow + wrap(i, ni)) * 16];
      dst[(drow + ie) * 16] = v;
    }
  }
}

// The same fill as 16-byte moves, for 16-byte aligned grids whose extended
// plane holds fewer than 2^31 double pairs: a workgroup takes kHaloChunk
// consecutive pairs of one extended plane of one tile (nje rows of nie
// points of 8 pairs, as stored), and each thread issues the loads of its
// kHaloBatch pairs before it stores any of them.  Grid: tiles16 x nke x
// ceil(plane pairs / kHaloChunk), plane-major.  On the 32^3 grid with 520
// bands and FDn 6, 16 KB per workgroup and 4 loads per thread ran at the
// copy kernel's rate (96-97 us); 32 and 64 KB workgroups took 103-111 us,
// 4 KB ones (one pair per thread) 127 us.
typedef double hpair __attribute__((ext_vector_type(2)));
constexpr int kHaloChunk = 1024;  // pairs (16 KB) per workgroup
constexpr int kHaloBatch = 4;     // pairs per thread in flight

__global__ __launch_bounds__(256) void halo_extend_vec_kernel(
    const hpair* __restrict__ local, int ni, int nj, int nk, int F, int mode,
    int chunks, hpair* __restrict__ ext) {
  const int nie = ni + 2 * F, nje = nj + 2 * F, nke = nk + 2 * F;
  const int nie8 = nie * 8;
  const int pp = nje * nie8;  // pairs per extended plane
  const int pl = blockIdx.x / chunks;  // t16 * nke + ke
  const int c = blockIdx.x - pl * chunks;
  const int t16 = pl / nke;
  const int ke = pl - t16 * nke;
  const int k = ke - F;
  const bool kin = k >= 0 && k < nk;
  if (!kin && mode == 2) return;  // a plane of ghost points only
  const int kk = kin ? k : wrap(k, nk);
  const hpair* src = local + ((int64_t)t16 * nk + kk) * nj * ni * 8;
  hpair* dst = ext + (int64_t)pl * pp;
  const int e1 = min(c * kHaloChunk + kHaloChunk, pp);
  for (int b = c * kHaloChunk + (int)threadIdx.x; b < e1;
       b += 256 * kHaloBatch) {
    hpair v[kHaloBatch];
    bool st[kHaloBatch];
#pragma unroll
    for (int q = 0; q < kHaloBatch; ++q) {
      const int e = b + 256 * q;
      const int je = e / nie8;
      const int r = e - je * nie8;
      const int j = je - F, i = (r >> 3) - F;
      const bool in = kin && j >= 0 && j < nj && i >= 0 && i < ni;
      st[q] = e < e1 && (in || mode != 2);
      v[q] = hpair{0.0, 0.0};
      if (e < e1 && (in || mode == 1)) {
        const int jj = in ? j : wrap(j, nj);
        const int ii = in ? i : wrap(i, ni);
        v[q] = src[((int64_t)jj * ni + ii) * 8 + (r & 7)];
      }
    }
#pragma unroll
    for (int q = 0; q < kHaloBatch; ++q)
      if (st[q]) dst[b + 256 * q] = v[q];
  }
}

// stencil_kernel with two adjacent bands per lane (16-byte loads, stores
// and LDS accesses) and no load waited for before its plane: 8 band pairs x
// the 8 x 8 block (512 threads), the same work units, LDS layout and order
// of terms.  The plane loop is fully unrolled and every plane runs (planes
// past nk load from the last plane and store nothing), so the column shift
// is resolved at compile time and no register whose load is in flight is
// copied (the rolled loop's copies waited for every load at each plane's
// end); loads are unconditional (clamped to valid points), and global
// accesses are buffer operations from resources based at the workgroup's
// first plane, with fixed per-thread offsets and a uniform per-plane one.
// Each plane's column value, diagonal and halo are loaded one plane ahead.
// For 16-byte aligned grids and outputs whose kPairPlanes + 2 FDn
// extended planes and kPairPlanes output planes span fewer than
// kStencilRecords bytes.  On the 32^3 grid with 520 bands and FDn 6:
// 120 us against 143 us for stencil_kernel (one band per lane in this form
// took 132 us; loads two planes ahead 127 us).
// Planes per workgroup: 16 here (the z prologue of 2 FDn planes is then
// shared by twice the outputs; 120 us against 127 us with 8 on one box),
// 8 for stencil_kernel (16 measured slower there).
constexpr int kPairPlanes = 16;
constexpr int kStencilRsrc3 = 0x00020000;        // raw buffer (gfx9 family)
constexpr int kStencilRecords = 0x7FFFFF00;      // bytes addressable

typedef unsigned int su4 __attribute__((ext_vector_type(4)));
typedef unsigned int su2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ hpair ld_pair(__amdgpu_buffer_rsrc_t r, int v,
                                         int s) {
  return __builtin_bit_cast(hpair,
                            __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0));
}

template <int F>
__global__ __launch_bounds__(512)
__attribute__((amdgpu_waves_per_eu(2, 8))) void stencil_pair_kernel(
    const double* ext, const double* A_diag, int ni, int nj, int nk, int nb,
    StencilCoef c, double* out, int tshift) {
  // (No __restrict__: loads through restrict-qualified read-only arguments
  // may be moved to their uses, past the barriers.)
  constexpr int E = kSB + 2 * F;             // LDS plane edge
  constexpr int kHalo = 4 * kSB * F;         // halo points per plane
  constexpr int kHs = (kHalo + 63) / 64;     // halo points per thread
  __shared__ hpair sl[E * E * 8];            // [j + F][i + F][band pair]
  const int tid = threadIdx.x;
  const int bp = tid & 7;
  const int p = tid >> 3;                    // 0..63
  const int nbi = (ni + kSB - 1) / kSB;
  const int nblk = nbi * ((nj + kSB - 1) / kSB);
  const int nkc = (nk + kPairPlanes - 1) / kPairPlanes;
  const int units = gridDim.x;
  const int xcd = blockIdx.x & 7, per = units >> 3, xr = units & 7;
  const int u = xcd * per + (xcd < xr ? xcd : xr) + (blockIdx.x >> 3);
  const int blkid = u % nblk;
  const int i0 = (blkid % nbi) * kSB, j0 = (blkid / nbi) * kSB;
  const int pi = p % kSB, pj = p / kSB;
  const int i = i0 + pi, j = j0 + pj;
  const int k0 = (u / nblk) % nkc * kPairPlanes;
  const int k1 = k0 + kPairPlanes < nk ? k0 + kPairPlanes : nk;
  const int t16 = u / (nblk * nkc);
  const int b = t16 * 16 + 2 * bp;
  const bool band_ok = b < nb;
  const bool two = b + 1 < nb;
  const bool own = band_ok && i < ni && j < nj;
  const int64_t K = (int64_t)ni * nj * nk;
  const int nie = ni + 2 * F;
  const int sk = nie * (nj + 2 * F);         // points per extended plane
  const int64_t Kex = (int64_t)sk * (nk + 2 * F);
  const int T = 1 << tshift;
  const int64_t plane = (int64_t)ni * nj;
  // Resources based at plane k0 (the extended grid's plane k0 holds grid
  // plane k0 - F); (t16 * 16 + 2 bp) >> tshift is t16 * 16 >> tshift.
  const __amdgpu_buffer_rsrc_t re = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<double*>(ext) + ((int64_t)t16 * Kex + (int64_t)k0 * sk) * 16,
      (short)0, kStencilRecords, kStencilRsrc3);
  const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<double*>(A_diag) + k0 * plane, (short)0, kStencilRecords,
      kStencilRsrc3);
  const __amdgpu_buffer_rsrc_t ro = __builtin_amdgcn_make_buffer_rsrc(
      out + (int64_t)((t16 * 16) >> tshift) * K * T + k0 * plane * T,
      (short)0, kStencilRecords, kStencilRsrc3);

  // Loads are unconditional: a block point outside the extended grid reads
  // the nearest column inside it, an unneeded halo slot this thread's own
  // column and a point outside the grid the nearest diagonal entry; none of
  // those values reaches an output (the neighbours of a point inside the
  // grid lie inside the extended grid, and only such points store).  Halo
  // slots that load nothing useful store to the plane's corner point
  // (i, j) = (-F, -F), which no neighbour read touches.
  const int ic = i < ni + F ? i : ni + F - 1, jc = j < nj + F ? j : nj + F - 1;
  const int col = (ic + F) + (jc + F) * nie;
  int hoff[kHs];     // LDS offset (points) of halo point m
  int hby[kHs];      // byte offset of halo point m in an extended plane
#pragma unroll
  for (int m = 0; m < kHs; ++m) {
    const int h = p + 64 * m;
    hoff[m] = 0;
    int hcol = col;
    if (h < kHalo && band_ok) {
      int hi, hj;
      const int x = h < 2 * kSB * F ? h : h - 2 * kSB * F;
      const int line = x / (2 * F), d = x % (2 * F);
      const int a = d < F ? d - F : kSB + d - F;
      if (h < 2 * kSB * F) {
        hi = a;
        hj = line;
      } else {
        hi = line;
        hj = a;
      }
      const int gi = i0 + hi, gj = j0 + hj;
      const bool need = (h < 2 * kSB * F) ? (j0 + line < nj && gi < ni + F)
                                           : (i0 + line < ni && gj < nj + F);
      if (need) {
        hoff[m] = (hj + F) * E + (hi + F);
        hcol = (gi + F) + (gj + F) * nie;
      }
    }
    hby[m] = hcol * 128 + bp * 16;
  }
  const int qby = col * 128 + bp * 16;
  const int dby = ((i < ni ? i : ni - 1) + (j < nj ? j : nj - 1) * ni) * 8;
  const int oby = ((i < ni ? i : 0) + (j < nj ? j : 0) * ni) * T * 8 +
                  (b & (T - 1)) * 8;
  const int pb = sk * 128;                   // bytes per extended plane
  hpair q[2 * F + 1];  // q[F + d] = psi at plane k + d
#pragma unroll
  for (int d = 0; d < 2 * F; ++d) q[d] = ld_pair(re, qby, d * pb);
  // Plane k0 + t's column value at extended plane k0 + t + 2F, diagonal and
  // halo; planes past nk (a short last chunk's tail) load from the last.
  const int tl = nk - 1 - k0;
  hpair qn, hv[kHs];
  double ak;
  auto load_next = [&](int t) __attribute__((always_inline)) {
    const int tc = t < tl ? t : tl;
    qn = ld_pair(re, qby, (tc + 2 * F) * pb);
    ak = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                        ra, dby, tc * (int)plane * 8, 0));
#pragma unroll
    for (int m = 0; m < kHs; ++m) hv[m] = ld_pair(re, hby[m], (tc + F) * pb);
  };
  load_next(0);
  const int me = ((pj + F) * E + (pi + F)) * 8 + bp;
#pragma unroll
  for (int t = 0; t < kPairPlanes; ++t) {
    q[2 * F] = qn;
    const double dg = ak + c.c0;
    sl[me] = q[F];
#pragma unroll
    for (int m = 0; m < kHs; ++m) sl[hoff[m] * 8 + bp] = hv[m];
    if (t + 1 < kPairPlanes) load_next(t + 1);
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): LDS stores done
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (own && k0 + t < k1) {
      // Per band: centre, x terms, y terms, z terms, r ascending, each
      // term one fused multiply-add (stencil_kernel's order).
      double r0 = q[F].x * dg, r1 = q[F].y * dg;
#pragma unroll
      for (int r = 1; r <= F; ++r) {
        const hpair a = sl[me + r * 8], bb = sl[me - r * 8];
        r0 = fma(c.x[r], a.x + bb.x, r0);
        r1 = fma(c.x[r], a.y + bb.y, r1);
      }
#pragma unroll
      for (int r = 1; r <= F; ++r) {
        const hpair a = sl[me + r * E * 8], bb = sl[me - r * E * 8];
        r0 = fma(c.y[r], a.x + bb.x, r0);
        r1 = fma(c.y[r], a.y + bb.y, r1);
      }
#pragma unroll
      for (int r = 1; r <= F; ++r) {
        r0 = fma(c.z[r], q[F + r].x + q[F - r].x, r0);
        r1 = fma(c.z[r], q[F + r].y + q[F - r].y, r1);
      }
      const int sby = t * (int)plane * T * 8;
      if (two)
        __builtin_amdgcn_raw_buffer_store_b128(
            __builtin_bit_cast(su4, hpair{r0, r1}), ro, oby, sby, 0);
      else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(su2, r0),
                                              ro, oby, sby, 0);
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);  // this thread's LDS reads done
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
#pragma unroll
    for (int d = 0; d < 2 * F; ++d) q[d] = q[d + 1];
  }
}

int check_grid(int ni, int nj, int nk, int nb, int F) {
  if (ni <= 0 || nj <= 0 || nk <= 0 || nb < 0)
    return set_error(SKGEMM_ERR_ARG, "grid %d x %d x %d, %d bands", ni, nj, nk,
                     nb);
  if (F < 1 || F > kMaxFDn)
    return set_error(SKGEMM_ERR_ARG, "FDn %d outside 1..%d", F, kMaxFDn);
  const int64_t ex = (int64_t)(ni + 2 * F) * (nj + 2 * F) * (nk + 2 * F);
  const int64_t tiles = (nb + 15) / 16;
  // Launch grids: halo (nj + 2F, ceil((nk + 2F) / kHaloPlanes) x tiles);
  // stencil (blocks of the i-j plane, k chunks, tiles).
  const int64_t halo_y = (nk + 2 * F + kHaloPlanes - 1) / kHaloPlanes * tiles;
  const int64_t st_x = (int64_t)((ni + kSB - 1) / kSB) * ((nj + kSB - 1) / kSB);
  const int64_t plane = (int64_t)(ni + 2 * F) * (nj + 2 * F);
  if (ex * 16 >= ((int64_t)1 << 40) || plane >= 0x7fffffff ||
      halo_y > 65535 || tiles > 65535 ||
      st_x > 0x7fffffff ||
      (nk + kStencilPlanes - 1) / kStencilPlanes > 65535)
    return set_error(SKGEMM_ERR_NOTSUP, "grid too large for one launch");
  return SKGEMM_OK;
}

}  // namespace
}  // namespace skg

using namespace skg;

extern "C" int skgemm_halo_extend(skgemm_ctx* c, int ni, int nj, int nk,
                                  int nb, int FDn, const double* local16,
                                  double* ext16, int boundary) {
  SKG_TRY(check_ctx(c));
  SKG_TRY(check_grid(ni, nj, nk, nb, FDn));
  if (boundary != SKGEMM_HALO_ZERO && boundary != SKGEMM_HALO_PERIODIC &&
      boundary != SKGEMM_HALO_KEEP)
    return set_error(SKGEMM_ERR_ARG, "unknown halo boundary %d", boundary);
  if (nb == 0) return SKGEMM_OK;
  if (local16 == nullptr || ext16 == nullptr)
    return set_error(SKGEMM_ERR_ARG, "local16 or ext16 is NULL");
  SKG_TRY(check_dev(local16, "local16"));
  SKG_TRY(check_dev(ext16, "ext16"));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  const int nie = ni + 2 * FDn, nje = nj + 2 * FDn, nke = nk + 2 * FDn;
  const int64_t tiles = (nb + 15) / 16;
  const int64_t pairs = (int64_t)nje * nie * 8;  // per extended plane
  const int64_t chunks = (pairs + kHaloChunk - 1) / kHaloChunk;
  const bool vec = ((uintptr_t)local16 & 15) == 0 &&
                   ((uintptr_t)ext16 & 15) == 0 &&
                   pairs <= 0x7fffffff - 2 * kHaloChunk &&
                   tiles * nke * chunks < 0x7fffffff;
  c->record_begin();
  if (vec) {
    hipLaunchKernelGGL(halo_extend_vec_kernel,
                       dim3((unsigned)(tiles * nke * chunks)), dim3(256), 0,
                       c->stream, (const hpair*)local16, ni, nj, nk, FDn,
                       boundary, (int)chunks, (hpair*)ext16);
  } else {
    const dim3 grid(nje, ((nke + kHaloPlanes - 1) / kHaloPlanes) * tiles);
    hipLaunchKernelGGL(halo_extend_kernel, grid, dim3(256), 0, c->stream,
                       local16, ni, nj, nk, FDn, boundary, ext16);
  }
  SKG_HIP(hipGetLastError(), "halo launch");
  c->record_end();
  return SKGEMM_OK;
}

extern "C" int skgemm_stencil_apply(skgemm_ctx* c, int ni, int nj, int nk,
                                    int nb, int FDn, const double* ext16,
                                    const double* A_diag,
                                    const double* coef_x,
                                    const double* coef_y,
                                    const double* coef_z, double coef_0,
                                    double* out, int out_tile) {
  SKG_TRY(check_ctx(c));
  SKG_TRY(check_grid(ni, nj, nk, nb, FDn));
  if (out_tile != 16 && out_tile != 32)
    return set_error(SKGEMM_ERR_ARG, "out_tile must be 16 or 32");
  if (coef_x == nullptr || coef_y == nullptr || coef_z == nullptr)
    return set_error(SKGEMM_ERR_ARG, "coefficients are NULL");
  if (nb == 0) return SKGEMM_OK;
  if (ext16 == nullptr || A_diag == nullptr || out == nullptr)
    return set_error(SKGEMM_ERR_ARG, "ext16, A_diag or out is NULL");
  SKG_TRY(check_dev(ext16, "ext16"));
  SKG_TRY(check_dev(A_diag, "A_diag"));
  SKG_TRY(check_dev(out, "out"));
  StencilCoef cf{};
  for (int r = 0; r <= FDn; ++r) {
    cf.x[r] = coef_x[r];
    cf.y[r] = coef_y[r];
    cf.z[r] = coef_z[r];
  }
  cf.c0 = coef_0;
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  // The band-pair kernel's resources span kPairPlanes + 2 FDn extended
  // planes and kPairPlanes diagonal and output planes from a chunk's first
  // plane.
  const int64_t ep = (int64_t)(ni + 2 * FDn) * (nj + 2 * FDn);
  const int64_t gp = (int64_t)ni * nj;
  const bool pair = ((uintptr_t)ext16 & 15) == 0 &&
                    ((uintptr_t)out & 15) == 0 &&
                    (kPairPlanes + 2 * FDn) * ep * 128 < kStencilRecords &&
                    kPairPlanes * gp * 8 * out_tile < kStencilRecords;
  const int planes = pair ? kPairPlanes : kStencilPlanes;
  const int kchunks = (nk + planes - 1) / planes;
  const int64_t units = (int64_t)((ni + kSB - 1) / kSB) *
                        ((nj + kSB - 1) / kSB) * kchunks * ((nb + 15) / 16);
  if (units >= 0x7fffffff)
    return set_error(SKGEMM_ERR_NOTSUP, "stencil: too many workgroups");
  const dim3 grid((unsigned)units);
  const int ts = out_tile == 32 ? 5 : 4;
  c->record_begin();
#define SKG_STENCIL(F)                                                     \
  case F:                                                                  \
    if (pair)                                                              \
      hipLaunchKernelGGL(stencil_pair_kernel<F>, grid, dim3(512), 0,       \
                         c->stream, ext16, A_diag, ni, nj, nk, nb, cf,     \
                         out, ts);                                         \
    else                                                                   \
      hipLaunchKernelGGL(stencil_kernel<F>, grid, dim3(1024), 0,           \
                         c->stream, ext16, A_diag, ni, nj, nk, nb, cf, out, \
                         ts);                                              \
    break;
  switch (FDn) {
    SKG_STENCIL(1) SKG_STENCIL(2) SKG_STENCIL(3) SKG_STENCIL(4)
    SKG_STENCIL(5) SKG_STENCIL(6) SKG_STENCIL(7) SKG_STENCIL(8)
  }
#undef SKG_STENCIL
  SKG_HIP(hipGetLastError(), "stencil launch");
  c->record_end();
  return SKGEMM_OK;
}
