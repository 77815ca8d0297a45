// skgemm_stencil.hip - real-space finite-difference stencil H psi on the
// 16-band-packed extended grid (include/stencil_extended_packed.hpp) and the
// single-domain halo fill (include/halo_packed.hpp).
//
// Layouts: local grid [ceil(nb/16)][K][16], K = ni nj nk, g = i + j ni +
// k ni nj; extended grid [ceil(nb/16)][K_ex][16] with the interior at offset
// (F, F, F), K_ex = (ni+2F)(nj+2F)(nk+2F); output [ceil(nb/T)][K][T], T = 32
// or 16.  A grid point's 16 bands are 128 contiguous bytes, so one wave (16
// band lanes x 4 consecutive grid points along i) reads and writes 512-byte
// runs.  Both kernels are HBM-bound: the stencil's 6F + 1 reads per output
// come from L2 / L1 (neighbours along i are in the same or adjacent lines,
// along j and k in lines that the neighbouring workgroups of the same plane
// and of the next planes read too), so its algorithmic traffic is the
// extended input once, the diagonal once and the output once.
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

// One workgroup: 16 band lanes x 16 consecutive i of one j row of one
// 16-band tile, marching through kStencilPlanes planes of k; grid
// (ceil(ni/16), nj, tiles16 * ceil(nk / kStencilPlanes)).  Each lane keeps
// its own column's 2F + 1 values along k in registers, so a plane step loads
// one new z value and the 4F x / y neighbours (lines that the lanes and the
// neighbouring rows' workgroups of the same plane share in L1 / L2), not 6F.
constexpr int kStencilPlanes = 8;

template <int F>
__global__ __launch_bounds__(256) void stencil_kernel(
    const double* __restrict__ ext, const double* __restrict__ A_diag, int ni,
    int nj, int nk, int nb, StencilCoef c, double* __restrict__ out,
    int tshift) {
  const int bl = threadIdx.x & 15;
  const int i = blockIdx.x * 16 + (threadIdx.x >> 4);
  const int j = blockIdx.y;
  const int kchunks = (nk + kStencilPlanes - 1) / kStencilPlanes;
  const int t16 = blockIdx.z / kchunks;
  const int k0 = (blockIdx.z % kchunks) * kStencilPlanes;
  const int k1 = k0 + kStencilPlanes < nk ? k0 + kStencilPlanes : nk;
  const int b = t16 * 16 + bl;
  if (i >= ni || b >= nb) return;
  const int64_t K = (int64_t)ni * nj * nk;
  const int64_t sj = ni + 2 * F;
  const int64_t sk = sj * (nj + 2 * F);
  const int64_t Kex = sk * (nk + 2 * F);
  const double* s = ext + (int64_t)t16 * Kex * 16 + bl;
  const int T = 1 << tshift;
  double* o = out + ((int64_t)(b >> tshift) * K) * T + (b & (T - 1));
  // Column (i, j): extended index of plane k is col + (k + F) sk.
  const int64_t col = (i + F) + (j + F) * sj;
  double q[2 * F + 1];  // q[F + d] = psi at plane k + d
#pragma unroll
  for (int d = 0; d < 2 * F; ++d) q[d] = s[(col + (k0 + d) * sk) * 16];
  for (int k = k0; k < k1; ++k) {
    q[2 * F] = s[(col + (k + 2 * F) * sk) * 16];
    const int64_t gex = col + (k + F) * sk;
    const int64_t g = i + (int64_t)j * ni + (int64_t)k * ni * nj;
    // Same order as the reference's scalar version: centre, x terms, y
    // terms, z terms, r ascending in each (each term one fused
    // multiply-add).
    double res = q[F] * (A_diag[g] + c.c0);
#pragma unroll
    for (int r = 1; r <= F; ++r)
      res = fma(c.x[r], s[(gex + r) * 16] + s[(gex - r) * 16], res);
#pragma unroll
    for (int r = 1; r <= F; ++r)
      res = fma(c.y[r], s[(gex + r * sj) * 16] + s[(gex - r * sj) * 16], res);
#pragma unroll
    for (int r = 1; r <= F; ++r) res = fma(c.z[r], q[F + r] + q[F - r], res);
    o[g * T] = res;
#pragma unroll
    for (int d = 0; d < 2 * F; ++d) q[d] = q[d + 1];
  }
}

// Every extended point of every tile: the interior copies the local point,
// the halo the wrapped one (mode 1), zero (mode 0) or is left alone (mode
// 2).  Grid (ceil(ni_ex/16), nj_ex, nk_ex * tiles16), 16 band lanes x 16
// points along i.
__global__ __launch_bounds__(256) void halo_extend_kernel(
    const double* __restrict__ local, int ni, int nj, int nk, int F, int mode,
    double* __restrict__ ext) {
  const int bl = threadIdx.x & 15;
  const int nie = ni + 2 * F, nje = nj + 2 * F, nke = nk + 2 * F;
  const int ie = blockIdx.x * 16 + (threadIdx.x >> 4);
  const int je = blockIdx.y;
  const int ke = blockIdx.z % nke;
  const int t16 = blockIdx.z / nke;
  if (ie >= nie) return;
  const int64_t K = (int64_t)ni * nj * nk;
  const int64_t Kex = (int64_t)nie * nje * nke;
  const int i = ie - F, j = je - F, k = ke - F;
  const bool inside =
      i >= 0 && i < ni && j >= 0 && j < nj && k >= 0 && k < nk;
  if (!inside && mode == 2) return;
  double v = 0.0;
  if (inside || mode == 1) {
    const int ii = ((i % ni) + ni) % ni, jj = ((j % nj) + nj) % nj,
              kk = ((k % nk) + nk) % nk;
    const int64_t g = ii + (int64_t)jj * ni + (int64_t)kk * ni * nj;
    v = local[((int64_t)t16 * K + g) * 16 + bl];
  }
  const int64_t gex = ie + (int64_t)je * nie + (int64_t)ke * nie * nje;
  ext[((int64_t)t16 * Kex + gex) * 16 + bl] = v;
}

int check_grid(int ni, int nj, int nk, int nb, int F) {
  if (ni <= 0 || nj <= 0 || nk <= 0 || nb < 0)
    return set_error(SKGEMM_ERR_ARG, "grid %d x %d x %d, %d bands", ni, nj, nk,
                     nb);
  if (F < 1 || F > kMaxFDn)
    return set_error(SKGEMM_ERR_ARG, "FDn %d outside 1..%d", F, kMaxFDn);
  const int64_t ex = (int64_t)(ni + 2 * F) * (nj + 2 * F) * (nk + 2 * F);
  // Launch grids: y = nj + 2F rows, z = (nk + 2F) planes x 16-band tiles.
  if (ex * 16 >= ((int64_t)1 << 40) || nj + 2 * F > 65535 ||
      (int64_t)(nk + 2 * F) * ((nb + 15) / 16) > 65535)
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
  const dim3 grid((nie + 15) / 16, nje, nke * ((nb + 15) / 16));
  c->record_begin();
  hipLaunchKernelGGL(halo_extend_kernel, grid, dim3(256), 0, c->stream,
                     local16, ni, nj, nk, FDn, boundary, ext16);
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
  const int kchunks = (nk + kStencilPlanes - 1) / kStencilPlanes;
  const dim3 grid((ni + 15) / 16, nj, kchunks * ((nb + 15) / 16));
  const int ts = out_tile == 32 ? 5 : 4;
  c->record_begin();
#define SKG_STENCIL(F)                                                     \
  case F:                                                                  \
    hipLaunchKernelGGL(stencil_kernel<F>, grid, dim3(256), 0, c->stream,   \
                       ext16, A_diag, ni, nj, nk, nb, cf, out, ts);        \
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
