// skgemm_nloc.hip - non-local projector application on band-packed orbitals,
// result += V_nloc psi (include/nloc_packed_opt2.hpp:229-655).
//
// Per projector p (grid indices idx_p, nrow_p of them; chi_p nrow_p x
// ncol_p; gamma_p):  X_p = dv chi_p^T psi[idx_p, :],  X_p *= gamma_p (first
// projector of each atom only),  result[idx_p, :] += chi_p X_p.
// Device sequence, all on the context's stream, no host work:
//   1. forward parts: the projectors' rows are cut into chunks of at most
//      256; one workgroup per (chunk, 64 bands) gathers its rows of psi
//      straight from the 16-band panels (32 rows per wave, all loads in
//      flight together) and forms the chunk's part of chi_p^T psi[idx_p, :];
//   2. X_p = (dv sum of the chunk parts, in chunk order) gamma_p (times 1.0,
//      exact, where the atom rule does not apply), one thread per (column,
//      band);
//   3. backward products and scatter together: the touched grid points,
//      ascending, are cut at create time into blocks of 32; one wave per
//      (block, 64 bands) holds the block's result values in LDS and adds
//      each row's sum_c chi_p[r][c] X_p[c][b] to its point's value, the rows
//      in projector order, projector by projector with X_p in registers.
//      Shared grid points need no atomics, the sums keep the reference's
//      order and no per-row products are stored.
// HBM-bound: psi is read once per projector row (repeats of overlapping
// projectors mostly hit L2), result read and written once per touched
// point; the forward parts and X (columns x bands) are small.
// (Before: the backward kernel stored every row's products, 8 x rows x nb
// bytes, for a scatter kernel to add them up: 92 + 109 us on 32768 points x
// 520 bands with 64 projectors.  A first fused form, one wave per point
// reading X from LDS row by row, took 0.64 ms: each point's result load,
// row entries, chi loads and FMA chain ran one after another.)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/skgemm.h"
#include "skgemm_ctx.h"

namespace skg {
namespace {

struct NlocProj {
  int64_t roff, choff, coff;
  int nrow, ncol;
  int chunk0, nchunks;  // this projector's forward chunks
  int64_t poff;         // its first row of forward parts (chunk-major)
};
struct NlocChunk {
  int proj, r0, nr;
  int64_t f0;  // its chi values in the forward A layout, in blocks of 4096
};
typedef double v4d __attribute__((ext_vector_type(4)));

// Unit u (of `units`) for hardware workgroup g of a one-dimensional grid:
// workgroup g runs on XCD g % 8, so each XCD gets a contiguous range of
// units (neighbouring units share data through that XCD's L2).
__device__ inline int xcd_unit(int g, int units) {
  const int per = units >> 3, rem = units & 7, xcd = g & 7;
  return xcd * per + (xcd < rem ? xcd : rem) + (g >> 3);
}
// A segment of a point block's rows: rows e0 .. e1 of the block-ordered row
// list, all of one projector (X columns coff .., ncol of them), touching
// distinct points of the block; its chi rows are stored from a0 on, in
// units of 256 doubles (one 16-row block and 16-column pass: row block m,
// pass q at a0 + m ceil(ncol / 16) + q).
struct NlocSeg {
  int e0, e1, coff, ncol;
  int64_t a0;
};
// A block of touched points u0 .. u0 + np - 1 (ascending grid index) and its
// segments s0 .. s1 - 1.
struct NlocBlk {
  int u0, np, s0, s1;
};
constexpr int kBlkPoints = 32;  // points per block of the apply kernel
constexpr int kApplyNJ = 4;     // its waves cover 16 kApplyNJ bands
constexpr int kFwdWaves = 8, kFwdRowsPerWave = 32;
constexpr int kFwdRows = kFwdWaves * kFwdRowsPerWave;  // rows per chunk

// Forward parts: part[poff_p + (k - chunk0_p) ncol_p + c][b] =
//   sum over the chunk's rows r of chi_p[r][c] psi(index[roff_p + r], b).
// Wave w takes rows 32 w .. 32 w + 31 of the chunk as eight steps of
// v_mfma_f64_16x16x4 per 16-band block (A = 16 chi columns x 4 rows, B = 4
// psi rows x 16 bands): lane l gathers psi at row 4 s + l / 16, bands
// 4 (l % 16) + j of the workgroup's 64 (four adjacent values of one 16-band
// panel; 32 loads issued at once, on clamped rows) and reads its chi values
// from a copy the host stores in
// the A layout (zero past the chunk's rows and the projector's columns, so
// the clamped rows add nothing).  The 8 waves' sums go through LDS and are
// added in wave order.  (A non-finite psi value at a chunk's last row can
// turn an infinite part into NaN: the clamped copies add 0 x inf.)
// (Before: FMAs with each row's 16 chi values read from LDS as broadcasts,
// 1 KB returned to the wave per 16-byte broadcast, which bound the kernel by
// LDS bandwidth: 76 us on 49665 rows x 520 bands.)
__global__ __launch_bounds__(kFwdWaves * 64, 4) void nloc_forward_kernel(
    const double* __restrict__ psi16, int Nd, int nb,
    const int* __restrict__ index, const double* __restrict__ chif,
    const NlocProj* __restrict__ proj, const NlocChunk* __restrict__ chunks,
    double* __restrict__ part) {
  __shared__ __align__(16) double red[kFwdWaves * 16 * 64];  // each wave's 16 x 64 sums
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // units chunk-major, the band groups of one chunk on one XCD
  const int u = xcd_unit(blockIdx.x, gridDim.x);
  const int bgroups = (nb + 63) >> 6;
  const int ci = u / bgroups;
  const int b0 = (u - ci * bgroups) * 64;
  const NlocChunk C = chunks[ci];
  const NlocProj P = proj[C.proj];
  const int* idx = index + P.roff;
  const int rlast = C.r0 + C.nr - 1;
  // lane l's bands: 4 (l % 16) + j, j = 0..3, four adjacent values of one
  // 16-band panel (a band group past nb reads the last panel's group: its
  // values only reach outputs that are not stored)
  const int bq = min(b0 + 4 * (lane & 15), ((nb + 15) & ~15) - 4);
  const double* src = psi16 + (int64_t)(bq >> 4) * Nd * 16 + (bq & 15);
  double v[8][4];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int g = idx[min(C.r0 + 32 * w + 4 * s + (lane >> 4), rlast)];
    const double2* pv = reinterpret_cast<const double2*>(src + (int64_t)g * 16);
    const double2 x0 = pv[0], x1 = pv[1];
    v[s][0] = x0.x;
    v[s][1] = x0.y;
    v[s][2] = x1.x;
    v[s][3] = x1.y;
  }
  double* dst = part + (P.poff + (int64_t)(ci - P.chunk0) * P.ncol) * nb;
  const int kp = (P.ncol + 15) >> 4;
  for (int q = 0; q < kp; ++q) {
    const double2* cf = reinterpret_cast<const double2*>(
        chif + (C.f0 + q) * (kFwdRows * 16) + (w * 64 + lane) * 8);
    double a[8];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const double2 x = cf[h];
      a[2 * h] = x.x;
      a[2 * h + 1] = x.y;
    }
    v4d d[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) d[j] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        d[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], v[s][j], d[j], 0, 0,
                                                    0);
    if (q > 0) __syncthreads();  // the previous pass's sums are read
    // lane l holds column 4 u + l / 16, band 4 (l % 16) + j
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double2* rp = reinterpret_cast<double2*>(
          red + (w * 16 + 4 * u + (lane >> 4)) * 64 + 4 * (lane & 15));
      rp[0] = make_double2(d[0][u], d[1][u]);
      rp[1] = make_double2(d[2][u], d[3][u]);
    }
    __syncthreads();
    // wave w adds up columns 2 w and 2 w + 1 over the waves, in wave order
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int c = 2 * w + h;
      double sum = red[c * 64 + lane];
#pragma unroll
      for (int x = 1; x < kFwdWaves; ++x) sum += red[(x * 16 + c) * 64 + lane];
      if (16 * q + c < P.ncol && b0 + lane < nb)
        dst[(int64_t)(16 * q + c) * nb + b0 + lane] = sum;
    }
  }
}

// X_p[c][b] = (dv sum_j part[poff_p + j ncol_p + c][b]) factor_p[c], the
// chunk parts added in chunk order; one thread per (column, band), four
// columns per workgroup.  X rows are xs = 64 ceil(nb / 64) long, zero past
// nb.
__global__ __launch_bounds__(256) void nloc_x_kernel(
    const double* __restrict__ part, const NlocProj* __restrict__ proj,
    const int* __restrict__ colproj, const double* __restrict__ factor,
    double dv, int nb, int xs, int64_t cols, double* __restrict__ X) {
  const int64_t col = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int b = blockIdx.y * 64 + (threadIdx.x & 63);
  if (col >= cols) return;
  double x = 0.0;
  if (b < nb) {
    const NlocProj P = proj[colproj[col]];
    const int c = (int)(col - P.coff);
    const double* src = part + (P.poff + c) * nb + b;
    for (int j = 0; j < P.nchunks; ++j) x += src[(int64_t)j * P.ncol * nb];
    x = (dv * x) * factor[col];
  }
  X[col * xs + b] = x;
}

// A wave of the apply kernel covers W = 16 NJ bands: lane l's bands are
// NJ (l % 16) + j, j < NJ (NJ adjacent values).

// Loads NJ adjacent doubles (16-byte pieces where NJ >= 2; p is 8 NJ-byte
// aligned).
template <int NJ>
__device__ inline void load_bands(const double* __restrict__ p, double* v) {
  if constexpr (NJ == 1) {
    v[0] = p[0];
  } else {
#pragma unroll
    for (int h = 0; h < NJ / 2; ++h) {
      const double2 x = reinterpret_cast<const double2*>(p)[h];
      v[2 * h] = x.x;
      v[2 * h + 1] = x.y;
    }
  }
}

// One segment's operands in registers, loaded ahead of its turn: its chi
// rows in the A layout of v_mfma_f64_16x16x4 (lane l: rows 16 m + l % 16,
// columns 4 s + l / 16, s = 0..3, for each block m of 16 rows; the host
// stores them so, zero past the segment's rows and the projector's
// columns), its first 16 X columns in the B layout (lane l: column
// 4 s + l / 16, bands NJ (l % 16) + j of the wave's W), and in lane k the
// block slot of the segment's row k.
template <int NP, int NJ>
struct NlocSegOps {
  NlocSeg S;
  double a[NP / 16][4];
  double x[4][NJ];
  int es;
};

template <int NP, int NJ>
__device__ inline void nloc_segment_loads(const NlocSeg& S,
                                          const double* __restrict__ chia,
                                          const int* __restrict__ eslot,
                                          const double* __restrict__ X, int xs,
                                          int b0, int lane,
                                          NlocSegOps<NP, NJ>& o) {
  o.S = S;
  const int n = S.e1 - S.e0;
  const int kp = (S.ncol + 15) >> 4;  // 16-column passes
  const int mbs = (n + 15) >> 4;      // 16-row blocks
#pragma unroll
  for (int m = 0; m < NP / 16; ++m) {
    const double2* src = reinterpret_cast<const double2*>(
        chia + (S.a0 + (int64_t)min(m, mbs - 1) * kp) * 256 + lane * 4);
    const double2 u = src[0], v = src[1];
    o.a[m][0] = u.x;
    o.a[m][1] = u.y;
    o.a[m][2] = v.x;
    o.a[m][3] = v.y;
  }
  o.es = eslot[S.e0 + min(lane, n - 1)];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int c = 4 * s + (lane >> 4);
    double v[NJ];
    load_bands<NJ>(X + (int64_t)(S.coff + min(c, S.ncol - 1)) * xs + b0 +
                       NJ * (lane & 15),
                   v);
#pragma unroll
    for (int j = 0; j < NJ; ++j) o.x[s][j] = c < S.ncol ? v[j] : 0.0;
  }
}

// Adds one segment's row products T = chi_seg X_p (rows x W bands) to the
// block's values in LDS ([point][W]).  Each 16-row block is 4 NJ MFMAs (4
// column steps x NJ band blocks; columns past 16 in further passes, loaded
// there); lane l then holds T[16 m + 4 v + l / 16][NJ (l % 16) + j] and adds
// it to its row's point.  The rows' points are distinct, so every (point,
// band) value is read and written once per segment.
template <int NP, int NJ>
__device__ inline void nloc_segment_apply(const NlocSegOps<NP, NJ>& o,
                                          double* __restrict__ acc,
                                          const double* __restrict__ chia,
                                          const double* __restrict__ X, int xs,
                                          int b0, int lane) {
  constexpr int W = 16 * NJ;
  const NlocSeg& S = o.S;
  const int n = S.e1 - S.e0;
  const int kp = (S.ncol + 15) >> 4;
#pragma unroll
  for (int m = 0; m < NP / 16; ++m) {
    if (16 * m >= n) break;
    v4d d[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) d[j] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int j = 0; j < NJ; ++j)
        d[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.a[m][s], o.x[s][j], d[j],
                                                    0, 0, 0);
    // Further 16-column passes.  Loads are unconditional (the column index
    // is clamped, so every address is valid; columns past ncol are zeroed
    // by a select), so none waits alone in a branch next to its use.
    for (int p = 1; p < kp; ++p) {
      const double* ap = chia + (S.a0 + (int64_t)m * kp + p) * 256 + lane * 4;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int c = 16 * p + 4 * s + (lane >> 4);
        const double* xc = X + (int64_t)(S.coff + min(c, S.ncol - 1)) * xs +
                           b0 + NJ * (lane & 15);
        const double av = ap[s];
        double xv[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const double x = xc[j];
          xv[j] = c < S.ncol ? x : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          d[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xv[j], d[j], 0, 0,
                                                      0);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int e = 16 * m + 4 * v + (lane >> 4);
      const int sl = __shfl(o.es, min(e, n - 1));
      double* ap = acc + sl * W + NJ * (lane & 15);
      if (e < n) {
        if constexpr (NJ == 1) {
          ap[0] += d[0][v];
        } else {
#pragma unroll
          for (int h = 0; h < NJ / 2; ++h) {
            double2* a2 = reinterpret_cast<double2*>(ap) + h;
            const double2 t = *a2;
            *a2 = make_double2(t.x + d[2 * h][v], t.y + d[2 * h + 1][v]);
          }
        }
      }
    }
  }
  __syncthreads();  // the next segment's lanes read what these wrote
}

// result[g][b] += sum over the rows (p, r) touching grid point g, in
// projector order (then row order), of sum_c chi_p[r][c] X_p[c][b].  The
// touched points, ascending, are cut into blocks of NP; a block's rows,
// ordered by (projector, row), are cut into segments of one projector whose
// rows touch distinct points of the block.  One wave per (block, W = 16 NJ
// bands): the block's result values go to LDS ([point][W]) and the
// segments are applied in order, each one's operands loaded one segment
// ahead (two register sets in turn, the loop unrolled by two so that no set
// is copied: a copy would wait for its loads).  A point's additions all
// happen on one wave in projector order, so the sums keep the reference's
// order of projectors with no atomics, and no per-row products are stored.
// Units are band-group-major (the result streams through one band range at
// a time) with a run of neighbouring blocks, which share projectors, on
// each XCD.
// Measured on 32768 points x 520 bands with 64 projectors, W = 64: the row
// products as FMAs on chi values read from LDS as broadcasts (16 per row,
// each returning 1 KB to the wave) were bound by LDS bandwidth at 102 us;
// with MFMA, three register sets two segments ahead took 248 registers (two
// waves per SIMD) and 75.5 us against 71.2 us for two sets (three waves per
// SIMD); blocks of 16 points took 89 us.
template <int NP, int NJ>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 3)))
void nloc_apply_kernel(
    const double* __restrict__ X, int nb, const double* __restrict__ chia,
    const int* __restrict__ eslot, const NlocSeg* __restrict__ seg,
    const NlocBlk* __restrict__ blk, const int* __restrict__ point, int Nd,
    int tshift, double* __restrict__ result) {
  static_assert(NP % 16 == 0 && NP <= 64, "block points");
  static_assert(NJ == 1 || NJ == 2 || NJ == 4, "band blocks per wave");
  constexpr int W = 16 * NJ;
  constexpr int kPer = NP * W / 64;  // result values per lane
  __shared__ __align__(16) double acc[NP * W];
  const int lane = threadIdx.x;
  const int u = xcd_unit(blockIdx.x, gridDim.x);
  const int nblk = gridDim.x / ((nb + W - 1) / W);
  const int bg = u / nblk;
  const int bi = u - bg * nblk;
  const int b0 = bg * W;
  const int xs = (nb + 63) & ~63;  // X row length
  const int T = 1 << tshift;
  const NlocBlk B = blk[bi];
  const int* pts = point + B.u0;  // NP entries (the table is padded)
  // value i of lane l: point (64 i + l) / W, band b0 + l % W
  const int bw = lane % W;
  const int bc = min(b0 + bw, nb - 1);
  const int64_t boff = (int64_t)(bc >> tshift) * Nd * T + (bc & (T - 1));
  // the grid point of value i (uniform when W = 64: scalar loads)
  int gp[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i)
    gp[i] = W == 64 ? pts[i] : pts[(64 * i + lane) / W];
  double r[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) r[i] = result[boff + (int64_t)gp[i] * T];
  NlocSegOps<NP, NJ> o0, o1;
  nloc_segment_loads<NP, NJ>(seg[B.s0], chia, eslot, X, xs, b0, lane, o0);
#pragma unroll
  for (int i = 0; i < kPer; ++i) acc[64 * i + lane] = r[i];
  __syncthreads();
  for (int si = B.s0;; si += 2) {
    nloc_segment_loads<NP, NJ>(seg[min(si + 1, B.s1 - 1)], chia, eslot, X, xs,
                               b0, lane, o1);
    nloc_segment_apply<NP, NJ>(o0, acc, chia, X, xs, b0, lane);
    if (si + 1 >= B.s1) break;
    nloc_segment_loads<NP, NJ>(seg[min(si + 2, B.s1 - 1)], chia, eslot, X, xs,
                               b0, lane, o0);
    nloc_segment_apply<NP, NJ>(o1, acc, chia, X, xs, b0, lane);
    if (si + 2 >= B.s1) break;
  }
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int sp = W == 64 ? i : (64 * i + lane) / W;
    const int g = W == 64 ? pts[i] : gp[i];  // W = 64: a scalar reload
    if (sp < B.np && b0 + bw < nb)
      result[boff + (int64_t)g * T] = acc[64 * i + lane];
  }
}

}  // namespace
}  // namespace skg

using namespace skg;

struct skgemm_nloc {
  skgemm_ctx* ctx = nullptr;
  int Nd = 0, num_proj = 0;
  int64_t rows = 0, cols = 0, nchunks = 0, parts = 0;
  // device copies
  int* index = nullptr;       // rows
  double* chif = nullptr;     // chunks' chi values in the forward A layout
  double* factor = nullptr;   // cols
  NlocProj* proj = nullptr;   // num_proj
  NlocChunk* chunks = nullptr;  // nchunks
  int* colproj = nullptr;     // cols: the projector of each X column
  int* point = nullptr;       // npoints, then kBlkPoints copies of the last
  NlocBlk* blk = nullptr;     // nblocks
  NlocSeg* seg = nullptr;     // segments of all blocks
  int* eslot = nullptr;       // rows in block order: their block slot
  double* chia = nullptr;     // segments' chi rows in the MFMA A layout
  int64_t npoints = 0, nblocks = 0;
  // workspace, grown with nb
  double* part = nullptr;     // parts x nb
  double* X = nullptr;        // cols x nb
  int ws_nb = 0;
};

static void nloc_free(skgemm_nloc* h) {
  for (void* p : {(void*)h->index, (void*)h->chif, (void*)h->factor,
                  (void*)h->proj, (void*)h->chunks, (void*)h->colproj,
                  (void*)h->point, (void*)h->blk, (void*)h->seg,
                  (void*)h->eslot, (void*)h->chia, (void*)h->part, (void*)h->X})
    if (p) hipFree(p);
}

template <typename T>
static int upload(T** dst, const T* src, size_t n, const char* what,
                  hipStream_t stream) {
  if (n == 0) return SKGEMM_OK;
  if (hipMalloc(dst, n * sizeof(T)) != hipSuccess)
    return set_error(SKGEMM_ERR_NOMEM, "nloc %s (%zu bytes)", what,
                     n * sizeof(T));
  if (hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice,
                     stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return set_error(SKGEMM_ERR_HIP, "nloc %s upload", what);
  return SKGEMM_OK;
}

extern "C" int skgemm_nloc_create(skgemm_ctx* c, int Nd, int num_proj,
                                  const int* nrow, const int* ncol,
                                  const int* atom_index, const int* index_all,
                                  const double* chi_all,
                                  const double* gamma_all, skgemm_nloc** out) {
  SKG_TRY(check_ctx(c));
  if (out == nullptr) return set_error(SKGEMM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (Nd <= 0 || num_proj < 0)
    return set_error(SKGEMM_ERR_ARG, "nloc: Nd > 0 and num_proj >= 0 needed");
  if (num_proj > 0 && (nrow == nullptr || ncol == nullptr ||
                       atom_index == nullptr))
    return set_error(SKGEMM_ERR_ARG, "nloc: nrow, ncol, atom_index needed");
  std::vector<int64_t> roff(num_proj + 1, 0), coff(num_proj + 1, 0),
      choff(num_proj + 1, 0);
  for (int p = 0; p < num_proj; ++p) {
    if (nrow[p] < 0 || ncol[p] < 0)
      return set_error(SKGEMM_ERR_ARG, "nloc: projector %d has a negative size",
                       p);
    roff[p + 1] = roff[p] + nrow[p];
    coff[p + 1] = coff[p] + ncol[p];
    choff[p + 1] = choff[p] + (int64_t)nrow[p] * ncol[p];
  }
  const int64_t rows = roff[num_proj], cols = coff[num_proj];
  if (rows >= INT32_MAX || cols >= INT32_MAX)
    return set_error(SKGEMM_ERR_NOTSUP, "nloc: more than 2^31 projector rows");
  if ((rows > 0 && index_all == nullptr) ||
      (choff[num_proj] > 0 && chi_all == nullptr) ||
      (cols > 0 && gamma_all == nullptr))
    return set_error(SKGEMM_ERR_ARG, "nloc: index, chi or gamma is NULL");
  for (int64_t R = 0; R < rows; ++R)
    if (index_all[R] < 0 || index_all[R] >= Nd)
      return set_error(SKGEMM_ERR_ARG, "nloc: grid index %d out of [0, %d)",
                       index_all[R], Nd);
  auto* h = new skgemm_nloc();
  h->ctx = c;
  h->Nd = Nd;
  h->num_proj = num_proj;
  h->rows = rows;
  h->cols = cols;
  // Row factors: gamma for the first projector of each atom (the reference
  // starts from the first projector's atom_index - 1), 1.0 elsewhere.
  std::vector<double> factor((size_t)cols, 1.0);
  if (num_proj > 0) {
    int atom = atom_index[0] - 1;
    for (int p = 0; p < num_proj; ++p) {
      if (atom == atom_index[p]) continue;
      atom = atom_index[p];
      for (int j = 0; j < ncol[p]; ++j)
        factor[coff[p] + j] = gamma_all[coff[p] + j];
    }
  }
  // Projector table and forward chunks (projectors without columns or rows
  // get none: they contribute nothing).
  std::vector<NlocProj> meta((size_t)num_proj);
  std::vector<NlocChunk> chunks;
  // Each chunk's chi values in the forward kernel's A layout: per 16-column
  // pass q a block of 4096 doubles, wave w's lane l holding at
  // 8 (64 w + l) + s the value of column 16 q + l % 16, chunk row
  // 32 w + 4 s + l / 16 (zero past the chunk's rows and the columns).
  std::vector<double> chif;
  int64_t poff = 0;
  for (int p = 0; p < num_proj; ++p) {
    NlocProj m{roff[p], choff[p], coff[p], nrow[p], ncol[p], 0, 0, poff};
    m.chunk0 = (int)chunks.size();
    const int kp = (ncol[p] + 15) / 16;
    if (ncol[p] > 0)
      for (int r0 = 0; r0 < nrow[p]; r0 += kFwdRows) {
        const int nr = std::min(kFwdRows, nrow[p] - r0);
        chunks.push_back(
            NlocChunk{p, r0, nr, (int64_t)(chif.size() / (kFwdRows * 16))});
        for (int q = 0; q < kp; ++q)
          for (int i = 0; i < kFwdRows * 16; ++i) {
            const int s = i % 8, l = (i / 8) % 64, w = i / 512;
            const int row = 32 * w + 4 * s + l / 16, col = 16 * q + l % 16;
            chif.push_back(row < nr && col < ncol[p]
                               ? chi_all[choff[p] + (int64_t)col * nrow[p] +
                                         r0 + row]
                               : 0.0);
          }
      }
    m.nchunks = (int)chunks.size() - m.chunk0;
    poff += (int64_t)m.nchunks * ncol[p];
    meta[p] = m;
  }
  h->nchunks = (int64_t)chunks.size();
  h->parts = poff;
  // Point blocks: the touched grid points in ascending order (projectors
  // without columns skipped), kBlkPoints to a block; a block's rows sorted
  // by global row, i.e. by (projector, row), and cut into segments at each
  // new projector and before a row whose point the segment already has.
  std::vector<std::pair<int, int>> touch;  // (grid point, global row)
  touch.reserve((size_t)rows);
  std::vector<int> rowproj((size_t)rows);
  for (int p = 0; p < num_proj; ++p) {
    for (int r = 0; r < nrow[p]; ++r) rowproj[roff[p] + r] = p;
    if (ncol[p] == 0) continue;
    for (int r = 0; r < nrow[p]; ++r)
      touch.push_back({index_all[roff[p] + r], (int)(roff[p] + r)});
  }
  std::sort(touch.begin(), touch.end());
  std::vector<int> point, colproj((size_t)cols);
  std::vector<NlocBlk> blks;
  std::vector<NlocSeg> segs;
  std::vector<int> eslot;   // the block slot of each row, in block order
  std::vector<int> erow;    // its global row (host only)
  std::vector<double> chia;  // chi rows of each segment in the MFMA A layout
  eslot.reserve(touch.size());
  erow.reserve(touch.size());
  for (int p = 0; p < num_proj; ++p)
    for (int j = 0; j < ncol[p]; ++j) colproj[coff[p] + j] = p;
  std::vector<std::pair<int, int>> brow;  // (global row, slot)
  for (size_t i = 0; i < touch.size();) {
    NlocBlk B{(int)point.size(), 0, (int)segs.size(), 0};
    brow.clear();
    while (i < touch.size() && B.np < kBlkPoints) {
      const int g = touch[i].first;
      for (; i < touch.size() && touch[i].first == g; ++i)
        brow.push_back({touch[i].second, B.np});
      point.push_back(g);
      ++B.np;
    }
    std::sort(brow.begin(), brow.end());
    uint64_t seen = 0;
    for (size_t e = 0; e < brow.size(); ++e) {
      const int R = brow[e].first, s = brow[e].second, p = rowproj[R];
      if (e == 0 || p != rowproj[brow[e - 1].first] || ((seen >> s) & 1)) {
        if (e > 0) segs.back().e1 = (int)eslot.size();
        segs.push_back(
            NlocSeg{(int)eslot.size(), 0, (int)coff[p], ncol[p], 0});
        seen = 0;
      }
      seen |= 1ull << s;
      eslot.push_back(s);
      erow.push_back(R);
    }
    segs.back().e1 = (int)eslot.size();
    B.s1 = (int)segs.size();
    blks.push_back(B);
  }
  // Each segment's chi rows in the A-operand layout of v_mfma_f64_16x16x4:
  // per 16-row block m and 16-column pass q, 256 doubles, lane l's four
  // values (s = 0..3) at 4 l + s: row 16 m + l % 16 of the segment, column
  // 16 q + 4 s + l / 16 (zero past the rows and columns).
  for (NlocSeg& S : segs) {
    const int n = S.e1 - S.e0, kp = (S.ncol + 15) / 16, p = colproj[S.coff];
    S.a0 = (int64_t)(chia.size() / 256);
    for (int m = 0; m < (n + 15) / 16; ++m)
      for (int q = 0; q < kp; ++q)
        for (int l = 0; l < 64; ++l)
          for (int s4 = 0; s4 < 4; ++s4) {
            const int e = 16 * m + l % 16, c = 16 * q + 4 * s4 + l / 16;
            chia.push_back(e < n && c < S.ncol
                               ? chi_all[choff[p] + (int64_t)c * nrow[p] +
                                         (erow[S.e0 + e] - roff[p])]
                               : 0.0);
          }
  }
  h->npoints = (int64_t)point.size();
  h->nblocks = (int64_t)blks.size();
  if (!point.empty()) point.resize(point.size() + kBlkPoints, point.back());
  int rc = hipSetDevice(c->device) == hipSuccess
               ? SKGEMM_OK
               : set_error(SKGEMM_ERR_HIP, "hipSetDevice");
  if (rc == SKGEMM_OK)
    rc = upload(&h->index, index_all, (size_t)rows, "index", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->chif, chif.data(), chif.size(), "chi (forward)",
                c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->factor, factor.data(), factor.size(), "gamma", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->proj, meta.data(), meta.size(), "projector table",
                c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->chunks, chunks.data(), chunks.size(), "chunk table",
                c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->colproj, colproj.data(), colproj.size(), "column table",
                c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->point, point.data(), point.size(), "points", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->blk, blks.data(), blks.size(), "block table", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->seg, segs.data(), segs.size(), "segment table", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->eslot, eslot.data(), eslot.size(), "row slots",
                c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->chia, chia.data(), chia.size(), "chi blocks", c->stream);
  if (rc != SKGEMM_OK) {
    nloc_free(h);
    delete h;
    return rc;
  }
  *out = h;
  return SKGEMM_OK;
}

extern "C" int skgemm_nloc_apply(skgemm_nloc* h, int nb, double dv,
                                 const double* psi16, double* result,
                                 int result_tile) {
  if (h == nullptr) return set_error(SKGEMM_ERR_ARG, "nloc handle is NULL");
  skgemm_ctx* c = h->ctx;
  SKG_TRY(check_ctx(c));
  if (nb < 0) return set_error(SKGEMM_ERR_ARG, "nb < 0");
  if (result_tile != 16 && result_tile != 32)
    return set_error(SKGEMM_ERR_ARG, "result_tile must be 16 or 32");
  if (nb == 0 || h->rows == 0 || h->cols == 0 || h->npoints == 0)
    return SKGEMM_OK;
  if (psi16 == nullptr || result == nullptr)
    return set_error(SKGEMM_ERR_ARG, "psi16 or result is NULL");
  SKG_TRY(check_dev(psi16, "psi16"));
  if ((uintptr_t)psi16 % 16 != 0)
    return set_error(SKGEMM_ERR_ARG, "psi16 must be 16-byte aligned");
  SKG_TRY(check_dev(result, "result"));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  if (nb > h->ws_nb) {
    SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (double** p : {&h->part, &h->X}) {
      if (*p) hipFree(*p);
      *p = nullptr;
    }
    h->ws_nb = 0;
    if (hipMalloc(&h->part, (size_t)std::max<int64_t>(h->parts, 1) * nb *
                                sizeof(double)) != hipSuccess ||
        hipMalloc(&h->X, (size_t)h->cols * ((nb + 63) & ~63) *
                             sizeof(double)) != hipSuccess)
      return set_error(SKGEMM_ERR_NOMEM, "nloc workspace for %d bands", nb);
    h->ws_nb = nb;
    ++c->mem_gen;
  }
  const unsigned bgroups = (unsigned)((nb + 63) / 64);
  if (h->nchunks == 0) return SKGEMM_OK;  // no projector has columns
  const int64_t agroups = (nb + 16 * kApplyNJ - 1) / (16 * kApplyNJ);
  if (h->nchunks * bgroups >= INT32_MAX || h->nblocks * agroups >= INT32_MAX)
    return set_error(SKGEMM_ERR_NOTSUP, "nloc: too many workgroups");
  hipLaunchKernelGGL(nloc_forward_kernel,
                     dim3((unsigned)(h->nchunks * bgroups)),
                     dim3(kFwdWaves * 64), 0,
                     c->stream, psi16, h->Nd, nb, h->index, h->chif, h->proj,
                     h->chunks, h->part);
  SKG_HIP(hipGetLastError(), "nloc forward launch");
  hipLaunchKernelGGL(nloc_x_kernel, dim3((unsigned)((h->cols + 3) / 4), bgroups),
                     dim3(256), 0, c->stream, h->part, h->proj, h->colproj,
                     h->factor, dv, nb, (int)(bgroups * 64), h->cols, h->X);
  SKG_HIP(hipGetLastError(), "nloc X launch");
  hipLaunchKernelGGL((nloc_apply_kernel<kBlkPoints, kApplyNJ>),
                     dim3((unsigned)(h->nblocks * agroups)), dim3(64), 0,
                     c->stream, h->X, nb, h->chia, h->eslot, h->seg, h->blk,
                     h->point, h->Nd, result_tile == 32 ? 5 : 4, result);
  SKG_HIP(hipGetLastError(), "nloc apply launch");
  return SKGEMM_OK;
}

extern "C" int skgemm_nloc_destroy(skgemm_nloc* h) {
  if (h == nullptr) return SKGEMM_OK;
  if (h->ctx) {
    hipSetDevice(h->ctx->device);
    hipStreamSynchronize(h->ctx->stream);
  }
  nloc_free(h);
  delete h;
  return SKGEMM_OK;
}
