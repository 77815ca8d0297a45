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
//   2. backward products: one workgroup per (chunk, 64 bands) forms X_p for
//      its bands, (dv sum of the chunk parts, in chunk order) gamma_p (times
//      1.0, exact, where the atom rule does not apply), in registers, and
//      writes the chunk's rows of temp_p = chi_p X_p;
//   3. scatter: one wave per (touched grid point g, 64 bands) adds the temp
//      rows of g in projector order (a CSR table built at create time), so
//      shared grid points need no atomics and the sums keep the reference's
//      order.
// HBM-bound: psi is read once per projector row (repeats of overlapping
// projectors mostly hit L2), temp written and read once, result read and
// written once per touched point.  (Forming the products inside the
// scatter instead reads X_p, ncol values per band, for every touched
// (point, row), 13x the bytes of one temp row at 13 columns: on 32768
// points x 520 bands with 64 projectors that scatter alone took 305 us,
// against 285 us for this whole sequence.)

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
};
constexpr int kFwdWaves = 8, kFwdRowsPerWave = 32;
constexpr int kFwdRows = kFwdWaves * kFwdRowsPerWave;  // rows per chunk
constexpr int kFwdCols = 16;                            // columns per pass

// Forward parts: part[poff_p + (k - chunk0_p) ncol_p + c][b] =
//   sum over the chunk's rows r of chi_p[r][c] psi(index[roff_p + r], b).
// Wave w takes rows w * 32 .. w * 32 + 31 of the chunk and issues its 32
// psi loads at once, on clamped addresses (a band past nb reads band
// nb - 1, a row past the chunk its last row), before the chunk's chi
// values are staged in LDS ([column][row], zero past the chunk's rows and
// the projector's columns, so the clamped rows add nothing); chi values
// are read as broadcasts.  The 8 waves' sums are added in wave order
// through LDS, 8 columns at a time.  chi_p is column-major nrow x ncol.
// (A non-finite psi value at a chunk's last row can turn an infinite part
// into NaN: the clamped copies add 0 x inf.)
__global__ __launch_bounds__(kFwdWaves * 64, 4) void nloc_forward_kernel(
    const double* __restrict__ psi16, int Nd, int nb,
    const int* __restrict__ index, const double* __restrict__ chi,
    const NlocProj* __restrict__ proj, const NlocChunk* __restrict__ chunks,
    double* __restrict__ part) {
  // chi of the chunk ([kFwdRows][kFwdCols], a row's 16 values read as
  // eight 16-byte broadcasts), then the wave sums ([kFwdWaves][8][64]):
  // 32 KB either way.
  __shared__ __align__(16) double sh[kFwdCols * kFwdRows];
  static_assert(kFwdWaves * 8 * 64 <= kFwdCols * kFwdRows, "LDS reuse");
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = blockIdx.y * 64 + lane;
  const NlocChunk C = chunks[blockIdx.x];
  const NlocProj P = proj[C.proj];
  const int bc = min(b, nb - 1);
  const double* src = psi16 + (int64_t)(bc >> 4) * Nd * 16 + (bc & 15);
  const int* idx = index + P.roff;
  const int rb = C.r0 + w * kFwdRowsPerWave;  // this wave's first row
  const int rlast = C.r0 + C.nr - 1;
  double v[kFwdRowsPerWave];
#pragma unroll
  for (int u = 0; u < kFwdRowsPerWave; ++u)
    v[u] = src[(int64_t)idx[min(rb + u, rlast)] * 16];
  double* dst = part + (P.poff + (int64_t)(blockIdx.x - P.chunk0) * P.ncol) * nb;
  for (int c0 = 0; c0 < P.ncol; c0 += kFwdCols) {
    const int nc = min(kFwdCols, P.ncol - c0);
    if (c0 > 0) __syncthreads();  // the previous pass's sums are read
    for (int e = tid; e < kFwdCols * kFwdRows; e += kFwdWaves * 64) {
      const int r = e / kFwdCols, c = e % kFwdCols;
      sh[e] = (c < nc && r < C.nr)
                  ? chi[P.choff + (int64_t)(c0 + c) * P.nrow + C.r0 + r]
                  : 0.0;
    }
    __syncthreads();
    double acc[kFwdCols];
#pragma unroll
    for (int c = 0; c < kFwdCols; ++c) acc[c] = 0.0;
    const double2* ch =
        reinterpret_cast<const double2*>(sh + w * kFwdRowsPerWave * kFwdCols);
#pragma unroll
    for (int u = 0; u < kFwdRowsPerWave; ++u) {
#pragma unroll
      for (int q = 0; q < kFwdCols / 2; ++q) {
        const double2 x = ch[u * (kFwdCols / 2) + q];
        acc[2 * q] = fma(x.x, v[u], acc[2 * q]);
        acc[2 * q + 1] = fma(x.y, v[u], acc[2 * q + 1]);
      }
    }
    // The sums are complete here: without this the compiler sinks the
    // second half's products below the first half's reduction and keeps
    // (spills) their chi values across it.
#pragma unroll
    for (int c = 0; c < kFwdCols; ++c) asm volatile("" : "+v"(acc[c]));
#pragma unroll
    for (int h = 0; h < kFwdCols / 8; ++h) {
      __syncthreads();  // chi (h = 0) or the previous half's sums are read
#pragma unroll
      for (int c = 0; c < 8; ++c) sh[(w * 8 + c) * 64 + lane] = acc[h * 8 + c];
      __syncthreads();
      const int c = h * 8 + w;  // wave w adds up column h * 8 + w
      if (c < nc) {
        double sum = sh[(c % 8) * 64 + lane];
#pragma unroll
        for (int q = 1; q < kFwdWaves; ++q)
          sum += sh[(q * 8 + c % 8) * 64 + lane];
        if (b < nb) dst[(int64_t)(c0 + c) * nb + b] = sum;
      }
    }
  }
}

// temp[roff_p + r][b] = sum_c chi_p[r][c] X_p[c][b] for the chunk's rows,
// with X_p[c][b] = (dv sum_j part[poff_p + j ncol_p + c][b]) factor (j in
// chunk order) formed in registers for this lane's band.  Wave w takes the
// chunk's rows w * 32 .. w * 32 + 31; chi is staged in LDS as in the
// forward pass and read as broadcasts.
__global__ __launch_bounds__(kFwdWaves * 64) void nloc_backward_kernel(
    const double* __restrict__ part, const NlocProj* __restrict__ proj,
    const NlocChunk* __restrict__ chunks, const double* __restrict__ factor,
    double dv, const double* __restrict__ chi, int nb,
    double* __restrict__ temp) {
  __shared__ __align__(16) double sh[kFwdCols * kFwdRows];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = blockIdx.y * 64 + lane;
  const int bc = min(b, nb - 1);
  const NlocChunk C = chunks[blockIdx.x];
  const NlocProj P = proj[C.proj];
  const int nr = min(kFwdRowsPerWave, C.nr - w * kFwdRowsPerWave);
  double* dst = temp + (P.roff + C.r0 + w * kFwdRowsPerWave) * nb + b;
  for (int c0 = 0; c0 < P.ncol; c0 += kFwdCols) {
    const int nc = min(kFwdCols, P.ncol - c0);
    // The 16 columns' parts of one chunk are loaded together, chunk after
    // chunk (each column's sum in chunk order).
    double xv[kFwdCols];
#pragma unroll
    for (int c = 0; c < kFwdCols; ++c) xv[c] = 0.0;
    const double* src = part + P.poff * nb + bc;
    for (int j = 0; j < P.nchunks; ++j) {
      double x[kFwdCols];
#pragma unroll
      for (int c = 0; c < kFwdCols; ++c)
        x[c] = src[((int64_t)j * P.ncol + min(c0 + c, P.ncol - 1)) * nb];
#pragma unroll
      for (int c = 0; c < kFwdCols; ++c) xv[c] += x[c];
    }
#pragma unroll
    for (int c = 0; c < kFwdCols; ++c)
      xv[c] = (dv * xv[c]) * factor[P.coff + min(c0 + c, P.ncol - 1)];
    if (c0 > 0) __syncthreads();  // the previous pass's chi is read
    for (int e = tid; e < kFwdCols * kFwdRows; e += kFwdWaves * 64) {
      const int r = e / kFwdCols, c = e % kFwdCols;
      sh[e] = (c < nc && r < C.nr)
                  ? chi[P.choff + (int64_t)(c0 + c) * P.nrow + C.r0 + r]
                  : 0.0;
    }
    __syncthreads();
    // One row at a time (its 16 chi values are eight 16-byte broadcasts);
    // a projector of more than 16 columns continues from the sums its
    // earlier passes stored.
    const double2* ch =
        reinterpret_cast<const double2*>(sh + w * kFwdRowsPerWave * kFwdCols);
    for (int u = 0; u < nr; ++u) {
      double t = (c0 > 0 && b < nb) ? dst[(int64_t)u * nb] : 0.0;
#pragma unroll
      for (int q = 0; q < kFwdCols / 2; ++q) {
        const double2 x = ch[u * (kFwdCols / 2) + q];
        t = fma(x.x, xv[2 * q], t);
        t = fma(x.y, xv[2 * q + 1], t);
      }
      if (b < nb) dst[(int64_t)u * nb] = t;
    }
  }
}

// result[g][b] += the temp rows of grid point g in projector order.  One
// lane per band, a wave per (grid point, 64 bands); grid point u (sorted
// list `point`) receives rows rows_of[ptr[u] .. ptr[u+1]), loaded four at a
// time from clamped positions and added in order.
__global__ void nloc_scatter_kernel(const double* __restrict__ temp, int nb,
                                    const int* __restrict__ point,
                                    const int* __restrict__ ptr,
                                    const int* __restrict__ rows_of,
                                    int64_t npoints, int Nd, int tshift,
                                    double* __restrict__ result) {
  const int b = blockIdx.x * 64 + (threadIdx.x & 63);
  const int bc = min(b, nb - 1);
  const int T = 1 << tshift;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int64_t u = (int64_t)blockIdx.y * 4 + wv; u < npoints;
       u += (int64_t)gridDim.y * 4) {
    double* dst =
        result + ((int64_t)(bc >> tshift) * Nd + point[u]) * T + (bc & (T - 1));
    double acc = *dst;
    const int j0 = ptr[u], j1 = ptr[u + 1];
    for (int j = j0; j < j1; j += 4) {
      double x[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        x[k] = temp[(int64_t)rows_of[min(j + k, j1 - 1)] * nb + bc];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (j + k < j1) acc += x[k];
    }
    if (b < nb) *dst = acc;
  }
}

dim3 rows_grid(int nb, int64_t rows) {
  const int64_t by = std::min<int64_t>((rows + 3) / 4, 65535);
  return dim3((nb + 63) / 64, (unsigned)std::max<int64_t>(by, 1));
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
  double* chi = nullptr;      // sum nrow * ncol
  double* factor = nullptr;   // cols
  NlocProj* proj = nullptr;   // num_proj
  NlocChunk* chunks = nullptr;  // nchunks
  int* point = nullptr;       // npoints
  int* ptr = nullptr;         // npoints + 1
  int* rows_of = nullptr;     // rows touching each point, in order
  int64_t npoints = 0;
  // workspace, grown with nb
  double* part = nullptr;     // parts x nb
  double* temp = nullptr;     // rows x nb
  int ws_nb = 0;
};

static void nloc_free(skgemm_nloc* h) {
  for (void* p : {(void*)h->index, (void*)h->chi, (void*)h->factor,
                  (void*)h->proj, (void*)h->chunks, (void*)h->point,
                  (void*)h->ptr, (void*)h->rows_of, (void*)h->part,
                  (void*)h->temp})
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
  int64_t poff = 0;
  for (int p = 0; p < num_proj; ++p) {
    NlocProj m{roff[p], choff[p], coff[p], nrow[p], ncol[p], 0, 0, poff};
    m.chunk0 = (int)chunks.size();
    if (ncol[p] > 0)
      for (int r0 = 0; r0 < nrow[p]; r0 += kFwdRows)
        chunks.push_back(NlocChunk{p, r0, std::min(kFwdRows, nrow[p] - r0)});
    m.nchunks = (int)chunks.size() - m.chunk0;
    poff += (int64_t)m.nchunks * ncol[p];
    meta[p] = m;
  }
  h->nchunks = (int64_t)chunks.size();
  h->parts = poff;
  // Scatter table: grid points in ascending order, each with its rows in
  // projector order (then row order), skipping projectors without columns.
  std::vector<std::pair<int, int64_t>> touch;
  touch.reserve((size_t)rows);
  for (int p = 0; p < num_proj; ++p) {
    if (ncol[p] == 0) continue;
    for (int r = 0; r < nrow[p]; ++r)
      touch.push_back({index_all[roff[p] + r], roff[p] + r});
  }
  std::stable_sort(touch.begin(), touch.end(),
                   [](const std::pair<int, int64_t>& a,
                      const std::pair<int, int64_t>& b) { return a.first < b.first; });
  std::vector<int> point, ptr(1, 0), rows_of;
  rows_of.reserve(touch.size());
  for (size_t i = 0; i < touch.size(); ++i) {
    if (i == 0 || touch[i].first != touch[i - 1].first) {
      point.push_back(touch[i].first);
      if (i > 0) ptr.push_back((int)i);
    }
    rows_of.push_back((int)touch[i].second);
  }
  if (!touch.empty()) ptr.push_back((int)touch.size());
  h->npoints = (int64_t)point.size();
  int rc = hipSetDevice(c->device) == hipSuccess
               ? SKGEMM_OK
               : set_error(SKGEMM_ERR_HIP, "hipSetDevice");
  if (rc == SKGEMM_OK)
    rc = upload(&h->index, index_all, (size_t)rows, "index", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->chi, chi_all, (size_t)choff[num_proj], "chi", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->factor, factor.data(), factor.size(), "gamma", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->proj, meta.data(), meta.size(), "projector table",
                c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->chunks, chunks.data(), chunks.size(), "chunk table",
                c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->point, point.data(), point.size(), "points", c->stream);
  if (rc == SKGEMM_OK && h->npoints > 0)
    rc = upload(&h->ptr, ptr.data(), ptr.size(), "point table", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->rows_of, rows_of.data(), rows_of.size(), "row list",
                c->stream);
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
  SKG_TRY(check_dev(result, "result"));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  if (nb > h->ws_nb) {
    SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (double** p : {&h->part, &h->temp}) {
      if (*p) hipFree(*p);
      *p = nullptr;
    }
    h->ws_nb = 0;
    if (hipMalloc(&h->part, (size_t)std::max<int64_t>(h->parts, 1) * nb *
                                sizeof(double)) != hipSuccess ||
        hipMalloc(&h->temp, (size_t)h->rows * nb * sizeof(double)) != hipSuccess)
      return set_error(SKGEMM_ERR_NOMEM, "nloc workspace for %d bands", nb);
    h->ws_nb = nb;
    ++c->mem_gen;
  }
  const unsigned bgroups = (unsigned)((nb + 63) / 64);
  if (h->nchunks == 0) return SKGEMM_OK;  // no projector has columns
  const dim3 cgrid((unsigned)h->nchunks, bgroups);
  hipLaunchKernelGGL(nloc_forward_kernel, cgrid, dim3(kFwdWaves * 64), 0,
                     c->stream, psi16, h->Nd, nb, h->index, h->chi, h->proj,
                     h->chunks, h->part);
  SKG_HIP(hipGetLastError(), "nloc forward launch");
  hipLaunchKernelGGL(nloc_backward_kernel, cgrid, dim3(kFwdWaves * 64), 0,
                     c->stream, h->part, h->proj, h->chunks, h->factor, dv,
                     h->chi, nb, h->temp);
  SKG_HIP(hipGetLastError(), "nloc backward launch");
  hipLaunchKernelGGL(nloc_scatter_kernel, rows_grid(nb, h->npoints), dim3(256),
                     0, c->stream, h->temp, nb, h->point, h->ptr, h->rows_of,
                     h->npoints, h->Nd, result_tile == 32 ? 5 : 4, result);
  SKG_HIP(hipGetLastError(), "nloc scatter launch");
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
