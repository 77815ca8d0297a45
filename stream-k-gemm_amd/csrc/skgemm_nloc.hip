// skgemm_nloc.hip - non-local projector application on band-packed orbitals,
// result += V_nloc psi (include/nloc_packed_opt2.hpp:229-655), as a client of
// the batched Stream-K scheduler.
//
// Per projector p (grid indices idx_p, nrow_p of them; chi_p nrow_p x
// ncol_p; gamma_p):  X_p = dv chi_p^T psi[idx_p, :],  X_p *= gamma_p (first
// projector of each atom only),  result[idx_p, :] += chi_p X_p.
// Device sequence, all on the context's stream:
//   1. forward products fused with the gather: X_p = dv chi_p^T psi[idx_p,
//      :] read straight from psi16 at the projector's grid points, times
//      gamma where the atom rule applies (times 1.0, exact, elsewhere); one
//      workgroup per (projector, 64 bands);
//   2. backward products: one persistent batched Stream-K launch, one
//      problem per projector (temp_p = chi_p X_p); its problem table is
//      described once per band count and kept in device memory, so an
//      application is three launches and no host work;
//   3. scatter: result[g, b] += temp[R, b] for every row R with grid index
//      g, each grid point by one thread in projector order (the rows that
//      touch g are listed in a CSR table built at create time), so shared
//      grid points need no atomics and the sums keep the reference's order.
// HBM-bound: the forward pass reads nrow_total x nb doubles of psi (repeats
// from overlapping projectors mostly hit L2), the backward products write
// and the scatter reads them; the products are small (ncol <= a few tens).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/skgemm.h"
#include "skgemm_ctx.h"

namespace skg {
namespace {

// Forward products fused with the gather:
//   X[coff_p + c][b] = (dv sum_r chi_p[r][c] psi(index[roff_p + r], b)) factor
// One workgroup per (projector, 64 bands) (grid x: projectors, y: groups
// of 64 bands), 8 waves.  Rows go in chunks of 64: the chunk's chi values
// are staged in LDS (row-major, read as broadcasts) and wave w takes rows w,
// w + 8, ... of the chunk (8 gathered loads in flight), accumulating 16
// columns per lane; the 8 partial sums are added in wave order through LDS.
// Columns past 16 take further passes.  chi_p is column-major nrow x ncol.
struct NlocProj {
  int64_t roff, choff, coff;
  int nrow, ncol;
};
constexpr int kNlocCols = 16, kNlocWaves = 8, kNlocChunk = 64;

__global__ __launch_bounds__(kNlocWaves * 64) void nloc_project_kernel(
    const double* __restrict__ psi16, int Nd, int nb,
    const int* __restrict__ index, const double* __restrict__ chi,
    const NlocProj* __restrict__ proj, const double* __restrict__ factor,
    double dv, double* __restrict__ X) {
  __shared__ double part[kNlocWaves][kNlocCols][64];
  __shared__ double chs[kNlocChunk][kNlocCols];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int b = blockIdx.y * 64 + lane;
  const NlocProj P = proj[blockIdx.x];
  const bool bok = b < nb;
  const double* src = psi16 + (int64_t)(b >> 4) * Nd * 16 + (b & 15);
  const int* idx = index + P.roff;
  constexpr int kPer = kNlocChunk / kNlocWaves;  // rows per wave and chunk
  for (int c0 = 0; c0 < P.ncol; c0 += kNlocCols) {
    const int nc = P.ncol - c0 < kNlocCols ? P.ncol - c0 : kNlocCols;
    const double* ch = chi + P.choff + (int64_t)c0 * P.nrow;
    double acc[kNlocCols];
#pragma unroll
    for (int c = 0; c < kNlocCols; ++c) acc[c] = 0.0;
    for (int r0 = 0; r0 < P.nrow; r0 += kNlocChunk) {
      __syncthreads();  // the previous chunk's chi reads are done
      for (int e = tid; e < kNlocChunk * kNlocCols; e += kNlocWaves * 64) {
        const int r = e / kNlocCols, c = e % kNlocCols;
        chs[r][c] =
            (c < nc && r0 + r < P.nrow) ? ch[(int64_t)c * P.nrow + r0 + r] : 0.0;
      }
      double v[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int r = r0 + w + kNlocWaves * u;
        v[u] = (bok && r < P.nrow) ? src[(int64_t)idx[r] * 16] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < kPer; ++u)
#pragma unroll
        for (int c = 0; c < kNlocCols; ++c)
          acc[c] = fma(chs[w + kNlocWaves * u][c], v[u], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < kNlocCols; ++c) part[w][c][lane] = acc[c];
    __syncthreads();
    for (int c = w; c < nc; c += kNlocWaves) {
      double sum = part[0][c][lane];
#pragma unroll
      for (int q = 1; q < kNlocWaves; ++q) sum += part[q][c][lane];
      if (bok) {
        const int64_t col = P.coff + c0 + c;
        X[col * nb + b] = (dv * sum) * factor[col];
      }
    }
    __syncthreads();
  }
}

// One lane per band, 64 consecutive bands per wave (a 16-band panel row is
// 128 contiguous bytes).  Grid point u (sorted list `point`) receives the rows ptr[u] .. ptr[u+1]
// of temp in projector order.
__global__ void nloc_scatter_kernel(const double* __restrict__ temp, int nb,
                                    const int* __restrict__ point,
                                    const int* __restrict__ ptr,
                                    const int* __restrict__ rows_of,
                                    int64_t npoints, int Nd, int tshift,
                                    double* __restrict__ result) {
  const int b = blockIdx.x * 64 + (threadIdx.x & 63);
  if (b >= nb) return;
  const int T = 1 << tshift;
  for (int64_t u = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6); u < npoints;
       u += (int64_t)gridDim.y * 4) {
    double* dst =
        result + ((int64_t)(b >> tshift) * Nd + point[u]) * T + (b & (T - 1));
    double acc = *dst;
    for (int j = ptr[u]; j < ptr[u + 1]; ++j)
      acc += temp[(int64_t)rows_of[j] * nb + b];
    *dst = acc;
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
  std::vector<int> nrow, ncol;
  std::vector<int64_t> roff, coff, choff;  // row, column, chi offsets
  int64_t rows = 0, cols = 0;
  // device copies
  int* index = nullptr;       // rows
  double* chi = nullptr;      // sum nrow * ncol
  double* factor = nullptr;   // cols
  NlocProj* proj = nullptr;   // num_proj (rows, chi and column offsets)
  int* point = nullptr;       // npoints
  int* ptr = nullptr;         // npoints + 1
  int* rows_of = nullptr;     // rows
  int64_t npoints = 0;
  // workspace, grown with nb
  double* X = nullptr;
  double* temp = nullptr;
  int ws_nb = 0;
  // The two grouped products, described for `table_nb` bands and kept in
  // device memory (their rows depend on nb and the workspace only).
  PlainTable bw;
  Problem* d_bw = nullptr;
  int64_t* d_hdr = nullptr;  // {rows, iterations} of bw
  int table_nb = -1, table_cfg = -2, table_grid = -2;
};

static void nloc_free(skgemm_nloc* h) {
  for (void* p : {(void*)h->index, (void*)h->chi, (void*)h->factor,
                  (void*)h->proj,
                  (void*)h->point, (void*)h->ptr, (void*)h->rows_of,
                  (void*)h->X, (void*)h->temp, (void*)h->d_bw,
                  (void*)h->d_hdr})
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
  auto* h = new skgemm_nloc();
  h->ctx = c;
  h->Nd = Nd;
  h->num_proj = num_proj;
  h->roff.assign(num_proj + 1, 0);
  h->coff.assign(num_proj + 1, 0);
  h->choff.assign(num_proj + 1, 0);
  for (int p = 0; p < num_proj; ++p) {
    if (nrow[p] < 0 || ncol[p] < 0) {
      delete h;
      return set_error(SKGEMM_ERR_ARG, "nloc: projector %d has a negative size",
                       p);
    }
    h->nrow.push_back(nrow[p]);
    h->ncol.push_back(ncol[p]);
    h->roff[p + 1] = h->roff[p] + nrow[p];
    h->coff[p + 1] = h->coff[p] + ncol[p];
    h->choff[p + 1] = h->choff[p] + (int64_t)nrow[p] * ncol[p];
  }
  h->rows = h->roff[num_proj];
  h->cols = h->coff[num_proj];
  if (h->rows >= INT32_MAX || h->cols >= INT32_MAX) {
    delete h;
    return set_error(SKGEMM_ERR_NOTSUP, "nloc: more than 2^31 projector rows");
  }
  if ((h->rows > 0 && index_all == nullptr) ||
      (h->choff[num_proj] > 0 && chi_all == nullptr) ||
      (h->cols > 0 && gamma_all == nullptr)) {
    delete h;
    return set_error(SKGEMM_ERR_ARG, "nloc: index, chi or gamma is NULL");
  }
  for (int64_t R = 0; R < h->rows; ++R)
    if (index_all[R] < 0 || index_all[R] >= Nd) {
      delete h;
      return set_error(SKGEMM_ERR_ARG, "nloc: grid index %d out of [0, %d)",
                       index_all[R], Nd);
    }
  // Row factors: gamma for the first projector of each atom (the reference
  // starts from the first projector's atom_index - 1), 1.0 elsewhere.
  std::vector<double> factor((size_t)h->cols, 1.0);
  if (num_proj > 0) {
    int atom = atom_index[0] - 1;
    for (int p = 0; p < num_proj; ++p) {
      if (atom == atom_index[p]) continue;
      atom = atom_index[p];
      for (int j = 0; j < ncol[p]; ++j)
        factor[h->coff[p] + j] = gamma_all[h->coff[p] + j];
    }
  }
  // Scatter table: grid points in ascending order, each with its rows in
  // projector order (then row order), skipping projectors without columns
  // (they contribute nothing).
  std::vector<std::pair<int, int64_t>> touch;
  touch.reserve((size_t)h->rows);
  for (int p = 0; p < num_proj; ++p) {
    if (ncol[p] == 0) continue;
    for (int r = 0; r < nrow[p]; ++r)
      touch.push_back({index_all[h->roff[p] + r], h->roff[p] + r});
  }
  std::stable_sort(touch.begin(), touch.end(),
                   [](const std::pair<int, int64_t>& a,
                      const std::pair<int, int64_t>& b) { return a.first < b.first; });
  std::vector<int> point, ptr(1, 0), rows_of;
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
    rc = upload(&h->index, index_all, (size_t)h->rows, "index", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->chi, chi_all, (size_t)h->choff[num_proj], "chi", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->factor, factor.data(), factor.size(), "gamma", c->stream);
  std::vector<NlocProj> meta((size_t)num_proj);
  for (int p = 0; p < num_proj; ++p)
    meta[p] = NlocProj{h->roff[p], h->choff[p], h->coff[p], nrow[p], ncol[p]};
  if (rc == SKGEMM_OK)
    rc = upload(&h->proj, meta.data(), meta.size(), "projector table",
                c->stream);
  if (rc == SKGEMM_OK) rc = upload(&h->point, point.data(), point.size(), "points", c->stream);
  if (rc == SKGEMM_OK && h->npoints > 0)
    rc = upload(&h->ptr, ptr.data(), ptr.size(), "point table", c->stream);
  if (rc == SKGEMM_OK)
    rc = upload(&h->rows_of, rows_of.data(), rows_of.size(), "row table", c->stream);
  if (rc != SKGEMM_OK) {
    nloc_free(h);
    delete h;
    return rc;
  }
  *out = h;
  return SKGEMM_OK;
}

// Describe the backward products temp_p (nrow x nb, row-major) = chi_p
// X_p (chi_p column-major nrow x ncol, X_p row-major ncol x nb) and upload
// their table.
static int nloc_tables(skgemm_nloc* h, int nb) {
  skgemm_ctx* c = h->ctx;
  std::vector<skgemm_problem> bw;
  for (int p = 0; p < h->num_proj; ++p) {
    const int nr = h->nrow[p], nc = h->ncol[p];
    if (nc == 0 || nr == 0) continue;
    skgemm_problem g{};
    g.kind = 0;
    g.layout_a = SKGEMM_COL_MAJOR;
    g.layout_b = SKGEMM_ROW_MAJOR;
    g.layout_d = SKGEMM_ROW_MAJOR;
    g.M = nr;
    g.N = nb;
    g.K = nc;
    g.alpha = 1.0;
    g.A = h->chi + h->choff[p];
    g.lda = nr;
    g.B = h->X + h->coff[p] * nb;
    g.ldb = nb;
    g.D = h->temp + h->roff[p] * nb;
    g.ldd = nb;
    bw.push_back(g);
  }
  h->table_nb = -1;
  SKG_TRY(build_plain_table(c, bw.data(), (int)bw.size(), &h->bw));
  SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (h->d_bw) hipFree(h->d_bw);
  h->d_bw = nullptr;
  const size_t nbk = std::max<size_t>(h->bw.rows.size(), 1);
  if (hipMalloc(&h->d_bw, nbk * sizeof(Problem)) != hipSuccess ||
      (h->d_hdr == nullptr &&
       hipMalloc(&h->d_hdr, 2 * sizeof(int64_t)) != hipSuccess))
    return set_error(SKGEMM_ERR_NOMEM, "nloc problem table");
  const int64_t hdr[2] = {(int64_t)h->bw.rows.size(), h->bw.iters};
  // On the context's stream, completed before returning (hdr is local).
  SKG_HIP(hipMemcpyAsync(h->d_bw, h->bw.rows.data(),
                         h->bw.rows.size() * sizeof(Problem),
                         hipMemcpyHostToDevice, c->stream),
          "nloc table upload");
  SKG_HIP(hipMemcpyAsync(h->d_hdr, hdr, sizeof(hdr), hipMemcpyHostToDevice,
                         c->stream),
          "nloc table upload");
  SKG_HIP(hipStreamSynchronize(c->stream), "nloc table upload");
  h->table_nb = nb;
  h->table_cfg = c->forced_cfg;
  h->table_grid = c->forced_grid;
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
    for (double** p : {&h->X, &h->temp}) {
      if (*p) hipFree(*p);
      *p = nullptr;
    }
    h->ws_nb = 0;
    if (hipMalloc(&h->X, (size_t)h->cols * nb * sizeof(double)) != hipSuccess ||
        hipMalloc(&h->temp, (size_t)h->rows * nb * sizeof(double)) != hipSuccess)
      return set_error(SKGEMM_ERR_NOMEM, "nloc workspace for %d bands", nb);
    h->ws_nb = nb;
    ++c->mem_gen;
  }
  if (nb != h->table_nb || h->ws_nb != nb || c->forced_cfg != h->table_cfg ||
      c->forced_grid != h->table_grid) {
    SKG_TRY(nloc_tables(h, nb));
  }
  hipLaunchKernelGGL(nloc_project_kernel,
                     dim3((unsigned)h->num_proj, (nb + 63) / 64),
                     dim3(kNlocWaves * 64), 0,
                     c->stream, psi16, h->Nd, nb, h->index, h->chi, h->proj,
                     h->factor, dv, h->X);
  SKG_HIP(hipGetLastError(), "nloc forward launch");
  SKG_TRY(launch_plain_table(c, h->bw, h->d_bw, h->d_hdr));
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
