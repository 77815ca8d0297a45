// Autotuner: picks (tile configuration, persistent grid, K chunks) per
// shape and layout by timing real launches, and persists the choice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <set>
#include <utility>
#include <tuple>
#include <vector>

#include "../../include/skgemm.h"
#include "skgemm_ctx.h"
#include "skgemm_internal.h"

using namespace skg;

namespace {

struct Saved {
  int cfg, grid, kchunks;
};

// Scratch operands of one shape; freed on scope exit.
struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) hipFree(p);
  }
  double* alloc(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess)
      return nullptr;
    ptrs.push_back(p);
    return static_cast<double*>(p);
  }
};

int64_t span(int layout, bool a_conv, int64_t ld, int ext, int K) {
  return span_operand(layout, a_conv, ld, ext, K);
}

int64_t min_ld(int layout, bool a_conv, int ext, int K) {
  if (layout == SKGEMM_PACK16 || layout == SKGEMM_PACK32) return std::max(K, 1);
  const bool o_major = (layout == SKGEMM_ROW_MAJOR) ? !a_conv : a_conv;
  return o_major ? std::max(ext, 1) : std::max(K, 1);
}

}  // namespace

extern "C" int skgemm_autotune(skgemm_ctx* c, int kind, int la, int lb, int M,
                               int N, int K, const char* cache_path,
                               int* best_tile, int* best_grid,
                               int* best_kchunks, double* best_us) {
  int rc = check_ctx(c);
  if (rc != SKGEMM_OK) return rc;
  rc = check_dims(M, N, K);
  if (rc != SKGEMM_OK) return rc;
  if (kind != 0 && kind != 1) return set_error(SKGEMM_ERR_ARG, "kind must be 0 or 1");
  if (kind == 1) {
    N = M;
    lb = la;
  }
  if (M == 0 || N == 0) return set_error(SKGEMM_ERR_ARG, "empty shape");
  if (hipSetDevice(c->device) != hipSuccess)
    return set_error(SKGEMM_ERR_HIP, "hipSetDevice");

  Scratch sc;
  const int64_t lda = min_ld(la, kind == 0, M, K);
  const int64_t ldb = min_ld(lb, false, N, K);
  double* A = sc.alloc(span(la, kind == 0, lda, M, K));
  double* B = kind == 0 ? sc.alloc(span(lb, false, ldb, N, K)) : A;
  const int ld_layout = (la == SKGEMM_COL_MAJOR && kind == 0) ? SKGEMM_COL_MAJOR
                                                             : SKGEMM_ROW_MAJOR;
  // Plans are keyed by the problem as launched (a column-major D runs
  // transposed).
  int kla = la, klb = lb, kM = M, kN = N;
  if (kind == 0) {
    int kld = ld_layout;
    const double *pa = nullptr, *pb = nullptr;
    int64_t x = 0, y = 0;
    transpose_launch(kla, klb, kld, kM, kN, pa, x, pb, y);
  }
  const int64_t ldd = ld_layout == SKGEMM_COL_MAJOR ? M : N;
  double* D = sc.alloc((size_t)M * N);
  if (!A || !B || !D) return set_error(SKGEMM_ERR_NOMEM, "autotune scratch");
  const int64_t na = span(la, kind == 0, lda, M, K);
  skgemm_fill_uniform(c, A, na, 7, -1.0, 1.0);
  if (kind == 0) skgemm_fill_uniform(c, B, span(lb, false, ldb, N, K), 8, -1.0, 1.0);

  const Saved saved{c->forced_cfg, c->forced_grid, c->forced_kchunks};
  auto run = [&]() {
    return kind == 0 ? skgemm_dgemm(c, la, lb, ld_layout, M, N, K, 1.0, A, lda,
                                    B, ldb, 0.0, D, ldd, 0.0)
                     : skgemm_dsyrk(c, la, N, K, 1.0, A, lda, 0.0, D, N);
  };
  hipEvent_t e0, e1;
  // Timing-only events without the system-scope fence, like the context's.
  hipEventCreateWithFlags(&e0, hipEventDisableSystemFence);
  hipEventCreateWithFlags(&e1, hipEventDisableSystemFence);
  double best = 1e30;
  Plan bp{-1, 0, 1};
  for (int cfg = 0; cfg < num_tile_cfgs(); ++cfg) {
    const TileCfg& t = tile_cfg(cfg);
    if (kind == 1 && t.bm != t.bn) continue;
    // A strip configuration differs from its twin only when the strip
    // applies to this shape.
    if (t.xs && strip_cols(cfg, kind, kN) == 0 &&
        strip_rows(cfg, kind, kM, kN) == 0)
      continue;
    const int cap = c->capacity(cfg);
    std::set<int> grids = {cap, std::min(cap, 2 * c->num_cus), std::min(cap, c->num_cus)};
    const int64_t ipt = std::max<int64_t>(1, (K + t.bk - 1) / t.bk);
    std::set<int> chunks = {1};
    if (ipt >= 16 * c->num_xcds) chunks.insert(c->num_xcds);
    std::set<std::pair<int, int>> cands;
    for (int g : grids)
      for (int kc : chunks) cands.insert({g, kc});
    // Data-parallel candidates: one (K chunk, tile) unit per workgroup, so
    // no fix-up and the workgroups of one XCD run tiles of one K chunk in
    // step (panel reuse in L2).
    const int64_t tiles = tiles_for(kind, cfg, kM, kN);
    for (int kc = 1; kc <= 16 && ipt / kc >= 8; ++kc) {
      const int64_t units = tiles * kc;
      if (units <= cap && 10 * units >= 7 * (int64_t)cap)
        cands.insert({(int)units, kc});
    }
    for (const auto& gk : cands) {
      const int g = gk.first, kc = gk.second;
      c->forced_cfg = cfg;
      c->forced_grid = g;
      c->forced_kchunks = kc;
      rc = SKGEMM_OK;
      for (int i = 0; i < 2 && rc == SKGEMM_OK; ++i) rc = run();
      if (rc != SKGEMM_OK) break;
      const int reps = 5;
      hipEventRecord(e0, c->stream);
      for (int i = 0; i < reps && rc == SKGEMM_OK; ++i) rc = run();
      hipEventRecord(e1, c->stream);
      if (rc != SKGEMM_OK) break;
      hipEventSynchronize(e1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      const double us = 1e3 * ms / reps;
      if (us < best) {
        best = us;
        bp = plan(c, kind, kla, klb, kM, kN, K);
      }
    }
    if (rc != SKGEMM_OK) break;
  }
  c->forced_cfg = saved.cfg;
  c->forced_grid = saved.grid;
  c->forced_kchunks = saved.kchunks;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != SKGEMM_OK) return rc;
  const int st = c->check_status();
  if (st != SKGEMM_OK) return st;
  c->tuned[std::make_tuple(kind, kla, klb, kM, kN, K)] = bp;
  if (cache_path != nullptr) {
    FILE* f = std::fopen(cache_path, "a");
    if (f == nullptr) return set_error(SKGEMM_ERR_ARG, "cannot open %s", cache_path);
    std::fprintf(f, "%d %d %d %d %d %d %d %d %d %.3f\n", kind, kla, klb, kM,
                 kN, K, bp.cfg, bp.grid, bp.kchunks, best);
    std::fclose(f);
  }
  if (best_tile) *best_tile = bp.cfg;
  if (best_grid) *best_grid = bp.grid;
  if (best_kchunks) *best_kchunks = bp.kchunks;
  if (best_us) *best_us = best;
  return SKGEMM_OK;
}

extern "C" int skgemm_load_tuning(skgemm_ctx* c, const char* cache_path) {
  int rc = check_ctx(c);
  if (rc != SKGEMM_OK) return rc;
  if (cache_path == nullptr) return set_error(SKGEMM_ERR_ARG, "path is NULL");
  FILE* f = std::fopen(cache_path, "r");
  if (f == nullptr) return set_error(SKGEMM_ERR_ARG, "cannot open %s", cache_path);
  int kind, la, lb, M, N, K, cfg, grid, kc;
  double us;
  while (std::fscanf(f, "%d %d %d %d %d %d %d %d %d %lf", &kind, &la, &lb, &M,
                     &N, &K, &cfg, &grid, &kc, &us) == 10) {
    if (cfg < 0 || cfg >= num_tile_cfgs() || grid <= 0 || kc <= 0) continue;
    c->tuned[std::make_tuple(kind, la, lb, M, N, K)] = Plan{cfg, grid, kc};
  }
  std::fclose(f);
  return SKGEMM_OK;
}
