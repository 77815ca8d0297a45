// Batched launches: many GEMM / SYRK problems of uneven sizes in one
// persistent Stream-K launch over their flattened MAC-iteration space
// (plain FP64), or one split-form sequence per problem (Ozaki INT8), both
// optionally captured in a hipGraph.  No counterpart in the reference,
// which runs one problem per call.
//
// Plain FP64: the problem table and its header {count, total iterations}
// live in device memory and are refreshed with stream-ordered copies from
// pinned host staging before every launch, so one graph per (tile, grid)
// replays any batch.  Fix-up flags are consumed and cleared by the kernel
// itself (no memset node in the graph).
//
// Split form: every launch argument of the sequence (operand pointers,
// sizes) is baked into the graph, so graphs are cached per exact problem
// list; all workspace is reserved before capture.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "../../include/skgemm.h"
#include "skgemm_ctx.h"

using namespace skg;

struct skgemm_ctx::Batch {
  Problem* dtab = nullptr;  // device problem table
  int64_t* dhdr = nullptr;  // device {count, total iterations}
  Problem* htab = nullptr;  // pinned staging
  int64_t* hhdr = nullptr;
  int cap = 0;
  hipEvent_t copied = nullptr;  // staging may be rewritten once it fired
  bool pending = false;
  struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    uint64_t gen = 0;  // mem_gen at capture
  };
  std::map<std::pair<int, int>, Graph> plain;
  std::map<std::vector<int64_t>, Graph> split;
  static void drop(Graph& x) {
    if (x.ge) hipGraphExecDestroy(x.ge);
    if (x.g) hipGraphDestroy(x.g);
    x.ge = nullptr;
    x.g = nullptr;
  }
};

void skgemm_ctx::free_batch() {
  if (batch == nullptr) return;
  Batch* b = batch;
  if (b->pending) hipEventSynchronize(b->copied);
  for (auto& kv : b->plain) Batch::drop(kv.second);
  for (auto& kv : b->split) Batch::drop(kv.second);
  if (b->dtab) hipFree(b->dtab);
  if (b->dhdr) hipFree(b->dhdr);
  if (b->htab) hipHostFree(b->htab);
  if (b->hhdr) hipHostFree(b->hhdr);
  if (b->copied) hipEventDestroy(b->copied);
  delete b;
  batch = nullptr;
}

namespace {

constexpr size_t kMaxSplitGraphs = 8;

int batch_state(skgemm_ctx* c, int count) {
  if (c->batch == nullptr) {
    c->batch = new skgemm_ctx::Batch();
    SKG_HIP(hipEventCreateWithFlags(&c->batch->copied, hipEventDisableTiming),
            "batch event");
    SKG_HIP(hipMalloc(&c->batch->dhdr, 2 * sizeof(int64_t)), "batch header");
    SKG_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->batch->hhdr),
                          2 * sizeof(int64_t), hipHostMallocDefault),
            "batch header staging");
  }
  auto* b = c->batch;
  if (count > b->cap) {
    const int cap = std::max(count, 2 * b->cap);
    SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    ++c->mem_gen;
    if (b->dtab) hipFree(b->dtab);
    if (b->htab) hipHostFree(b->htab);
    b->dtab = nullptr;
    b->htab = nullptr;
    b->cap = 0;
    b->pending = false;
    SKG_HIP(hipMalloc(&b->dtab, sizeof(Problem) * cap), "batch table");
    SKG_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->htab),
                          sizeof(Problem) * cap, hipHostMallocDefault),
            "batch table staging");
    b->cap = cap;
  }
  return SKGEMM_OK;
}

// Problem descriptor of q and the (M, N) it is launched with (with
// `transpose`, a GEMM with a column-major D runs transposed).
int describe(const skgemm_problem& q, Problem* p, int* Ml, int* Nl,
             bool transpose) {
  if (q.kind == 0) {
    int la = q.layout_a, lb = q.layout_b, ld_ = q.layout_d, M = q.M, N = q.N;
    const double *A = q.A, *B = q.B;
    int64_t lda = q.lda, ldb = q.ldb;
    if (transpose) transpose_launch(la, lb, ld_, M, N, A, lda, B, ldb);
    *Ml = M;
    *Nl = N;
    return prepare_gemm(la, lb, ld_, M, N, q.K, A, lda, B, ldb, q.D, q.ldd, p);
  }
  if (q.kind == 1) {
    if (q.M != q.N)
      return set_error(SKGEMM_ERR_ARG, "SYRK problem needs M == N (got %d, %d)",
                       q.M, q.N);
    *Ml = *Nl = q.N;
    return prepare_syrk(q.layout_a, q.N, q.K, q.A, q.lda, q.D, q.ldd, p);
  }
  return set_error(SKGEMM_ERR_ARG, "problem kind must be 0 (GEMM) or 1 (SYRK)");
}

// One tile configuration for the whole batch (square tiles: SYRK needs
// them).  The 128x128 / 8-wave tile unless every problem is small.
// The 128 x 128 tiles run their LDS-DMA variant (cfg 8) when every operand
// of the batch has 16-byte aligned pairs (checked by run_plain).
int batch_cfg(skgemm_ctx* c, const skgemm_problem* q, int n) {
  // Strip tiles run single problems only: a forced strip configuration
  // batches with its twin.
  if (c->forced_cfg >= 0 && tile_cfg(c->forced_cfg).bm == tile_cfg(c->forced_cfg).bn)
    return strip_twin(c->forced_cfg);
  int big = 0;
  for (int i = 0; i < n; ++i) big = std::max(big, std::min(q[i].M, q[i].N));
  return big >= 256 ? 8 : 1;
}

}  // namespace

namespace skg {

int build_plain_table(skgemm_ctx* c, const skgemm_problem* q, int n,
                      PlainTable* out) {
  PlainTable& tb = *out;
  tb.rows.clear();
  int cfg = batch_cfg(c, q, n);
  struct Row {
    Problem p;
    int M, N, i;
  };
  std::vector<Row> rows;
  rows.reserve(n);
  bool aligned = true;
  for (int i = 0; i < n; ++i) {
    Row r{};
    SKG_TRY(describe(q[i], &r.p, &r.M, &r.N, true));
    if (r.M == 0 || r.N == 0) continue;
    aligned = aligned && r.p.a.vec2 && r.p.b.vec2;
    r.i = i;
    rows.push_back(r);
  }
  // DMA tiles move 16-byte pairs; same tiles, register-staged loads else.
  if (tile_cfg(cfg).ns >= 2 && !aligned) cfg = register_sibling(cfg);
  int grid = c->capacity(cfg);
  if (c->forced_grid > 0) grid = std::min(grid, c->forced_grid);
  grid = std::max(8, grid - grid % 8);
  int64_t iters = 0, tiles = 0;
  for (Row& r : rows) {
    const skgemm_problem& x = q[r.i];
    const Plan pl{cfg, grid, 1};
    SKG_TRY(fill_problem(c, &r.p, x.kind, pl, r.M, r.N, x.K));
    r.p.alpha = x.alpha;
    r.p.beta = x.beta;
    r.p.gamma = x.kind == 1 ? 0.0 : x.gamma;
    r.p.pred = nullptr;
    r.p.tile_begin = tiles;
    r.p.iter_begin = iters;
    tiles += r.p.num_tiles;
    iters += r.p.num_tiles * (int64_t)r.p.kci;
    tb.rows.push_back(r.p);
  }
  tb.cfg = cfg;
  tb.grid = grid;
  tb.iters = iters;
  return SKGEMM_OK;
}

int launch_plain_table(skgemm_ctx* c, const PlainTable& tb,
                       const Problem* dtab, const int64_t* dhdr) {
  if (tb.rows.empty()) return SKGEMM_OK;
  SKG_TRY(c->ensure_ws(c->ws_doubles(tb.cfg, tb.grid), tb.grid));
  Launch L{};
  L.problems = dtab;
  L.hdr = dhdr;
  L.num_problems = (int)tb.rows.size();
  L.total_iters = tb.iters;
  L.grid = tb.grid;
  L.ws = c->ws;
  L.flags = c->flags;
  L.status = c->status;
  L.spin_limit = c->fixup_spin_limit();
  L.waitfree = c->waitfree ? 1 : 0;
  L.hold = c->fixup_hold ? 1 : 0;
  SKG_HIP(launch_dgemm_sk(tb.cfg, L, c->stream), "batched launch");
  return SKGEMM_OK;
}

}  // namespace skg

namespace {

int run_plain(skgemm_ctx* c, const skgemm_problem* q, int n, int use_graph) {
  PlainTable tb;
  SKG_TRY(build_plain_table(c, q, n, &tb));
  const int cfg = tb.cfg, grid = tb.grid;
  const int used = (int)tb.rows.size();
  const int64_t iters = tb.iters;
  if (used == 0) return SKGEMM_OK;
  SKG_TRY(batch_state(c, n));
  auto* b = c->batch;
  if (b->pending) {
    SKG_HIP(hipEventSynchronize(b->copied), "batch staging wait");
    b->pending = false;
  }
  std::memcpy(b->htab, tb.rows.data(), sizeof(Problem) * used);
  SKG_TRY(c->ensure_ws(c->ws_doubles(cfg, grid), grid));
  b->hhdr[0] = used;
  b->hhdr[1] = iters;
  SKG_HIP(hipMemcpyAsync(b->dtab, b->htab, sizeof(Problem) * used,
                         hipMemcpyHostToDevice, c->stream),
          "batch table upload");
  SKG_HIP(hipMemcpyAsync(b->dhdr, b->hhdr, 2 * sizeof(int64_t),
                         hipMemcpyHostToDevice, c->stream),
          "batch header upload");
  SKG_HIP(hipEventRecord(b->copied, c->stream), "batch event");
  b->pending = true;

  Launch L{};
  L.problems = b->dtab;
  L.hdr = b->dhdr;
  L.num_problems = used;
  L.total_iters = iters;
  L.grid = grid;
  L.ws = c->ws;
  L.flags = c->flags;
  L.status = c->status;
  L.spin_limit = c->fixup_spin_limit();
  L.waitfree = c->waitfree ? 1 : 0;
  L.hold = c->fixup_hold ? 1 : 0;
  c->record_begin();
  if (!use_graph) {
    SKG_HIP(launch_dgemm_sk(cfg, L, c->stream), "batched launch");
  } else {
    skgemm_ctx::Batch::Graph& gr = b->plain[std::make_pair(cfg, grid)];
    if (gr.ge != nullptr && gr.gen != c->mem_gen) {
      SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
      skgemm_ctx::Batch::drop(gr);
    }
    if (gr.ge == nullptr) {
      SKG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed),
              "begin capture");
      c->capturing = true;
      const hipError_t le = launch_dgemm_sk(cfg, L, c->stream);
      c->capturing = false;
      const hipError_t ce = hipStreamEndCapture(c->stream, &gr.g);
      if (le != hipSuccess || ce != hipSuccess) {
        skgemm_ctx::Batch::drop(gr);
        return hip_error(le != hipSuccess ? le : ce, "batched launch capture");
      }
      const hipError_t ie = hipGraphInstantiate(&gr.ge, gr.g, nullptr, nullptr, 0);
      if (ie != hipSuccess) {
        skgemm_ctx::Batch::drop(gr);
        return hip_error(ie, "graph instantiate");
      }
      gr.gen = c->mem_gen;
    }
    SKG_HIP(hipGraphLaunch(gr.ge, c->stream), "graph launch");
  }
  c->record_end();
  return SKGEMM_OK;
}

int run_split(skgemm_ctx* c, const skgemm_problem* q, int n, int slices,
              int use_graph) {
  std::vector<Problem> ps(n);
  for (int i = 0; i < n; ++i) {
    int Ml = 0, Nl = 0;
    SKG_TRY(describe(q[i], &ps[i], &Ml, &Nl, false));
    Problem& p = ps[i];
    p.M = q[i].kind == 1 ? q[i].N : q[i].M;
    p.N = q[i].N;
    p.K = q[i].K;
    p.alpha = q[i].alpha;
    p.beta = q[i].beta;
    p.gamma = q[i].kind == 1 ? 0.0 : q[i].gamma;
  }
  auto run_all = [&]() -> int {
    for (int i = 0; i < n; ++i) {
      if (ps[i].M == 0 || ps[i].N == 0) continue;
      const int lb = q[i].kind == 1 ? q[i].layout_a : q[i].layout_b;
      SKG_TRY(ozaki_run(c, q[i].kind, q[i].layout_a, lb, ps[i], slices,
                        nullptr));
    }
    return SKGEMM_OK;
  };
  if (!use_graph) return run_all();

  // Cache key: everything baked into the captured launches.
  std::vector<int64_t> key;
  key.reserve((size_t)n * 16 + 4);
  key.push_back(slices);
  key.push_back(c->forced_grid);
  for (int i = 0; i < n; ++i) {
    const skgemm_problem& r = q[i];
    int64_t a;
    const double x[3] = {r.alpha, r.beta, r.gamma};
    key.insert(key.end(), {r.kind, r.layout_a, r.layout_b, r.layout_d, r.M,
                           r.N, r.K, (int64_t)(intptr_t)r.A, r.lda,
                           (int64_t)(intptr_t)r.B, r.ldb,
                           (int64_t)(intptr_t)r.D, r.ldd});
    for (double v : x) {
      std::memcpy(&a, &v, sizeof(a));
      key.push_back(a);
    }
  }
  SKG_TRY(batch_state(c, 1));
  auto* b = c->batch;
  auto it = b->split.find(key);
  if (it != b->split.end() && it->second.gen != c->mem_gen) {
    SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    skgemm_ctx::Batch::drop(it->second);
    b->split.erase(it);
    it = b->split.end();
  }
  if (it == b->split.end()) {
    for (int i = 0; i < n; ++i) {
      if (ps[i].M == 0 || ps[i].N == 0) continue;
      const int lb = q[i].kind == 1 ? q[i].layout_a : q[i].layout_b;
      SKG_TRY(ozaki_reserve(c, q[i].kind, q[i].layout_a, lb, ps[i], slices));
    }
    SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (b->split.size() >= kMaxSplitGraphs) {
      skgemm_ctx::Batch::drop(b->split.begin()->second);
      b->split.erase(b->split.begin());
    }
    skgemm_ctx::Batch::Graph gr;
    SKG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed),
            "begin capture");
    const uint64_t gen0 = c->mem_gen;
    c->capturing = true;
    const int rc = run_all();
    c->capturing = false;
    const hipError_t ce = hipStreamEndCapture(c->stream, &gr.g);
    if (rc != SKGEMM_OK || ce != hipSuccess) {
      skgemm_ctx::Batch::drop(gr);
      if (rc != SKGEMM_OK) return rc;
      return hip_error(ce, "end capture");
    }
    if (c->mem_gen != gen0) {  // a buffer grew during capture: not replayable
      skgemm_ctx::Batch::drop(gr);
      return set_error(SKGEMM_ERR_STATE, "workspace grew during capture");
    }
    SKG_HIP(hipGraphInstantiate(&gr.ge, gr.g, nullptr, nullptr, 0),
            "graph instantiate");
    gr.gen = c->mem_gen;
    it = b->split.emplace(key, gr).first;
  }
  c->record_begin();
  SKG_HIP(hipGraphLaunch(it->second.ge, c->stream), "graph launch");
  c->record_end();
  return SKGEMM_OK;
}

}  // namespace

extern "C" int skgemm_batched(skgemm_ctx* c, const skgemm_problem* problems,
                              int count, int precision, int slices,
                              int use_graph) {
  SKG_TRY(check_ctx(c));
  if (count < 0 || (count > 0 && problems == nullptr))
    return set_error(SKGEMM_ERR_ARG, "problems is NULL or count < 0");
  if (count == 0) return SKGEMM_OK;
  if (c->capturing)
    return set_error(SKGEMM_ERR_STATE, "skgemm_batched inside a capture");
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  if (precision == SKGEMM_FP64) return run_plain(c, problems, count, use_graph);
  if (precision == SKGEMM_OZAKI_INT8)
    return run_split(c, problems, count, slices, use_graph);
  return set_error(SKGEMM_ERR_ARG, "precision must be SKGEMM_FP64 or "
                                   "SKGEMM_OZAKI_INT8");
}
