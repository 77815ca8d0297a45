// The context behind the opaque skgemm_ctx handle, and helpers shared by
// the host translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/skgemm.h"
#include "skgemm_internal.h"

namespace skg {
// Launch plan of one shape: tile configuration, persistent grid, K chunks.
struct Plan {
  int cfg;
  int grid;
  int kchunks;
};
}  // namespace skg

struct skgemm_ctx {
  static constexpr int kStages = 4;
  int magic = 0;
  int device = 0;
  int num_cus = 0, num_xcds = 8, lds_per_cu = 0, clock_mhz = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t last0 = nullptr, last1 = nullptr;  // pair of the last launch
  bool timed = false;
  // Optional ring of event pairs, one per launch (skgemm_timing_*).
  std::vector<std::pair<hipEvent_t, hipEvent_t>> tring;
  int tring_used = 0;
  void record_begin();
  void record_end();

  // Stream-K fix-up workspace: one partial tile per persistent workgroup.
  double* ws = nullptr;
  size_t ws_n = 0;  // doubles
  int* flags = nullptr;
  int flags_n = 0;
  int* status = nullptr;
  int64_t spin_limit = 1ll << 26;
  // Last-arriver fix-up (no waits between workgroups): twice the slots.
  bool waitfree = false;
  // Failure-path test hook: peers never raise their ready flags, so every
  // waiting owner runs out its (clamped) spin limit.
  bool fixup_hold = false;
  size_t ws_doubles(int cfg, int grid) const;
  // The spin limit a launch carries: at most 64 rounds while the hold
  // hook is on, so a held wait ends within microseconds.
  int64_t fixup_spin_limit() const {
    return fixup_hold && spin_limit > 64 ? 64 : spin_limit;
  }
  // SYRK outputs: SKGEMM_SYRK_LOWER or SKGEMM_SYRK_DIAG_BLOCKS.
  int syrk_store = SKGEMM_SYRK_LOWER;
  // Per-chunk sums of K-chunked launches.
  double* part = nullptr;
  size_t part_n = 0;

  // Device staging for host-pointer calls and probes.
  void* stage[kStages] = {nullptr, nullptr, nullptr, nullptr};
  size_t stage_n[kStages] = {0, 0, 0, 0};

  // Tuning: forced configuration, autotuned plans per exact shape
  // (kind, layout_a, layout_b, M, N, K) -> plan, occupancy per cfg.
  int forced_cfg = -1;
  int forced_grid = 0;
  int forced_kchunks = 0;
  std::map<std::tuple<int, int, int, int, int, int>, skg::Plan> tuned;
  std::map<int, int> occ_cache;
  // Cost-weighted Stream-K splits (device tables of grid + 1 first
  // iterations) per problem shape and plan; kept until the context ends, so
  // a captured graph may hold one.
  std::map<std::vector<int64_t>, int64_t*> split_tables;

  // Split-form workspace (skgemm_ozaki.hip) and batched-launch state
  // (skgemm_batch.cpp).  While a launch sequence is being captured into a
  // hipGraph, `capturing` suppresses timing-ring records.
  struct Extra;
  Extra* extra = nullptr;
  struct Batch;
  Batch* batch = nullptr;
  bool capturing = false;
  // Bumped whenever a workspace buffer is reallocated: captured graphs that
  // hold the old addresses are recaptured.
  uint64_t mem_gen = 0;
  // Split form: never pick a schedule whose workgroups wait on each other
  // (set on serving slots, whose launches run concurrently).
  bool oz_no_wait = false;
  // Modulus count of the last split-form call (0: computed in plain FP64).
  int oz_last_moduli = 0;
  // skgemm_timing_stages: split-form calls record one event pair per stage.
  bool stage_timing = false;

  int check_status();
  int ensure_ws(size_t ws_doubles, int nflags);
  int ensure_part(size_t doubles);
  double* staging(int i, size_t bytes);
  int capacity(int cfg);
  void free_all();
  void free_extra();
  void free_batch();
};

// Early return of an error code / of a failed HIP call as SKGEMM_ERR_HIP.
#define SKG_TRY(expr)                 \
  do {                                \
    int _rc = (expr);                 \
    if (_rc != SKGEMM_OK) return _rc; \
  } while (0)
#define SKG_HIP(expr, what)                                 \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return skg::hip_error(_e, what);  \
  } while (0)

namespace skg {

int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* what);

int make_operand(int layout, bool a_convention, const double* p, int64_t ld,
                 int ext, int K, Operand* op, const char* name);
int make_output(int layout, double* p, int64_t ld, int M, int N, Output* out,
                const char* name);
int64_t span_operand(int layout, bool a_convention, int64_t ld, int ext, int K);
int64_t span_output(int layout, int64_t ld, int M, int N);
int check_dev(const void* p, const char* name);
// Split-form driver (skgemm_ozaki.hip).  kind 0: GEMM, 1: SYRK; pr holds
// operands, output, sizes and alpha/beta/gamma.  ozaki_reserve grows every
// buffer the call would use, so that a later call allocates nothing (for
// graph capture).
int ozaki_run(skgemm_ctx* c, int kind, int la, int lb, const Problem& pr,
              int slices, double* err_bound);
int ozaki_reserve(skgemm_ctx* c, int kind, int la, int lb, const Problem& pr,
                  int slices);

int64_t tiles_for(int kind, int cfg, int M, int N);
Plan plan(skgemm_ctx* c, int kind, int la, int lb, int M, int N, int K);
// A GEMM whose D is column-major is launched as its transpose, D^T = B^T A^T
// with D^T row-major (the same memory): the epilogue then stores contiguous
// 128-byte rows instead of one 8-byte element per column.  Rewrites the
// arguments in place and returns true when it did.
bool transpose_launch(int& la, int& lb, int& ld_, int& M, int& N,
                      const double*& A, int64_t& lda, const double*& B,
                      int64_t& ldb);
int check_ctx(skgemm_ctx* c);
int check_dims(int M, int N, int K);
// alloc_part = false leaves p->part NULL when the plan has K chunks (the
// caller provides the chunk buffer).
int fill_problem(skgemm_ctx* c, Problem* p, int kind, const Plan& pl, int M,
                 int N, int K, bool alloc_part = true);
// pred: nullptr, or a device flag; the launch does its work only when the
// flag is non-zero (decided on the device).  timed: record the launch in
// the context's timing ring.
// A plain FP64 batch described once (tile configuration, persistent grid,
// problem rows with their iteration ranges) for callers that keep the rows
// in device memory of their own and launch them many times.
struct PlainTable {
  int cfg = 0, grid = 0;
  int64_t iters = 0;
  std::vector<Problem> rows;
};
int build_plain_table(skgemm_ctx* c, const skgemm_problem* q, int n,
                      PlainTable* out);
// dtab: the rows in device memory; dhdr: device {row count, iterations}.
int launch_plain_table(skgemm_ctx* c, const PlainTable& t, const Problem* dtab,
                       const int64_t* dhdr);
// Fix-up buffers a caller provides instead of the context's (no
// allocation, no weighted-split table: the even split).
struct CallerBuffers {
  double* ws;
  size_t ws_doubles;
  int* flags;
  int nflags;
};
int run_single(skgemm_ctx* c, Problem& p, const Plan& pl,
               const int* pred = nullptr, bool timed = true,
               const CallerBuffers* cb = nullptr);
// Looks up the device address of the allocation-free entry point's scratch
// array on `device` (current), once per device (skgemm_noalloc.hip).
int noalloc_init(int device);
int prepare_gemm(int la, int lb, int ld_, int M, int N, int K, const double* A,
                 int64_t lda, const double* B, int64_t ldb, double* D,
                 int64_t ldd, Problem* p);
int prepare_syrk(int la, int N, int K, const double* A, int64_t lda, double* C,
                 int64_t ldc, Problem* p);

}  // namespace skg
