// C ABI of the FP64 Stream-K GEMM / SYRK library: contexts, argument
// checking, operand descriptors, launch planning and the host-pointer
// wrappers.  Kernels live in skgemm_kernels.hip (plain FP64) and
// skgemm_ozaki.hip (split form).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/skgemm.h"
#include "skgemm_ctx.h"
#include "skgemm_internal.h"

using namespace skg;

namespace skg {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_error(hipError_t e, const char* what) {
  return set_error(SKGEMM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// ---- operand descriptors --------------------------------------------------

static bool valid_layout(int l) { return l >= SKGEMM_ROW_MAJOR && l <= SKGEMM_PACK32; }
static int panel_shift(int l) { return l == SKGEMM_PACK16 ? 4 : 5; }

// Input operand with free index o (extent `ext`) and contraction index k
// (extent K).  `a_convention`: A (ROW_MAJOR = o*ld + k) vs B (ROW_MAJOR =
// k*ld + o).
int make_operand(int layout, bool a_convention, const double* p, int64_t ld,
                 int ext, int K, Operand* op, const char* name) {
  if (!valid_layout(layout))
    return set_error(SKGEMM_ERR_ARG, "%s: unknown layout %d", name, layout);
  Operand o{};
  o.ptr = p;
  o.extent = ext;
  bool o_major;  // o varies fastest
  if (layout == SKGEMM_ROW_MAJOR) o_major = !a_convention;
  else if (layout == SKGEMM_COL_MAJOR) o_major = a_convention;
  else o_major = true;
  if (layout == SKGEMM_PACK16 || layout == SKGEMM_PACK32) {
    const int s = panel_shift(layout);
    if (ld < std::max(K, 1))
      return set_error(SKGEMM_ERR_ARG, "%s: packed ld %lld < K %d", name,
                       (long long)ld, K);
    o.pshift = s;
    o.pstride = ld << s;
    o.os = 1;
    o.ks = 1ll << s;
    o.kcontig = 0;
  } else if (o_major) {
    if (ld < std::max(ext, 1))
      return set_error(SKGEMM_ERR_ARG, "%s: ld %lld < %d", name, (long long)ld,
                       ext);
    o.pshift = kNoPanelShift;
    o.pstride = 0;
    o.os = 1;
    o.ks = ld;
    o.kcontig = 0;
  } else {
    if (ld < std::max(K, 1))
      return set_error(SKGEMM_ERR_ARG, "%s: ld %lld < K %d", name,
                       (long long)ld, K);
    o.pshift = kNoPanelShift;
    o.pstride = 0;
    o.os = ld;
    o.ks = 1;
    o.kcontig = 1;
  }
  // 16-byte pair loads need a 16-byte aligned pair start everywhere and an
  // even extent along the contiguous index (a pair never straddles the edge).
  const bool aligned = (reinterpret_cast<uintptr_t>(p) & 15) == 0;
  o.vec2 = aligned && (o.kcontig ? (o.os % 2 == 0 && K % 2 == 0)
                                 : (o.ks % 2 == 0 && o.pstride % 2 == 0 &&
                                    ext % 2 == 0));
  // Per-thread tile offsets are 32-bit byte offsets from the tile origin
  // (largest tile: 256 along o, 32 along k).
  const int64_t pm = (1ll << o.pshift) - 1;
  const int64_t max_rel =
      ((255 >> o.pshift) * o.pstride + std::min<int64_t>(255, pm) * o.os +
       31 * o.ks + 2) * 8;
  if (max_rel >= 0x7FFFFF00ll)
    return set_error(SKGEMM_ERR_NOTSUP,
                     "%s: leading dimension %lld too large for 32-bit tile "
                     "offsets", name, (long long)ld);
  *op = o;
  return SKGEMM_OK;
}

// Output D (m, n), M x N.
int make_output(int layout, double* p, int64_t ld, int M, int N, Output* out,
                const char* name) {
  if (!valid_layout(layout))
    return set_error(SKGEMM_ERR_ARG, "%s: unknown layout %d", name, layout);
  Output d{};
  d.ptr = p;
  if (layout == SKGEMM_ROW_MAJOR) {
    if (ld < std::max(N, 1))
      return set_error(SKGEMM_ERR_ARG, "%s: ld %lld < N %d", name, (long long)ld, N);
    d.pshift = kNoPanelShift; d.pstride = 0; d.ns = 1; d.ms = ld;
  } else if (layout == SKGEMM_COL_MAJOR) {
    if (ld < std::max(M, 1))
      return set_error(SKGEMM_ERR_ARG, "%s: ld %lld < M %d", name, (long long)ld, M);
    d.pshift = kNoPanelShift; d.pstride = 0; d.ns = ld; d.ms = 1;
  } else {
    if (ld < std::max(M, 1))
      return set_error(SKGEMM_ERR_ARG, "%s: packed ld %lld < M %d", name,
                       (long long)ld, M);
    const int s = panel_shift(layout);
    d.pshift = s; d.pstride = ld << s; d.ns = 1; d.ms = 1ll << s;
  }
  *out = d;
  return SKGEMM_OK;
}

// Number of elements from the base pointer to one past the last element an
// operand of this layout touches (0 if empty).
int64_t span_operand(int layout, bool a_convention, int64_t ld, int ext, int K) {
  if (ext <= 0 || K <= 0) return 0;
  if (layout == SKGEMM_PACK16 || layout == SKGEMM_PACK32) {
    const int s = panel_shift(layout);
    const int64_t P = 1ll << s;
    return ((int64_t)(ext - 1) >> s) * ld * P + (int64_t)(K - 1) * P +
           ((ext - 1) & (P - 1)) + 1;
  }
  const bool o_major = (layout == SKGEMM_ROW_MAJOR) ? !a_convention : a_convention;
  return o_major ? (int64_t)(K - 1) * ld + ext : (int64_t)(ext - 1) * ld + K;
}
int64_t span_output(int layout, int64_t ld, int M, int N) {
  if (M <= 0 || N <= 0) return 0;
  if (layout == SKGEMM_ROW_MAJOR) return (int64_t)(M - 1) * ld + N;
  if (layout == SKGEMM_COL_MAJOR) return (int64_t)(N - 1) * ld + M;
  const int s = panel_shift(layout);
  const int64_t P = 1ll << s;
  return ((int64_t)(N - 1) >> s) * ld * P + (int64_t)(M - 1) * P +
         ((N - 1) & (P - 1)) + 1;
}

// Device-pointer entry points reject pointers the GPU cannot dereference
// (pageable host memory) with SKGEMM_ERR_ARG instead of faulting the device.
int check_dev(const void* p, const char* name) {
  if (p == nullptr) return SKGEMM_OK;
  hipPointerAttribute_t a;
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(SKGEMM_ERR_ARG,
                     "%s is not device-accessible memory (use the _host "
                     "entry point for host arrays)", name);
  }
  if (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged &&
      a.type != hipMemoryTypeHost)
    return set_error(SKGEMM_ERR_ARG, "%s is not device-accessible memory",
                     name);
  return SKGEMM_OK;
}

}  // namespace skg

// ---- context ----------------------------------------------------------------

static const int kCtxMagic = 0x5ca1ab1e;

int skg::check_ctx(skgemm_ctx* c) {
  if (c == nullptr || c->magic != kCtxMagic)
    return set_error(SKGEMM_ERR_STATE, "invalid or destroyed context");
  return SKGEMM_OK;
}

extern "C" const char* skgemm_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* skgemm_version(void) { return "skgemm 0.1 (gfx950)"; }

extern "C" int skgemm_create(int device, skgemm_ctx** out) {
  if (out == nullptr) return set_error(SKGEMM_ERR_ARG, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  SKG_HIP(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
  if (device < 0 || device >= ndev)
    return set_error(SKGEMM_ERR_ARG, "device %d out of range (%d devices)",
                     device, ndev);
  SKG_HIP(hipSetDevice(device), "hipSetDevice");
  hipDeviceProp_t prop;
  SKG_HIP(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(SKGEMM_ERR_NOTSUP,
                     "device %d is %s; this library is built for gfx950 only",
                     device, prop.gcnArchName);
  skgemm_ctx* c = new skgemm_ctx();
  c->magic = kCtxMagic;
  c->device = device;
  c->num_cus = prop.multiProcessorCount;
  c->num_xcds = 8;
  c->lds_per_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
  c->clock_mhz = prop.clockRate / 1000;
  hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_error(e, "hipStreamCreate");
  }
  c->stream = c->own_stream;
  // Timing-only events (no system-scope fence on record: a fenced record
  // also writes back and invalidates the caches between the launches it
  // brackets, which slowed the next launch measurably).
  if ((e = hipEventCreateWithFlags(&c->ev0, hipEventDisableSystemFence)) !=
          hipSuccess ||
      (e = hipEventCreateWithFlags(&c->ev1, hipEventDisableSystemFence)) !=
          hipSuccess ||
      (e = hipMalloc(&c->status, sizeof(int))) != hipSuccess ||
      (e = launch_zero_ints(c->status, 1, c->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(c->stream)) != hipSuccess) {
    skgemm_destroy(c);
    return hip_error(e, "context setup");
  }
  c->last0 = c->ev0;
  c->last1 = c->ev1;
  c->spin_limit = 1ll << 26;
  // Device address of the allocation-free entry point's scratch array on
  // this device (once per device; a failure is reported by that entry point).
  noalloc_init(device);
  *out = c;
  return SKGEMM_OK;
}

extern "C" int skgemm_destroy(skgemm_ctx* c) {
  SKG_TRY(check_ctx(c));
  hipSetDevice(c->device);
  if (c->own_stream) hipStreamSynchronize(c->own_stream);
  if (c->stream && c->stream != c->own_stream) hipStreamSynchronize(c->stream);
  c->free_all();
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->own_stream) hipStreamDestroy(c->own_stream);
  c->magic = 0;
  delete c;
  return SKGEMM_OK;
}

extern "C" int skgemm_set_stream(skgemm_ctx* c, void* stream) {
  SKG_TRY(check_ctx(c));
  c->stream = stream ? reinterpret_cast<hipStream_t>(stream) : c->own_stream;
  return SKGEMM_OK;
}

extern "C" int skgemm_synchronize(skgemm_ctx* c) {
  SKG_TRY(check_ctx(c));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return c->check_status();
}

extern "C" int skgemm_device_info(skgemm_ctx* c, int* num_cus, int* num_xcds,
                                  int* lds, int* clock_mhz) {
  SKG_TRY(check_ctx(c));
  if (num_cus) *num_cus = c->num_cus;
  if (num_xcds) *num_xcds = c->num_xcds;
  if (lds) *lds = c->lds_per_cu;
  if (clock_mhz) *clock_mhz = c->clock_mhz;
  return SKGEMM_OK;
}

extern "C" int skgemm_set_waitfree(skgemm_ctx* c, int on) {
  SKG_TRY(check_ctx(c));
  if (on != 0 && on != 1) return set_error(SKGEMM_ERR_ARG, "on must be 0 or 1");
  c->waitfree = on != 0;
  return SKGEMM_OK;
}

extern "C" int skgemm_set_spin_limit(skgemm_ctx* c, int64_t limit) {
  SKG_TRY(check_ctx(c));
  if (limit < 1) return set_error(SKGEMM_ERR_ARG, "spin limit must be >= 1");
  c->spin_limit = limit;
  return SKGEMM_OK;
}

extern "C" int skgemm_set_fixup_hold(skgemm_ctx* c, int on) {
  SKG_TRY(check_ctx(c));
  if (on != 0 && on != 1) return set_error(SKGEMM_ERR_ARG, "on must be 0 or 1");
  c->fixup_hold = on != 0;
  return SKGEMM_OK;
}

extern "C" int skgemm_set_syrk_store(skgemm_ctx* c, int mode) {
  SKG_TRY(check_ctx(c));
  if (mode != SKGEMM_SYRK_LOWER && mode != SKGEMM_SYRK_DIAG_BLOCKS)
    return set_error(SKGEMM_ERR_ARG, "SYRK store mode must be %d or %d",
                     SKGEMM_SYRK_LOWER, SKGEMM_SYRK_DIAG_BLOCKS);
  c->syrk_store = mode;
  return SKGEMM_OK;
}

// ---- memory -----------------------------------------------------------------

extern "C" int skgemm_dev_alloc(skgemm_ctx* c, size_t bytes, int zero,
                                void** out) {
  SKG_TRY(check_ctx(c));
  if (out == nullptr) return set_error(SKGEMM_ERR_ARG, "out is NULL");
  *out = nullptr;
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess)
    return set_error(SKGEMM_ERR_NOMEM, "hipMalloc(%zu): %s", bytes,
                     hipGetErrorString(e));
  if (zero) {
    // On the context's stream: a null-stream memset is not ordered against
    // the context's non-blocking stream.
    hipError_t ez = hipMemsetAsync(p, 0, bytes, c->stream);
    if (ez == hipSuccess) ez = hipStreamSynchronize(c->stream);
    if (ez != hipSuccess) {
      hipFree(p);
      return hip_error(ez, "zero-fill");
    }
  }
  *out = p;
  return SKGEMM_OK;
}

extern "C" int skgemm_dev_free(skgemm_ctx* c, void* p) {
  SKG_TRY(check_ctx(c));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  if (p) SKG_HIP(hipFree(p), "hipFree");
  return SKGEMM_OK;
}

extern "C" int skgemm_memcpy_h2d(skgemm_ctx* c, void* dst, const void* src,
                                 size_t bytes) {
  SKG_TRY(check_ctx(c));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  SKG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream),
          "hipMemcpyAsync H2D");
  SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return SKGEMM_OK;
}

extern "C" int skgemm_memcpy_d2h(skgemm_ctx* c, void* dst, const void* src,
                                 size_t bytes) {
  SKG_TRY(check_ctx(c));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  SKG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream),
          "hipMemcpyAsync D2H");
  SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return SKGEMM_OK;
}

// ---- skgemm_ctx members -------------------------------------------------------

// Status readback on the context's own stream (called after it has been
// synchronized, never inside a capture): operations on the legacy stream
// are refused while another thread captures a graph.
int skgemm_ctx::check_status() {
  int h = 0;
  SKG_HIP(hipMemcpyAsync(&h, status, sizeof(int), hipMemcpyDeviceToHost,
                         stream),
          "status readback");
  SKG_HIP(hipStreamSynchronize(stream), "status readback");
  if (h & 2) {
    launch_zero_ints(status, 1, stream);
    hipStreamSynchronize(stream);
    return set_error(SKGEMM_ERR_NOTSUP,
                     "internal: an LDS-DMA tile configuration was given an "
                     "operand without 16-byte aligned pairs");
  }
  if (h != 0) {
    launch_zero_ints(status, 1, stream);
    if (flags) launch_zero_ints(flags, flags_n, stream);
    hipStreamSynchronize(stream);
    return set_error(SKGEMM_ERR_TIMEOUT,
                     "a Stream-K fix-up wait exceeded its spin limit (%lld)",
                     (long long)spin_limit);
  }
  return SKGEMM_OK;
}

size_t skgemm_ctx::ws_doubles(int cfg, int grid) const {
  return (size_t)grid * slot_doubles(cfg) * (waitfree ? 2 : 1);
}

int skgemm_ctx::ensure_ws(size_t ws_doubles, int nflags) {
  if (ws_doubles > ws_n) {
    if (ws) {
      hipStreamSynchronize(stream);
      hipFree(ws);
      ws = nullptr;
    }
    ++mem_gen;
    hipError_t e = hipMalloc(&ws, ws_doubles * sizeof(double));
    if (e != hipSuccess) {
      ws_n = 0;
      return set_error(SKGEMM_ERR_NOMEM, "workspace of %zu doubles: %s",
                       ws_doubles, hipGetErrorString(e));
    }
    ws_n = ws_doubles;
  }
  if (nflags > flags_n) {
    if (flags) {
      hipStreamSynchronize(stream);
      hipFree(flags);
      flags = nullptr;
    }
    ++mem_gen;
    hipError_t e = hipMalloc(&flags, sizeof(int) * nflags);
    if (e != hipSuccess) {
      flags_n = 0;
      return set_error(SKGEMM_ERR_NOMEM, "flags: %s", hipGetErrorString(e));
    }
    // Flags start at zero: a fill kernel on this stream, completed before
    // returning (a later call may run on another stream); inside a capture
    // it becomes a graph node ahead of the launch, harmless on replays since
    // the flags are zero between launches.  Every consumer resets the flag
    // it consumed, so they are zero again at the end of every launch.
    SKG_HIP(launch_zero_ints(flags, nflags, stream), "flags init");
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (!capturing && hipStreamIsCapturing(stream, &cs) == hipSuccess &&
        cs == hipStreamCaptureStatusNone)
      SKG_HIP(hipStreamSynchronize(stream), "flags init");
    flags_n = nflags;
  }
  return SKGEMM_OK;
}

int skgemm_ctx::ensure_part(size_t doubles) {
  if (doubles > part_n) {
    if (part) {
      hipStreamSynchronize(stream);
      hipFree(part);
      part = nullptr;
      part_n = 0;
    }
    hipError_t e = hipMalloc(&part, doubles * sizeof(double));
    if (e != hipSuccess)
      return set_error(SKGEMM_ERR_NOMEM, "K-chunk buffer of %zu doubles: %s",
                       doubles, hipGetErrorString(e));
    part_n = doubles;
  }
  return SKGEMM_OK;
}

double* skgemm_ctx::staging(int i, size_t bytes) {
  if (bytes > stage_n[i]) {
    if (stage[i]) {
      hipStreamSynchronize(stream);
      hipFree(stage[i]);
      stage[i] = nullptr;
      stage_n[i] = 0;
    }
    if (hipMalloc(&stage[i], bytes) != hipSuccess) return nullptr;
    stage_n[i] = bytes;
  }
  return reinterpret_cast<double*>(stage[i]);
}

void skgemm_ctx::free_all() {
  for (auto& pr : tring) {
    hipEventDestroy(pr.first);
    hipEventDestroy(pr.second);
  }
  tring.clear();
  if (ws) hipFree(ws);
  if (part) hipFree(part);
  part = nullptr;
  if (flags) hipFree(flags);
  if (status) hipFree(status);
  for (int i = 0; i < kStages; ++i)
    if (stage[i]) hipFree(stage[i]);
  ws = nullptr;
  flags = nullptr;
  status = nullptr;
  free_extra();
  free_batch();
  for (auto& kv : split_tables) hipFree(kv.second);
  split_tables.clear();
}

// One event pair per launch: the timing ring's next pair while the ring has
// room, else ev0 / ev1 (a second pair at the same boundaries would only add
// device time between launches).  last0 / last1 name the pair of the most
// recent launch.
void skgemm_ctx::record_begin() {
  if (capturing) return;
  last0 = tring_used < (int)tring.size() ? tring[tring_used].first : ev0;
  hipEventRecord(last0, stream);
}

void skgemm_ctx::record_end() {
  if (capturing) return;
  if (tring_used < (int)tring.size()) {
    last1 = tring[tring_used].second;
    ++tring_used;
  } else {
    last1 = ev1;
  }
  hipEventRecord(last1, stream);
  timed = true;
}

int skgemm_ctx::capacity(int cfg) {
  auto it = occ_cache.find(cfg);
  int occ;
  if (it == occ_cache.end()) {
    occ = occupancy_dgemm_sk(cfg);
    occ_cache[cfg] = occ;
  } else {
    occ = it->second;
  }
  return std::max(occ, 1) * num_cus;
}

// ---- planning ---------------------------------------------------------------

namespace skg {

int64_t tiles_for(int kind, int cfg, int M, int N) {
  const TileCfg& t = tile_cfg(cfg);
  if (kind == 1) {
    const int64_t T =
        strip_rows(cfg, kind, N, N) ? N / t.bm : (N + t.bm - 1) / t.bm;
    return T * (T + 1) / 2;
  }
  const int tn = strip_cols(cfg, kind, N) ? N / t.bn : (N + t.bn - 1) / t.bn;
  const int tm =
      strip_rows(cfg, kind, M, N) ? M / t.bm : (M + t.bm - 1) / t.bm;
  return (int64_t)tm * tn;
}

static int64_t iters_per_tile(int cfg, int K) {
  const TileCfg& t = tile_cfg(cfg);
  return std::max<int64_t>(1, (K + t.bk - 1) / t.bk);
}

// K chunks: with fewer tiles than CUs and a long K, one chunk per XCD so
// that the workgroups of an XCD stream the same rows of the operands
// through its L2 for all tiles; otherwise one chunk.
static int default_kchunks(skgemm_ctx* c, int kind, int cfg, int M, int N,
                           int K) {
  const int64_t tiles = tiles_for(kind, cfg, M, N);
  const int64_t ipt = iters_per_tile(cfg, K);
  return (tiles < c->num_cus && ipt >= 64 * c->num_xcds) ? c->num_xcds : 1;
}

// Default tile when the autotuner has no entry for this shape.
// 128 x 128 tiles use the 4-stage LDS-DMA variant (cfg 8; the host takes
// its register-staged sibling, cfg 4, for unaligned operands).
static int default_cfg(skgemm_ctx* c, int kind, int M, int N, int K) {
  (void)K;
  if (kind == 1) return (N >= 768) ? 8 : 1;
  if (M >= 768 && N >= 768) {
    const int64_t t128 = tiles_for(0, 4, M, N);
    return (t128 >= 4 * c->num_cus) ? 8 : 6;
  }
  if (M >= 128 && N >= 128) return 1;
  if (M >= 128) return 2;
  if (N >= 128) return 3;
  return 1;
}

// Whole XCD rounds, except for a grid of exactly one (K chunk, tile) unit
// per workgroup, which is kept (data-parallel: no fix-up).
static int clamp_grid(skgemm_ctx* c, int cfg, int g, int64_t total,
                      int64_t units) {
  const int cap = c->capacity(cfg);
  g = std::min(g, cap);
  if (g >= 16 && g != units) g &= ~7;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, total));
}

Plan plan(skgemm_ctx* c, int kind, int la, int lb, int M, int N, int K) {
  Plan pl{};
  if (kind == 1) lb = la;
  auto it = c->tuned.find(std::make_tuple(kind, la, lb, M, N, K));
  if (c->forced_cfg >= 0) {
    pl.cfg = c->forced_cfg;
    if (kind == 1 && tile_cfg(pl.cfg).bm != tile_cfg(pl.cfg).bn) pl.cfg = 1;
    pl.grid = c->forced_grid > 0 ? c->forced_grid : c->capacity(pl.cfg);
    pl.kchunks = c->forced_kchunks > 0
                     ? c->forced_kchunks
                     : default_kchunks(c, kind, pl.cfg, M, N, K);
  } else if (it != c->tuned.end()) {
    pl = it->second;
  } else {
    pl.cfg = default_cfg(c, kind, M, N, K);
    pl.grid = c->capacity(pl.cfg);
    pl.kchunks = default_kchunks(c, kind, pl.cfg, M, N, K);
    // Inner-product GEMM (few output tiles, long K, e.g. the projection
    // Psi^T (H Psi)): 128 x 128 tiles, as many K chunks as fill the CUs
    // once, one workgroup per (tile, chunk).
    const int64_t t4 = tiles_for(0, 4, M, N);
    if (kind == 0 && M >= 256 && N >= 256 && 2 * t4 <= c->num_cus &&
        iters_per_tile(4, K) >= 64 * c->num_xcds) {
      pl.cfg = 8;
      pl.kchunks = (int)std::min<int64_t>(16, std::max<int64_t>(
                                                  c->num_xcds, c->num_cus / t4));
      pl.grid = (int)std::min<int64_t>(t4 * pl.kchunks, c->capacity(8));
    }
  }
  if (c->forced_grid > 0) pl.grid = c->forced_grid;
  if (c->forced_kchunks > 0) pl.kchunks = c->forced_kchunks;
  const int64_t ipt = iters_per_tile(pl.cfg, K);
  pl.kchunks = (int)std::max<int64_t>(1, std::min<int64_t>(pl.kchunks, ipt));
  const int64_t kci = (ipt + pl.kchunks - 1) / pl.kchunks;
  const int64_t units = tiles_for(kind, pl.cfg, M, N) * pl.kchunks;
  pl.grid = clamp_grid(c, pl.cfg, pl.grid, units * kci, units);
  return pl;
}

// Fill a Problem for the Stream-K kernel (tile counts depend on the cfg).
int fill_problem(skgemm_ctx* c, Problem* p, int kind, const Plan& pl, int M,
                 int N, int K, bool alloc_part) {
  const TileCfg& t = tile_cfg(pl.cfg);
  p->M = M;
  p->N = N;
  p->K = K;
  p->syrk = kind == 1 ? (c->syrk_store ? 2 : 1) : 0;
  p->xcol = strip_cols(pl.cfg, kind, N);
  p->xrow = strip_rows(pl.cfg, kind, M, N);
  p->tiles_m = p->xrow ? M / t.bm : (M + t.bm - 1) / t.bm;
  p->tiles_n = p->xcol ? N / t.bn : (N + t.bn - 1) / t.bn;
  if (kind == 1) p->tiles_n = p->tiles_m;
  p->iters_per_tile = (int)iters_per_tile(pl.cfg, K);
  // Tile order: keep the smaller operand cached and read the larger once.
  p->row_order = (kind == 0 && (int64_t)M >= (int64_t)N) ? 1 : 0;
  // Short M, long N, no strip: bands of ceil(tiles_n / 8) tile columns, so
  // that each XCD's share of the tiles is one band and the tile rows that
  // read one B panel run on one XCD together (c3 rotation 2302 -> 2196 us).
  // With a strip the column order stays: each workgroup's tiles are then a
  // whole column, strip tile included, and the split needs no fix-up.
  if (kind == 0 && p->row_order == 0 && p->tiles_m > 1 && p->tiles_n >= 16 &&
      p->xrow == 0 && p->xcol == 0)
    p->row_order = (p->tiles_n + 7) / 8;
  p->kchunks = pl.kchunks;
  p->kci = (p->iters_per_tile + pl.kchunks - 1) / pl.kchunks;
  p->num_tiles = tiles_for(kind, pl.cfg, M, N);
  p->part = nullptr;
  if (pl.kchunks > 1 && alloc_part) {
    SKG_TRY(c->ensure_part((size_t)pl.kchunks * M * N));
    p->part = c->part;
  }
  return SKGEMM_OK;
}

// Cost-weighted Stream-K split for problems with strip tiles.  A strip
// adds one MFMA block per wave and sub-step to its tile (two for wave 0 of
// a SYRK's corner tile), so its iterations take (wb + 1) / wb of a plain
// tile's (wb = blocks per wave); an even split of iterations would leave
// the workgroups that hold strip tiles on the critical path.  Workgroup w
// starts at the first iteration whose weighted prefix reaches w / G of the
// total.  The table is computed once per shape and plan and kept on the
// device; while a stream capture is active and no table exists yet, the
// even split is used (no synchronous copy inside a capture).
static const int64_t* split_table(skgemm_ctx* c, const Problem& p, int cfg,
                                  int G) {
  if (p.xrow == 0 && p.xcol == 0) return nullptr;
  // One whole unit per workgroup: the even split is already aligned to the
  // units (no fix-up, and the workgroups of one XCD advance through K
  // together, sharing panels in its L2).
  if (p.num_tiles * p.kchunks == G) return nullptr;
  const std::vector<int64_t> key = {p.syrk,     p.M,        p.N,
                                    p.K,        cfg,        G,
                                    p.kchunks,  p.kci,      p.row_order,
                                    p.xrow,     p.xcol,     p.tiles_m,
                                    p.tiles_n};
  auto found = c->split_tables.find(key);
  if (found != c->split_tables.end()) return found->second;
  if (c->capturing || c->split_tables.size() >= 64) return nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess ||
      cs != hipStreamCaptureStatusNone)
    return nullptr;
  const TileCfg& t = tile_cfg(cfg);
  const int wb = (t.bm / t.wm / 16) * (t.bn / t.wn / 16);
  std::vector<int> wt((size_t)p.num_tiles);
  for (int64_t tile = 0; tile < p.num_tiles; ++tile) {
    int ti, tj;
    tile_coords(tile, p.syrk, p.row_order, p.tiles_m, p.tiles_n, ti, tj);
    const bool strip = (p.xrow && ti == p.tiles_m - 1) ||
                       (p.xcol && tj == p.tiles_n - 1);
    wt[tile] = wb + (strip ? 1 : 0) + (strip && p.syrk && ti == tj ? 1 : 0);
  }
  const int64_t units = p.num_tiles * p.kchunks;
  const int64_t total = units * p.kci;
  double per_chunk = 0.0;
  for (int w : wt) per_chunk += w;
  const double W = per_chunk * p.kchunks * p.kci;
  std::vector<int64_t> st((size_t)G + 1);
  st[0] = 0;
  st[G] = total;
  int64_t u = 0;
  double cum = 0.0;  // weighted prefix at the start of unit u
  for (int w = 1; w < G; ++w) {
    const double target = W * (double)w / (double)G;
    while (u < units &&
           cum + (double)wt[u % p.num_tiles] * p.kci <= target) {
      cum += (double)wt[u % p.num_tiles] * p.kci;
      ++u;
    }
    int64_t i = u * p.kci;
    if (u < units) {
      const double wu = wt[u % p.num_tiles];
      i += std::min<int64_t>(p.kci, (int64_t)std::ceil((target - cum) / wu));
    }
    st[w] = std::max(st[w - 1], std::min(i, total));
  }
  int64_t* d = nullptr;
  if (hipMalloc(&d, sizeof(int64_t) * st.size()) != hipSuccess) return nullptr;
  if (hipMemcpyAsync(d, st.data(), sizeof(int64_t) * st.size(),
                     hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    hipFree(d);
    return nullptr;
  }
  c->split_tables[key] = d;
  return d;
}

// Launch one problem (already described) with the given plan.
int run_single(skgemm_ctx* c, Problem& p, const Plan& plan_in,
               const int* pred, bool timed, const CallerBuffers* cb) {
  Plan pl = plan_in;
  if (tile_cfg(pl.cfg).ns >= 2 && !(p.a.vec2 && p.b.vec2)) {
    // LDS-DMA tiles move 16-byte pairs: unaligned operands take the
    // register-staged tile of the same shape (same tiles and units).
    pl.cfg = register_sibling(pl.cfg);
    pl.grid = std::max(1, std::min(pl.grid, c->capacity(pl.cfg)));
    if (p.xcol || p.xrow) {  // no strips there: plain tiling
      p.xcol = p.xrow = 0;
      p.tiles_m = (p.M + tile_cfg(pl.cfg).bm - 1) / tile_cfg(pl.cfg).bm;
      p.tiles_n = (p.N + tile_cfg(pl.cfg).bn - 1) / tile_cfg(pl.cfg).bn;
      p.num_tiles = p.syrk ? (int64_t)p.tiles_m * (p.tiles_m + 1) / 2
                           : (int64_t)p.tiles_m * p.tiles_n;
    }
  }
  p.pred = pred;
  p.tile_begin = 0;
  p.iter_begin = 0;
  Launch L{};
  L.problems = nullptr;
  L.single = p;
  L.num_problems = 1;
  L.total_iters = p.num_tiles * p.kchunks * p.kci;
  L.grid = pl.grid;
  if (cb != nullptr) {
    // Buffers the caller provides (nothing is allocated): the even split,
    // whose starts need no device table.
    if (c->ws_doubles(pl.cfg, pl.grid) > cb->ws_doubles ||
        pl.grid > cb->nflags)
      return set_error(SKGEMM_ERR_NOMEM, "internal: caller buffers too small");
    L.starts = nullptr;
    L.ws = cb->ws;
    L.flags = cb->flags;
  } else {
    L.starts = split_table(c, p, pl.cfg, pl.grid);
    SKG_TRY(c->ensure_ws(c->ws_doubles(pl.cfg, pl.grid), pl.grid));
    L.ws = c->ws;
    L.flags = c->flags;
  }
  L.status = c->status;
  L.spin_limit = c->fixup_spin_limit();
  L.waitfree = c->waitfree ? 1 : 0;
  L.hold = c->fixup_hold ? 1 : 0;
  if (timed) c->record_begin();
  SKG_HIP(launch_dgemm_sk(pl.cfg, L, c->stream), "Stream-K launch");
  if (p.kchunks > 1)
    SKG_HIP(launch_chunk_reduce(p, c->stream), "chunk reduce launch");
  if (timed) c->record_end();
  return SKGEMM_OK;
}

}  // namespace skg

int skg::check_dims(int M, int N, int K) {
  if (M < 0 || N < 0 || K < 0)
    return set_error(SKGEMM_ERR_ARG, "negative dimension (M=%d N=%d K=%d)", M, N, K);
  if (M >= (1 << 30) || N >= (1 << 30) || K >= (1 << 30))
    return set_error(SKGEMM_ERR_ARG, "dimension >= 2^30");
  return SKGEMM_OK;
}

namespace skg {

// Operand layout of B in the A convention (element (n, k)) and of A in the
// B convention: plain layouts exchange names, panel layouts keep them.
static int flip_layout(int l) {
  if (l == SKGEMM_ROW_MAJOR) return SKGEMM_COL_MAJOR;
  if (l == SKGEMM_COL_MAJOR) return SKGEMM_ROW_MAJOR;
  return l;
}

bool transpose_launch(int& la, int& lb, int& ld_, int& M, int& N,
                      const double*& A, int64_t& lda, const double*& B,
                      int64_t& ldb) {
  if (ld_ != SKGEMM_COL_MAJOR || !valid_layout(la) || !valid_layout(lb))
    return false;
  const int la2 = flip_layout(lb), lb2 = flip_layout(la);
  la = la2;
  lb = lb2;
  ld_ = SKGEMM_ROW_MAJOR;
  std::swap(M, N);
  std::swap(A, B);
  std::swap(lda, ldb);
  return true;
}

int prepare_gemm(int la, int lb, int ld_, int M, int N, int K, const double* A,
                 int64_t lda, const double* B, int64_t ldb, double* D,
                 int64_t ldd, Problem* p) {
  SKG_TRY(check_dims(M, N, K));
  if (M > 0 && N > 0) {
    if (D == nullptr) return set_error(SKGEMM_ERR_ARG, "D is NULL");
    if (K > 0 && (A == nullptr || B == nullptr))
      return set_error(SKGEMM_ERR_ARG, "A or B is NULL");
  }
  SKG_TRY(make_operand(la, true, A, lda, M, K, &p->a, "A"));
  SKG_TRY(make_operand(lb, false, B, ldb, N, K, &p->b, "B"));
  SKG_TRY(make_output(ld_, D, ldd, M, N, &p->d, "D"));
  if (M > 0 && N > 0) {
    if (K > 0) {
      SKG_TRY(check_dev(A, "A"));
      SKG_TRY(check_dev(B, "B"));
    }
    SKG_TRY(check_dev(D, "D"));
  }
  return SKGEMM_OK;
}

int prepare_syrk(int la, int N, int K, const double* A, int64_t lda, double* C,
                 int64_t ldc, Problem* p) {
  SKG_TRY(check_dims(N, N, K));
  if (N > 0) {
    if (C == nullptr) return set_error(SKGEMM_ERR_ARG, "C is NULL");
    if (K > 0 && A == nullptr) return set_error(SKGEMM_ERR_ARG, "A is NULL");
  }
  SKG_TRY(make_operand(la, false, A, lda, N, K, &p->a, "A"));
  p->b = p->a;
  SKG_TRY(make_output(SKGEMM_ROW_MAJOR, C, ldc, N, N, &p->d, "C"));
  if (N > 0) {
    if (K > 0) SKG_TRY(check_dev(A, "A"));
    SKG_TRY(check_dev(C, "C"));
  }
  return SKGEMM_OK;
}

}  // namespace skg

extern "C" int skgemm_dgemm(skgemm_ctx* c, int la, int lb, int ld_, int M,
                            int N, int K, double alpha, const double* A,
                            int64_t lda, const double* B, int64_t ldb,
                            double beta, double* D, int64_t ldd, double gamma) {
  SKG_TRY(check_ctx(c));
  transpose_launch(la, lb, ld_, M, N, A, lda, B, ldb);
  Problem p{};
  SKG_TRY(prepare_gemm(la, lb, ld_, M, N, K, A, lda, B, ldb, D, ldd, &p));
  if (M == 0 || N == 0) return SKGEMM_OK;
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  const Plan pl = plan(c, 0, la, lb, M, N, K);
  SKG_TRY(fill_problem(c, &p, 0, pl, M, N, K));
  p.alpha = alpha;
  p.beta = beta;
  p.gamma = gamma;
  return run_single(c, p, pl);
}

extern "C" int skgemm_dsyrk(skgemm_ctx* c, int la, int N, int K, double alpha,
                            const double* A, int64_t lda, double beta,
                            double* C, int64_t ldc) {
  SKG_TRY(check_ctx(c));
  Problem p{};
  SKG_TRY(prepare_syrk(la, N, K, A, lda, C, ldc, &p));
  if (N == 0) return SKGEMM_OK;
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  const Plan pl = plan(c, 1, la, la, N, N, K);
  SKG_TRY(fill_problem(c, &p, 1, pl, N, N, K));
  p.alpha = alpha;
  p.beta = beta;
  p.gamma = 0.0;
  return run_single(c, p, pl);
}

// ---- host-pointer wrappers ----------------------------------------------------

namespace {

// Copy host operands of a GEMM-shaped call to device staging, run `fn` on
// the device copies, copy D back; synchronous.
template <typename F>
int stage_gemm(skgemm_ctx* c, int la, int lb, int ld_, int M, int N, int K,
               const double* A, int64_t lda, const double* B, int64_t ldb,
               double* D, int64_t ldd, F fn) {
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  const int64_t na = span_operand(la, true, lda, M, K);
  const int64_t nb = span_operand(lb, false, ldb, N, K);
  const int64_t nd = span_output(ld_, ldd, M, N);
  double* dA = c->staging(0, std::max<int64_t>(na, 1) * 8);
  double* dB = c->staging(1, std::max<int64_t>(nb, 1) * 8);
  double* dD = c->staging(2, std::max<int64_t>(nd, 1) * 8);
  if (!dA || !dB || !dD) return set_error(SKGEMM_ERR_NOMEM, "staging buffers");
  if (na) SKG_HIP(hipMemcpyAsync(dA, A, na * 8, hipMemcpyHostToDevice, c->stream), "H2D A");
  if (nb) SKG_HIP(hipMemcpyAsync(dB, B, nb * 8, hipMemcpyHostToDevice, c->stream), "H2D B");
  // D is copied in even when beta == 0: packed / padded layouts have
  // elements outside the matrix that must come back unchanged.
  SKG_HIP(hipMemcpyAsync(dD, D, nd * 8, hipMemcpyHostToDevice, c->stream), "H2D D");
  SKG_TRY(fn(dA, dB, dD));
  SKG_HIP(hipMemcpyAsync(D, dD, nd * 8, hipMemcpyDeviceToHost, c->stream), "D2H D");
  SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return c->check_status();
}

template <typename F>
int stage_syrk(skgemm_ctx* c, int la, int N, int K, const double* A,
               int64_t lda, double* C, int64_t ldc, F fn) {
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  const int64_t na = span_operand(la, false, lda, N, K);
  const int64_t nc = span_output(SKGEMM_ROW_MAJOR, ldc, N, N);
  double* dA = c->staging(0, std::max<int64_t>(na, 1) * 8);
  double* dC = c->staging(2, std::max<int64_t>(nc, 1) * 8);
  if (!dA || !dC) return set_error(SKGEMM_ERR_NOMEM, "staging buffers");
  if (na) SKG_HIP(hipMemcpyAsync(dA, A, na * 8, hipMemcpyHostToDevice, c->stream), "H2D A");
  SKG_HIP(hipMemcpyAsync(dC, C, nc * 8, hipMemcpyHostToDevice, c->stream), "H2D C");
  SKG_TRY(fn(dA, dC));
  SKG_HIP(hipMemcpyAsync(C, dC, nc * 8, hipMemcpyDeviceToHost, c->stream), "D2H C");
  SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return c->check_status();
}

// Argument checks of the host forms (host pointers are not device memory,
// so the device-pointer checks are skipped here).
int host_args_gemm(int la, int lb, int ld_, int M, int N, int K,
                   const double* A, int64_t lda, const double* B, int64_t ldb,
                   double* D, int64_t ldd) {
  SKG_TRY(check_dims(M, N, K));
  if (M > 0 && N > 0) {
    if (D == nullptr) return set_error(SKGEMM_ERR_ARG, "D is NULL");
    if (K > 0 && (A == nullptr || B == nullptr))
      return set_error(SKGEMM_ERR_ARG, "A or B is NULL");
  }
  Operand o;
  Output d;
  SKG_TRY(make_operand(la, true, A, lda, M, K, &o, "A"));
  SKG_TRY(make_operand(lb, false, B, ldb, N, K, &o, "B"));
  SKG_TRY(make_output(ld_, D, ldd, M, N, &d, "D"));
  return SKGEMM_OK;
}

int host_args_syrk(int la, int N, int K, const double* A, int64_t lda,
                   double* C, int64_t ldc) {
  SKG_TRY(check_dims(N, N, K));
  if (N > 0) {
    if (C == nullptr) return set_error(SKGEMM_ERR_ARG, "C is NULL");
    if (K > 0 && A == nullptr) return set_error(SKGEMM_ERR_ARG, "A is NULL");
  }
  Operand o;
  Output d;
  SKG_TRY(make_operand(la, false, A, lda, N, K, &o, "A"));
  SKG_TRY(make_output(SKGEMM_ROW_MAJOR, C, ldc, N, N, &d, "C"));
  return SKGEMM_OK;
}

}  // namespace

extern "C" int skgemm_dgemm_host(skgemm_ctx* c, int la, int lb, int ld_,
                                 int M, int N, int K, double alpha,
                                 const double* A, int64_t lda, const double* B,
                                 int64_t ldb, double beta, double* D,
                                 int64_t ldd, double gamma) {
  SKG_TRY(check_ctx(c));
  SKG_TRY(host_args_gemm(la, lb, ld_, M, N, K, A, lda, B, ldb, D, ldd));
  if (M == 0 || N == 0) return SKGEMM_OK;
  return stage_gemm(c, la, lb, ld_, M, N, K, A, lda, B, ldb, D, ldd,
                    [&](const double* dA, const double* dB, double* dD) {
                      return skgemm_dgemm(c, la, lb, ld_, M, N, K, alpha, dA,
                                          lda, dB, ldb, beta, dD, ldd, gamma);
                    });
}

extern "C" int skgemm_dsyrk_host(skgemm_ctx* c, int la, int N, int K,
                                 double alpha, const double* A, int64_t lda,
                                 double beta, double* C, int64_t ldc) {
  SKG_TRY(check_ctx(c));
  SKG_TRY(host_args_syrk(la, N, K, A, lda, C, ldc));
  if (N == 0) return SKGEMM_OK;
  return stage_syrk(c, la, N, K, A, lda, C, ldc,
                    [&](const double* dA, double* dC) {
                      return skgemm_dsyrk(c, la, N, K, alpha, dA, lda, beta, dC,
                                          ldc);
                    });
}

extern "C" int skgemm_dgemm_ozaki_host(skgemm_ctx* c, int la, int lb, int ld_,
                                       int M, int N, int K, double alpha,
                                       const double* A, int64_t lda,
                                       const double* B, int64_t ldb,
                                       double beta, double* D, int64_t ldd,
                                       double gamma, int slices,
                                       double* err_bound) {
  SKG_TRY(check_ctx(c));
  SKG_TRY(host_args_gemm(la, lb, ld_, M, N, K, A, lda, B, ldb, D, ldd));
  if (err_bound) *err_bound = 0.0;
  if (M == 0 || N == 0) return SKGEMM_OK;
  return stage_gemm(c, la, lb, ld_, M, N, K, A, lda, B, ldb, D, ldd,
                    [&](const double* dA, const double* dB, double* dD) {
                      return skgemm_dgemm_ozaki(c, la, lb, ld_, M, N, K, alpha,
                                                dA, lda, dB, ldb, beta, dD, ldd,
                                                gamma, slices, err_bound);
                    });
}

extern "C" int skgemm_dsyrk_ozaki_host(skgemm_ctx* c, int la, int N, int K,
                                       double alpha, const double* A,
                                       int64_t lda, double beta, double* C,
                                       int64_t ldc, int slices,
                                       double* err_bound) {
  SKG_TRY(check_ctx(c));
  SKG_TRY(host_args_syrk(la, N, K, A, lda, C, ldc));
  if (err_bound) *err_bound = 0.0;
  if (N == 0) return SKGEMM_OK;
  return stage_syrk(c, la, N, K, A, lda, C, ldc,
                    [&](const double* dA, double* dC) {
                      return skgemm_dsyrk_ozaki(c, la, N, K, alpha, dA, lda,
                                                beta, dC, ldc, slices,
                                                err_bound);
                    });
}

// ---- layout conversion --------------------------------------------------------

// Lane order for one side of a layout conversion: along the index with unit
// stride, in groups of the panel width when that index is a packed one.
static void pack_side(bool o_unit, bool r_unit, int pshift, bool* fast_o,
                      int* wshift) {
  *fast_o = o_unit || !r_unit;
  *wshift = (*fast_o && pshift != kNoPanelShift) ? pshift : 6;
}

static int convert(skgemm_ctx* c, bool swap, int src_layout, const double* src,
                   int64_t ld_src, int dst_layout, double* dst, int64_t ld_dst,
                   int o_extent, int r_extent, const char* name) {
  SKG_TRY(check_ctx(c));
  SKG_TRY(check_dims(o_extent, r_extent, 1));
  if (o_extent == 0 || r_extent == 0) return SKGEMM_OK;
  if (src == nullptr || dst == nullptr)
    return set_error(SKGEMM_ERR_ARG, "%s: src or dst is NULL", name);
  // Both sides in the B convention: element (r, o), ROW_MAJOR = r*ld + o.
  Operand s{};
  SKG_TRY(make_operand(src_layout, false, src, ld_src, o_extent, r_extent, &s, "src"));
  Output d{};
  if (swap)  // dst element (r' = o, o' = r)
    SKG_TRY(make_output(dst_layout, dst, ld_dst, o_extent, r_extent, &d, "dst"));
  else
    SKG_TRY(make_output(dst_layout, dst, ld_dst, r_extent, o_extent, &d, "dst"));
  SKG_TRY(check_dev(src, "src"));
  SKG_TRY(check_dev(dst, "dst"));
  PackMap map{};
  map.swap = swap ? 1 : 0;
  bool fo;
  pack_side(s.os == 1, s.ks == 1, s.pshift, &fo, &map.src_wshift);
  map.src_fast_o = fo ? 1 : 0;
  // dst: n is the packed / unit-stride index of an Output; n = o, or r
  // when swapped.
  const bool n_unit = d.ns == 1, m_unit = d.ms == 1;
  bool n_fast;
  int w;
  pack_side(n_unit, m_unit, d.pshift, &n_fast, &w);
  map.dst_fast_o = (n_fast != swap) ? 1 : 0;
  map.dst_wshift = w;
  // 16-byte pairs along a side's lane order: a 16-byte aligned base, unit
  // stride along that index, the other index's stride and the panel stride
  // even, panels at least two wide (a pair starts at an even index).
  const bool s_al = ((uintptr_t)src & 15) == 0 && s.pstride % 2 == 0 &&
                    s.pshift >= 1;
  const bool vs = s_al && (map.src_fast_o ? (s.os == 1 && s.ks % 2 == 0)
                                          : (s.ks == 1 && s.os % 2 == 0));
  const bool d_al = ((uintptr_t)dst & 15) == 0 && d.pstride % 2 == 0 &&
                    d.pshift >= 1;
  const bool vd = d_al && (n_fast ? (d.ns == 1 && d.ms % 2 == 0)
                                  : (d.ms == 1 && d.ns % 2 == 0));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  c->record_begin();
  SKG_HIP(launch_pack(s, r_extent, d, map, vs, vd, c->stream), name);
  c->record_end();
  return SKGEMM_OK;
}

extern "C" int skgemm_pack(skgemm_ctx* c, int src_layout, const double* src,
                           int64_t ld_src, int dst_layout, double* dst,
                           int64_t ld_dst, int o_extent, int r_extent) {
  return convert(c, false, src_layout, src, ld_src, dst_layout, dst, ld_dst,
                 o_extent, r_extent, "pack launch");
}

extern "C" int skgemm_transpose(skgemm_ctx* c, int src_layout,
                                const double* src, int64_t ld_src,
                                int dst_layout, double* dst, int64_t ld_dst,
                                int o_extent, int r_extent) {
  return convert(c, true, src_layout, src, ld_src, dst_layout, dst, ld_dst,
                 o_extent, r_extent, "transpose launch");
}

// ---- tuning -------------------------------------------------------------------

extern "C" int skgemm_num_tiles(void) { return num_tile_cfgs(); }

extern "C" int skgemm_tile_info(int tile, int* bm, int* bn, int* bk, int* waves) {
  if (tile < 0 || tile >= num_tile_cfgs())
    return set_error(SKGEMM_ERR_ARG, "tile %d out of range", tile);
  const TileCfg& t = tile_cfg(tile);
  if (bm) *bm = t.bm;
  if (bn) *bn = t.bn;
  if (bk) *bk = t.bk;
  if (waves) *waves = t.wm * t.wn;
  return SKGEMM_OK;
}

extern "C" int skgemm_set_tuning(skgemm_ctx* c, int tile, int grid) {
  SKG_TRY(check_ctx(c));
  if (tile >= num_tile_cfgs())
    return set_error(SKGEMM_ERR_ARG, "tile %d out of range", tile);
  c->forced_cfg = tile < 0 ? -1 : tile;
  c->forced_grid = grid <= 0 ? 0 : grid;
  return SKGEMM_OK;
}

extern "C" int skgemm_query_plan(skgemm_ctx* c, int kind, int la, int lb,
                                 int M, int N, int K, int* tile, int* grid,
                                 int* kchunks, int64_t* tiles, int64_t* iters) {
  SKG_TRY(check_ctx(c));
  SKG_TRY(check_dims(M, N, K));
  if (kind != 0 && kind != 1) return set_error(SKGEMM_ERR_ARG, "kind must be 0 or 1");
  if (kind == 0 && la == SKGEMM_COL_MAJOR) {
    // The autotuner's convention: a column-major A comes with a
    // column-major D, which is launched transposed.
    int ld_ = SKGEMM_COL_MAJOR;
    const double *a = nullptr, *b = nullptr;
    int64_t x = 0, y = 0;
    transpose_launch(la, lb, ld_, M, N, a, x, b, y);
  }
  const Plan pl = plan(c, kind, la, lb, M, N, K);
  const TileCfg& t = tile_cfg(pl.cfg);
  const int64_t nt = tiles_for(kind, pl.cfg, M, N);
  const int64_t ipt = std::max<int64_t>(1, (K + t.bk - 1) / t.bk);
  const int64_t kci = (ipt + pl.kchunks - 1) / pl.kchunks;
  if (tile) *tile = pl.cfg;
  if (grid) *grid = pl.grid;
  if (kchunks) *kchunks = pl.kchunks;
  if (tiles) *tiles = nt;
  if (iters) *iters = nt * pl.kchunks * kci;
  return SKGEMM_OK;
}

extern "C" int skgemm_set_kchunks(skgemm_ctx* c, int kchunks) {
  SKG_TRY(check_ctx(c));
  if (kchunks < 0) return set_error(SKGEMM_ERR_ARG, "kchunks < 0");
  c->forced_kchunks = kchunks;
  return SKGEMM_OK;
}

extern "C" int skgemm_timing_stages(skgemm_ctx* c, int on) {
  SKG_TRY(check_ctx(c));
  c->stage_timing = on != 0;
  return SKGEMM_OK;
}

extern "C" int skgemm_split_info(skgemm_ctx* c, int* moduli) {
  SKG_TRY(check_ctx(c));
  if (moduli == nullptr) return set_error(SKGEMM_ERR_ARG, "moduli is NULL");
  *moduli = c->oz_last_moduli;
  return SKGEMM_OK;
}

extern "C" int skgemm_timing_enable(skgemm_ctx* c, int capacity) {
  SKG_TRY(check_ctx(c));
  if (capacity < 0) return set_error(SKGEMM_ERR_ARG, "capacity < 0");
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  SKG_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  for (auto& pr : c->tring) {
    hipEventDestroy(pr.first);
    hipEventDestroy(pr.second);
  }
  c->tring.clear();
  c->tring_used = 0;
  c->last0 = c->ev0;
  c->last1 = c->ev1;
  c->timed = false;
  // Timing-only events, no system-scope fence on record (as ev0 / ev1).
  for (int i = 0; i < capacity; ++i) {
    hipEvent_t a, b;
    SKG_HIP(hipEventCreateWithFlags(&a, hipEventDisableSystemFence),
            "hipEventCreate");
    hipError_t e = hipEventCreateWithFlags(&b, hipEventDisableSystemFence);
    if (e != hipSuccess) {
      hipEventDestroy(a);
      return hip_error(e, "hipEventCreate");
    }
    c->tring.emplace_back(a, b);
  }
  return SKGEMM_OK;
}

extern "C" int skgemm_timing_read(skgemm_ctx* c, double* ms, int max,
                                  int* count) {
  SKG_TRY(check_ctx(c));
  const int n = c->tring_used;
  if (n > 0)
    SKG_HIP(hipEventSynchronize(c->tring[n - 1].second), "hipEventSynchronize");
  int w = 0;
  for (int i = 0; i < n && i < max; ++i) {
    float f = 0.f;
    SKG_HIP(hipEventElapsedTime(&f, c->tring[i].first, c->tring[i].second),
            "hipEventElapsedTime");
    if (ms) ms[i] = f;
    ++w;
  }
  if (count) *count = w;
  c->tring_used = 0;
  return c->check_status();
}

extern "C" int skgemm_last_kernel_ms(skgemm_ctx* c, double* ms) {
  SKG_TRY(check_ctx(c));
  if (!c->timed) return set_error(SKGEMM_ERR_STATE, "no timed launch yet");
  SKG_HIP(hipEventSynchronize(c->last1), "hipEventSynchronize");
  float f = 0.f;
  SKG_HIP(hipEventElapsedTime(&f, c->last0, c->last1), "hipEventElapsedTime");
  *ms = f;
  return SKGEMM_OK;
}

// ---- probes and fills -----------------------------------------------------------

extern "C" int skgemm_peak_probe(skgemm_ctx* c, int which, double* tflops) {
  SKG_TRY(check_ctx(c));
  if (which != 0 && which != 1) return set_error(SKGEMM_ERR_ARG, "which must be 0 or 1");
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  double* sink = c->staging(3, 65536 * sizeof(double));
  if (!sink) return set_error(SKGEMM_ERR_NOMEM, "probe sink");
  const int blocks = c->num_cus * 8;
  const int iters = which == 0 ? 2000 : 20000;
  SKG_HIP(launch_peak_probe(which, blocks, iters / 10, sink, c->stream), "probe warmup");
  SKG_HIP(hipEventRecord(c->ev0, c->stream), "hipEventRecord");
  SKG_HIP(launch_peak_probe(which, blocks, iters, sink, c->stream), "probe");
  SKG_HIP(hipEventRecord(c->ev1, c->stream), "hipEventRecord");
  SKG_HIP(hipEventSynchronize(c->ev1), "hipEventSynchronize");
  float ms = 0.f;
  SKG_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1), "hipEventElapsedTime");
  const double per_mfma = which == 0 ? 2.0 * 16 * 16 * 4 : 2.0 * 32 * 32 * 32;
  const double n_mfma = (double)blocks * 4 /*waves*/ * iters * (which == 0 ? 8 : 4);
  *tflops = n_mfma * per_mfma / (ms * 1e-3) / 1e12;
  return SKGEMM_OK;
}

extern "C" int skgemm_copy_probe(skgemm_ctx* c, size_t bytes, double* gbps) {
  SKG_TRY(check_ctx(c));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  bytes &= ~(size_t)15;
  void *a = nullptr, *b = nullptr;
  SKG_HIP(hipMalloc(&a, bytes), "hipMalloc probe");
  hipError_t e = hipMalloc(&b, bytes);
  if (e != hipSuccess) {
    hipFree(a);
    return hip_error(e, "hipMalloc probe");
  }
  hipMemsetAsync(a, 0, bytes, c->stream);
  const int64_t n = bytes / 8;
  e = launch_copy_probe((double*)a, (double*)b, n, c->stream);
  hipEventRecord(c->ev0, c->stream);
  const int reps = 10;
  for (int i = 0; i < reps && e == hipSuccess; ++i)
    e = launch_copy_probe((double*)a, (double*)b, n, c->stream);
  hipEventRecord(c->ev1, c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
  else hipEventSynchronize(c->ev1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess) return hip_error(e, "copy probe");
  *gbps = 2.0 * bytes * reps / (ms * 1e-3) / 1e9;
  return SKGEMM_OK;
}

extern "C" int skgemm_fill_uniform(skgemm_ctx* c, double* x, int64_t n,
                                   uint64_t seed, double lo, double hi) {
  SKG_TRY(check_ctx(c));
  if (n < 0 || (n > 0 && x == nullptr)) return set_error(SKGEMM_ERR_ARG, "bad buffer");
  if (n == 0) return SKGEMM_OK;
  SKG_TRY(check_dev(x, "x"));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  SKG_HIP(launch_fill_uniform(x, n, seed, lo, hi, c->stream), "fill_uniform");
  return SKGEMM_OK;
}

extern "C" int skgemm_fill_normal(skgemm_ctx* c, double* x, int64_t n,
                                  uint64_t seed, double sigma) {
  SKG_TRY(check_ctx(c));
  if (n < 0 || (n > 0 && x == nullptr)) return set_error(SKGEMM_ERR_ARG, "bad buffer");
  if (n == 0) return SKGEMM_OK;
  SKG_TRY(check_dev(x, "x"));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  SKG_HIP(launch_fill_normal(x, n, seed, sigma, c->stream), "fill_normal");
  return SKGEMM_OK;
}

extern "C" int skgemm_scale_columns_pow10(skgemm_ctx* c, double* x, int rows,
                                          int cols, int64_t ld, uint64_t seed,
                                          double umax) {
  SKG_TRY(check_ctx(c));
  if (rows < 0 || cols < 0 || ld < rows || (rows > 0 && cols > 0 && !x))
    return set_error(SKGEMM_ERR_ARG, "bad matrix");
  if (rows == 0 || cols == 0) return SKGEMM_OK;
  SKG_TRY(check_dev(x, "x"));
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  SKG_HIP(launch_scale_columns_pow10(x, rows, cols, ld, seed, umax, c->stream),
          "scale_columns");
  return SKGEMM_OK;
}
