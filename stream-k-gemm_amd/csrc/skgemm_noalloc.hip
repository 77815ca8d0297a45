// Allocation-free FP64 SYRK (skgemm_dsyrk_noalloc): the Stream-K fix-up
// partials, their flags and the K-chunk partial sums live in one file-scope
// device array, so a call makes no hipMalloc, hipFree or other allocating
// call.  For callers that cannot allocate (code under HIP graph capture,
// tight serving loops, harnesses that forbid it).
//
// The array's name is a host-side stand-in in host code: kernels are given
// the device address that hipGetSymbolAddress returns for the current
// device, looked up once per device when a context is created there
// (noalloc_init) and kept per device.  Every call on a device shares that
// device's array and it keeps the last call's contents: the flags, which
// must start at zero, are cleared by a kernel at the start of each call.
// Calls that overlap on one device (two streams, two contexts) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/skgemm.h"
#include "skgemm_ctx.h"
#include "skgemm_internal.h"

namespace skg {
namespace {

// 128 MiB per device: a c3 overlap (65536 x 1024, 128 x 128 tiles on 252
// workgroups, 7 K chunks) needs 92 MB.
constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
bool g_done[kMaxDevices];
void* g_ptr[kMaxDevices];       // device address of `scratch`, per device
hipError_t g_err[kMaxDevices];

size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(SKGEMM_ERR_NOTSUP, "device %d beyond %d", device,
                     kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_done[device]) {
    // `device` is current (the caller set it).
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_err[device] = e;
    g_ptr[device] = e == hipSuccess ? p : nullptr;
    g_done[device] = true;
  }
  if (g_ptr[device] == nullptr)
    return hip_error(g_err[device], "address of the no-alloc scratch array");
  return SKGEMM_OK;
}

}  // namespace skg

using namespace skg;

extern "C" int skgemm_dsyrk_noalloc(skgemm_ctx* c, int la, int N, int K,
                                    double alpha, const double* A,
                                    int64_t lda, double beta, double* C,
                                    int64_t ldc) {
  SKG_TRY(check_ctx(c));
  Problem p{};
  SKG_TRY(prepare_syrk(la, N, K, A, lda, C, ldc, &p));
  if (N == 0) return SKGEMM_OK;
  SKG_HIP(hipSetDevice(c->device), "hipSetDevice");
  SKG_TRY(noalloc_init(c->device));  // a lookup: done when c was created
  char* base = (char*)g_ptr[c->device];
  const Plan pl = plan(c, 1, la, la, N, N, K);
  SKG_TRY(fill_problem(c, &p, 1, pl, N, N, K, /*alloc_part=*/false));
  p.alpha = alpha;
  p.beta = beta;
  p.gamma = 0.0;
  // Scratch layout: flags | fix-up partial tiles | K-chunk sums.
  const size_t flags_b = round256(sizeof(int) * (size_t)pl.grid);
  const size_t ws_n = c->ws_doubles(pl.cfg, pl.grid);
  const size_t ws_b = round256(sizeof(double) * ws_n);
  const size_t part_b =
      pl.kchunks > 1 ? sizeof(double) * (size_t)pl.kchunks * N * N : 0;
  const size_t need = flags_b + ws_b + part_b;
  if (need > kScratchBytes)
    return set_error(SKGEMM_ERR_NOMEM,
                     "skgemm_dsyrk_noalloc: N=%d K=%d under its plan (tile %d, "
                     "%d workgroups, %d K chunks) needs %zu bytes of scratch, "
                     "beyond the %zu-byte limit (one K chunk, "
                     "skgemm_set_kchunks(ctx, 1), needs %zu)",
                     N, K, pl.cfg, pl.grid, pl.kchunks, need, kScratchBytes,
                     flags_b + ws_b);
  CallerBuffers cb;
  cb.flags = (int*)base;
  cb.nflags = pl.grid;
  cb.ws = (double*)(base + flags_b);
  cb.ws_doubles = ws_n;
  p.part = pl.kchunks > 1 ? (double*)(base + flags_b + ws_b) : nullptr;
  // A kernel, not a memset node, so the call stays valid inside a capture.
  SKG_HIP(launch_zero_ints(cb.flags, pl.grid, c->stream), "clear flags");
  return run_single(c, p, pl, nullptr, true, &cb);
}
