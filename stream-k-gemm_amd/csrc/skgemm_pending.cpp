// Entry points whose implementation lands in later milestones; each reports
// SKGEMM_ERR_NOTSUP so callers see a clear error rather than a missing symbol.
#include "../../include/skgemm.h"
#include "skgemm_ctx.h"

using namespace skg;

struct skgemm_ctx::Extra {};
void skgemm_ctx::free_extra() {}

extern "C" int skgemm_dgemm_ozaki(skgemm_ctx*, int, int, int, int, int, int,
                                  double, const double*, int64_t,
                                  const double*, int64_t, double, double*,
                                  int64_t, double, int, double*) {
  return set_error(SKGEMM_ERR_NOTSUP, "skgemm_dgemm_ozaki: not built yet");
}
extern "C" int skgemm_dsyrk_ozaki(skgemm_ctx*, int, int, int, double,
                                  const double*, int64_t, double, double*,
                                  int64_t, int, double*) {
  return set_error(SKGEMM_ERR_NOTSUP, "skgemm_dsyrk_ozaki: not built yet");
}
extern "C" int skgemm_batched(skgemm_ctx*, const skgemm_problem*, int, int,
                              int, int) {
  return set_error(SKGEMM_ERR_NOTSUP, "skgemm_batched: not built yet");
}
extern "C" int skgemm_server_create(int, size_t, int, int, int, int,
                                    skgemm_server**) {
  return set_error(SKGEMM_ERR_NOTSUP, "server: not built yet");
}
extern "C" int skgemm_server_destroy(skgemm_server*) {
  return set_error(SKGEMM_ERR_NOTSUP, "server: not built yet");
}
extern "C" int skgemm_server_add_system(skgemm_server*, const double*, int,
                                        int, int*) {
  return set_error(SKGEMM_ERR_NOTSUP, "server: not built yet");
}
extern "C" int skgemm_server_submit(skgemm_server*, int, int, const double*,
                                    double*, int64_t*) {
  return set_error(SKGEMM_ERR_NOTSUP, "server: not built yet");
}
extern "C" int skgemm_server_wait(skgemm_server*, int64_t, double*) {
  return set_error(SKGEMM_ERR_NOTSUP, "server: not built yet");
}
extern "C" int skgemm_server_stats(skgemm_server*, int64_t*, int64_t*,
                                   double*, double*, int*) {
  return set_error(SKGEMM_ERR_NOTSUP, "server: not built yet");
}
