// Entry points whose implementation lands in later milestones; each reports
// SKGEMM_ERR_NOTSUP so callers see a clear error rather than a missing symbol.
#include "../../include/skgemm.h"
#include "skgemm_ctx.h"

using namespace skg;

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
