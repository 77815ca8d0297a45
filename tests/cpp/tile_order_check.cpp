// Prints the tile coordinates that tile_coords (stream-k-gemm_amd/csrc/
// skgemm_internal.h, the function both the Stream-K kernel and the host's
// weighted split call) gives for every flattened tile index of one problem:
//   tile_order_check syrk row_order tiles_m tiles_n
// one "ti tj" line per tile, in tile order.
#include <cstdio>
#include <cstdlib>

#include "../../stream-k-gemm_amd/csrc/skgemm_internal.h"

int main(int argc, char** argv) {
  if (argc != 5) return 2;
  const int syrk = std::atoi(argv[1]), order = std::atoi(argv[2]);
  const int tm = std::atoi(argv[3]), tn = std::atoi(argv[4]);
  const long long n = syrk ? (long long)tm * (tm + 1) / 2 : (long long)tm * tn;
  for (long long t = 0; t < n; ++t) {
    int ti, tj;
    skg::tile_coords(t, syrk, order, tm, tn, ti, tj);
    std::printf("%d %d\n", ti, tj);
  }
  return 0;
}
