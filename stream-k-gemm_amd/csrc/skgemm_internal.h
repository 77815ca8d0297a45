// Internal declarations shared by the kernel and host translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace skg {

// Free-index panel shift used for unpacked layouts: o >> 30 == 0 for every
// extent the C ABI accepts (int, < 2^30 checked on entry).
constexpr int kNoPanelShift = 30;

// One input operand, addressed by its free index o (m of A, n of B) and the
// contraction index k:
//   addr(o, k) = (o >> pshift) * pstride + (o & pmask) * os + k * ks
struct Operand {
  const double* ptr;
  int64_t pstride;
  int64_t os;
  int64_t ks;
  int pshift;
  int extent;   // M or N
  int kcontig;  // ks == 1: pairs are loaded along k, else along o
  int vec2;     // 16-byte loads of a pair are aligned
};

// Output D (m, n): addr = (n >> pshift) * pstride + (n & pmask) * ns + m * ms
struct Output {
  double* ptr;
  int64_t pstride;
  int64_t ns;
  int64_t ms;
  int pshift;
};

// One problem as the Stream-K kernel sees it.
struct Problem {
  Operand a, b;
  Output d;
  int M, N, K;
  int syrk;            // 1: tiles of the lower triangle only, diagonal masked;
                       // 2: the same, storing whole 16 x 16 diagonal blocks
  double alpha, beta, gamma;
  int tiles_m, tiles_n;
  int iters_per_tile;  // ceil(K / BK), at least 1
  // GEMM tile order: 0 = ti fastest, 1 = tj fastest (A panels read once),
  // >= 2 = bands of row_order tile columns, tj fastest within a band.
  int row_order;
  // K is cut into kchunks chunks of kci BK-steps; a unit is (chunk, tile).
  // With kchunks > 1 each unit writes its raw sum to part[c] (row-major
  // M x N, ld = N) and a reduce pass forms D.
  int kchunks;
  int kci;
  double* part;
  const int* pred;     // single launches: run only when *pred != 0
  // Columns N - xcol .. N - 1 (xcol <= 16) ride on the last tile column as
  // an extra strip, or rows M - xrow .. M - 1 on the last tile row
  // (configurations with xs; at most one of the two is non-zero).
  int xcol, xrow;
  int64_t tile_begin;  // first flattened tile index of this problem
  int64_t iter_begin;  // first flattened MAC-iteration index
  int64_t num_tiles;
};

// SYRK store predicate.  syrk == 1: the lower triangle (row >= col) of the
// row-major C.  syrk == 2: that plus the whole of every 16 x 16 diagonal
// block, which is what the reference's tile store writes
// (include/sme_syrk.hpp:57-109, syrk_crsp.cpp:137-244: full 16 x 16 blocks
// of the tile pairs tm1 <= tm2, the diagonal pairs included).
__device__ __forceinline__ bool syrk_keep(int syrk, int row, int col) {
  return syrk == 0 || row >= col || (syrk == 2 && (row >> 4) == (col >> 4));
}

// Flattened tile index -> tile coordinates (ti, tj) of one problem, shared
// by the kernel and the host's cost-weighted split so both walk the same
// order.  SYRK: the row-major lower triangle (ti >= tj), t = ti (ti + 1) / 2
// + tj.  GEMM: row_order 0 = ti fastest, 1 = tj fastest, >= 2 = bands of
// row_order tile columns (the last band may be narrower), tj fastest
// within a band, bands in order.
__host__ __device__ __forceinline__ void tile_coords(int64_t t, int syrk,
                                                     int row_order,
                                                     int tiles_m, int tiles_n,
                                                     int& ti, int& tj) {
  if (syrk) {
    int64_t i = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (i * (i + 1) / 2 > t) --i;
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    ti = (int)i;
    tj = (int)(t - i * (i + 1) / 2);
  } else if (row_order == 1) {
    ti = (int)(t / tiles_n);
    tj = (int)(t % tiles_n);
  } else if (row_order >= 2) {
    const int64_t per = (int64_t)tiles_m * row_order;
    const int bnd = (int)(t / per);
    const int r = (int)(t - bnd * per);
    const int left = tiles_n - bnd * row_order;
    const int bw = row_order < left ? row_order : left;
    ti = r / bw;
    tj = bnd * row_order + r % bw;
  } else {
    ti = (int)(t % tiles_m);
    tj = (int)(t / tiles_m);
  }
}

// Launch-wide state of one Stream-K launch.
struct Launch {
  const Problem* problems;  // device table (batched) or nullptr
  const int64_t* hdr;       // batched: device {num_problems, total_iters}
  Problem single;           // used when problems == nullptr
  int num_problems;         // single launches (batched: from hdr)
  int64_t total_iters;
  int grid;             // persistent workgroups
  // Single launches: first flattened iteration of each workgroup (grid + 1
  // entries, a cost-weighted split), or nullptr for the even split.
  const int64_t* starts;
  double* ws;           // grid (wait-free: 2 grid) partial tiles
  int* flags;           // grid flags: 0 empty, 1 partial ready (wait-free:
                        // arrival counters of split units, by first
                        // contributor)
  int* status;          // bit 0: a fix-up wait hit its spin limit
  int64_t spin_limit;   // s_sleep rounds before a wait gives up
  int waitfree;         // 1: last-arriver fix-up, no inter-workgroup waits
  int hold;             // 1: peers store partials but never raise their
                        // flags (failure-path tests; waiting fix-up only)
};

// Tile configurations (index = kernel table entry).
struct TileCfg {
  int bm, bn, bk, wm, wn;
  int ns;  // 0: register-staged LDS double buffer; >= 3: LDS-DMA stages
  int xs;  // 1: an N (else M) remainder of up to 16 rides on the last tile
           // column (row) (single problems, 16-byte aligned operands; a
           // SYRK takes the row strip)
};
int num_tile_cfgs();
const TileCfg& tile_cfg(int i);
// The register-staged configuration with the same tile shape (cfg itself
// when it has no DMA stages).
int register_sibling(int cfg);
// The configuration with the same tile and stages but no strip.
int strip_twin(int cfg);
// Strip columns (rows) of a problem under `cfg` (0: none; then N (M) is
// tiled by BN (BM)).  A column strip takes precedence.
int strip_cols(int cfg, int kind, int N);
int strip_rows(int cfg, int kind, int M, int N);
// Doubles of one workgroup's partial-tile slot in the fix-up workspace.
size_t slot_doubles(int cfg);

// Launch the FP64 Stream-K kernel with tile configuration `cfg`.
hipError_t launch_dgemm_sk(int cfg, const Launch& L, hipStream_t s);
// D = alpha * sum_c part[c] + beta * D + gamma for a problem whose units
// were written to part (kchunks > 1).
hipError_t launch_chunk_reduce(const Problem& p, hipStream_t s);
// Workgroups of configuration `cfg` that fit on one CU at once.
int occupancy_dgemm_sk(int cfg);
// LDS bytes one workgroup of `cfg` uses.
size_t lds_bytes_dgemm_sk(int cfg);

// Utility kernels.
hipError_t launch_fill_uniform(double* x, int64_t n, uint64_t seed, double lo,
                               double hi, hipStream_t s);
hipError_t launch_fill_normal(double* x, int64_t n, uint64_t seed,
                              double sigma, hipStream_t s);
hipError_t launch_scale_columns_pow10(double* x, int rows, int cols,
                                      int64_t ld, uint64_t seed, double umax,
                                      hipStream_t s);
// Lane order of the layout-conversion kernel on each side: lanes run along
// o (fast_o) or r in groups of 2^wshift (4 <= wshift <= 6).
struct PackMap {
  int swap;  // dst element (o, r) is out_addr(dst, o, r), else (r, o)
  int src_fast_o, src_wshift;
  int dst_fast_o, dst_wshift;
};
// vs / vd: 16-byte pairs along the source / destination lane order are
// contiguous and aligned (pack2_kernel), else one element per lane.
hipError_t launch_pack(const Operand& src, int r_extent, const Output& dst,
                       const PackMap& map, bool vs, bool vd, hipStream_t s);
hipError_t launch_zero_ints(int* p, int n, hipStream_t s);
hipError_t launch_peak_probe(int which, int blocks, int iters, double* sink,
                             hipStream_t s);
hipError_t launch_copy_probe(const double* src, double* dst, int64_t n,
                             hipStream_t s);

}  // namespace skg
