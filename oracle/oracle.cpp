// Oracle: CPU restatement of the reference's FP64 GEMM / SYRK / rotation path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the library links or calls this file;
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// liboracle.so, and only as the checker (or, for cpu_baseline, as the timed
// CPU port of the reference path).  It is never the thing measured on the GPU.
//
// The reference (shriramch/stream-k-gemm) is C++ on ARM SME intrinsics and
// cannot be compiled on x86_64 (no arm_sme.h, no AArch64 sysroot, no SME
// hardware or emulator).  This file restates its decompositions with the
// 16x32 ZA outer-product microkernel replaced by a portable register-blocked
// FMA loop.  Each function cites the reference lines it follows.
//
// Pinning (what checks each function; tests/test_oracle.py):
//   oracle_gemm_crsp_packed / oracle_gemm_opt_packed / oracle_gemm_naive
//       closed form of the reference's own test inputs (gemm_crsp.cpp:451-460)
//       for all four reference KATs, brute force vs numpy float64 matmul, and
//       the double-double oracle.
//   oracle_syrk_tiles / oracle_syrk_reference
//       numpy A.T@A lower triangle on the reference's glibc-rand inputs
//       (syrk_crsp.cpp:379-395, seeds 42..46), orthonormal-column identity,
//       power-of-two column scaling identity.
//   oracle_trp_3d / oracle_trp_reference
//       closed form of trp_3d_test.cpp:125-138 inputs, Q = I, Q = permutation.
//   oracle_dgemm_dd / oracle_dsyrk_dd (double-double accumulation)
//       exact small-integer products, and agreement with numpy to 1e-15.
//   pack / unpack, band_to_grid / grid_to_band
//       round trips, element-wise layout checks (pack.cpp:37-258), the
//       reference's band<->grid known-answer pattern g + 1000 b
//       (pack.cpp:190-258), and numpy reshape/transpose restatements of
//       both packed layouts.
//   oracle_nloc_apply
//       the reference test's unpacked formulation (nloc_packed_opt2_test.cpp:
//       84-165) restated in numpy, closed forms (one projector, one column,
//       identity chi), the gamma-once-per-atom rule, overlapping indices.
//   oracle_halo_extend / oracle_stencil_apply
//       the reference test's naive unpacked stencil (stencil_extended_test.cpp:
//       18-52) restated with numpy slices, polynomial closed forms, the
//       reference's periodic-halo spot checks (halo_packed_test.cpp:
//       151-200), and numpy np.pad(mode="wrap") of the local grid.
// The split form (Ozaki scheme II) is restated separately in oracle/ozaki.py
// and pinned in tests/test_ozaki_oracle.py.

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

extern "C" {

// ---------------------------------------------------------------------------
// Deterministic input generators (SplitMix64).  Not in the reference; these
// give CPU and GPU code the same seeded inputs.  Counter-based: element i of
// stream `seed` is mix(seed * 0x9E3779B97F4A7C15 + (i+1) * golden) so a GPU
// thread can produce element i independently.
// ---------------------------------------------------------------------------
static inline uint64_t sm64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline uint64_t sm64_at(uint64_t seed, uint64_t i) {
  return sm64_mix(seed * 0x9E3779B97F4A7C15ULL + (i + 1) * 0x9E3779B97F4A7C15ULL);
}
static inline double u01(uint64_t z) {  // [0,1) with 53 random bits
  return (double)(z >> 11) * 0x1.0p-53;
}

void oracle_fill_uniform(double* x, long n, uint64_t seed, double lo,
                         double hi) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) x[i] = lo + (hi - lo) * u01(sm64_at(seed, i));
}

// ---------------------------------------------------------------------------
// Reference test-input generators.
// ---------------------------------------------------------------------------

// gemm_crsp.cpp:451-460 (same in gemm_opt.cpp): A column-major M x K,
// B row-major K x N, closed-form entries.
void oracle_gemm_test_inputs(int M, int N, int K, double* A_colmajor,
                             double* B_rowmajor) {
  for (int j = 0; j < K; ++j)
    for (int i = 0; i < M; ++i)
      A_colmajor[i + (long)j * M] = (i + 1) * 0.01 + (j + 1) * 0.001;
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < N; ++j)
      B_rowmajor[(long)i * N + j] = (i + 1) * 0.001 + (j + 1) * 0.01;
}

// syrk_crsp.cpp:379-395: glibc srand/rand()%100/100 into the 16-packed layout,
// zero padding for m >= M.  glibc's rand() is deterministic, so this
// regenerates the reference's exact SYRK test inputs.
void oracle_init_A_packed(int K, int M, double* A_packed, int seed) {
  srand(seed);
  int tiles_m = (M + 15) / 16;
  for (int tm = 0; tm < tiles_m; ++tm)
    for (int k = 0; k < K; ++k)
      for (int ml = 0; ml < 16; ++ml) {
        int m = tm * 16 + ml;
        A_packed[((long)tm * K + k) * 16 + ml] =
            (m < M) ? (double)(rand() % 100) / 100.0 : 0.0;
      }
}

// trp_3d_test.cpp:125-138: grid-packed A [Nd/16][nb][16], row-major Q.
void oracle_trp_test_inputs(int Nd, int nb, double* A_grid, double* B) {
  for (int g = 0; g < Nd; ++g)
    for (int b = 0; b < nb; ++b)
      A_grid[(long)(g / 16) * nb * 16 + b * 16 + g % 16] =
          (g + 1) * 0.001 + (b + 1) * 0.01;
  for (int b = 0; b < nb; ++b)
    for (int bp = 0; bp < nb; ++bp)
      B[(long)b * nb + bp] = (b + 1) * 0.01 + (bp + 1) * 0.001;
}

// ---------------------------------------------------------------------------
// Packing (gemm_crsp.cpp:342-366; pack.cpp / sme_pack.hpp layouts)
// ---------------------------------------------------------------------------

// gemm_crsp.cpp:342-353: column-major A -> [M/16][K][16]; truncates M/16.
void oracle_pack_A(int M, int K, const double* A_colmajor, double* A_packed) {
  const int tiles_m = M / 16;
  for (int tm = 0; tm < tiles_m; ++tm)
    for (int k = 0; k < K; ++k)
      for (int i = 0; i < 16; ++i)
        A_packed[((long)tm * K + k) * 16 + i] =
            A_colmajor[(tm * 16 + i) + (long)k * M];
}

// gemm_crsp.cpp:355-366: row-major B -> [N/32][K][32]; truncates N/32.
void oracle_pack_B(int N, int K, const double* B_rowmajor, double* B_packed) {
  const int tiles_n = N / 32;
  for (int tn = 0; tn < tiles_n; ++tn)
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < 32; ++j)
        B_packed[((long)tn * K + k) * 32 + j] =
            B_rowmajor[(long)k * N + (tn * 32 + j)];
}

// sme_pack.hpp:41-60 / pack.cpp:1-13: column-major [Nd, nb] ->
// grid-packed [Nd/16][nb][16] (dst[(g/16)*nb*16 + b*16 + g%16] = src[g+b*Nd]).
void oracle_pack_c16(int Nd, int nb, const double* src, double* dst) {
  for (int g = 0; g < (Nd / 16) * 16; ++g)
    for (int b = 0; b < nb; ++b)
      dst[(long)(g / 16) * nb * 16 + (long)b * 16 + g % 16] =
          src[g + (long)b * Nd];
}
void oracle_unpack_c16(int Nd, int nb, const double* src, double* dst) {
  for (int g = 0; g < (Nd / 16) * 16; ++g)
    for (int b = 0; b < nb; ++b)
      dst[g + (long)b * Nd] =
          src[(long)(g / 16) * nb * 16 + (long)b * 16 + g % 16];
}

// sme_pack.hpp:178-244 (band_to_grid_tile): band-packed [nb/16][Nd][16]
// (src[(b/16)*Nd*16 + g*16 + b%16]) -> grid-packed [Nd/16][nb][16]
// (dst[(g/16)*nb*16 + b*16 + g%16]); whole 16 x 16 blocks only
// (g_tiles = Nd/16, b_tiles = nb/16), as the reference's loops.
void oracle_band_to_grid(int Nd, int nb, const double* src, double* dst) {
  for (int g = 0; g < (Nd / 16) * 16; ++g)
    for (int b = 0; b < (nb / 16) * 16; ++b)
      dst[(long)(g / 16) * nb * 16 + (long)b * 16 + g % 16] =
          src[(long)(b / 16) * Nd * 16 + (long)g * 16 + b % 16];
}
// sme_pack.hpp:246-313 (grid_to_band_tile): the inverse map.
void oracle_grid_to_band(int Nd, int nb, const double* src, double* dst) {
  for (int g = 0; g < (Nd / 16) * 16; ++g)
    for (int b = 0; b < (nb / 16) * 16; ++b)
      dst[(long)(b / 16) * Nd * 16 + (long)g * 16 + b % 16] =
          src[(long)(g / 16) * nb * 16 + (long)b * 16 + g % 16];
}

// ---------------------------------------------------------------------------
// Portable register-blocked microkernel standing in for the 16x32 FMOPA
// microkernel (gemm_crsp.cpp:73-111): C[16][32] (+)= sum_k a[k][0:16] x
// b[k][0:32] with a = PA panel (stride 16 per k) and b = PB panel (32 per k).
// The accumulation per element is one FMA chain in k order, as with FMOPA.
// ---------------------------------------------------------------------------
static inline void ukernel_16x32(int k_len, const double* __restrict a,
                                 const double* __restrict b, double* __restrict c,
                                 long ldc, bool accumulate) {
  for (int i0 = 0; i0 < 16; i0 += 4) {
    for (int j0 = 0; j0 < 32; j0 += 8) {
      double acc[4][8];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j)
          acc[i][j] = accumulate ? c[(i0 + i) * ldc + j0 + j] : 0.0;
      for (int k = 0; k < k_len; ++k) {
        const double* ak = a + (long)k * 16 + i0;
        const double* bk = b + (long)k * 32 + j0;
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 8; ++j) acc[i][j] = std::fma(ak[i], bk[j], acc[i][j]);
      }
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) c[(i0 + i) * ldc + j0 + j] = acc[i][j];
    }
  }
}

// ---------------------------------------------------------------------------
// gemm_crsp: split-K over threads ("Stream-K" in the reference).
// gemm_crsp.cpp:189-243 (per-thread tile loop, KC/NC/MC blocking, partial C
// reload/store per KC block) and :261-311 (k_chunk = ceil(K/T), per-thread
// full M x N partial, thread-ordered reduction, D = alpha*sum + gamma).
// Works on the packed layouts PA [M/16][K][16], PB [N/32][K][32]; like the
// reference the product covers only 16*floor(M/16) rows and 32*floor(N/32)
// columns, and the reduction writes alpha * 0 + gamma to the rest of D.
// ---------------------------------------------------------------------------
void oracle_gemm_crsp_packed(int M, int N, int K, const double* A_packed,
                             const double* B_packed, double* D, double alpha,
                             double gamma, int nthreads) {
  const int KC = 256, MC = 640, NC = 128;  // gemm_crsp.cpp:39-49
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  const int tiles_m = M / 16, tiles_n = N / 32;
  const int k_chunk = (K + nthreads - 1) / nthreads;
  std::vector<std::vector<double>> C_local(nthreads,
                                           std::vector<double>((size_t)M * N, 0.0));
#pragma omp parallel num_threads(nthreads)
  {
    const int tid = omp_get_thread_num();
    const int k_start = tid * k_chunk;
    const int k_end = std::min(k_start + k_chunk, K);
    double* Cl = C_local[tid].data();
    if (k_start < K) {
      const int k_len = k_end - k_start;
      for (int kc = 0; kc < k_len; kc += KC) {
        const int kb_len = std::min(KC, k_len - kc);
        const bool is_first = (kc == 0);
        for (int nc = 0; nc < tiles_n; nc += NC / 32) {
          const int tn_end = std::min(nc + NC / 32, tiles_n);
          for (int mc = 0; mc < tiles_m; mc += MC / 16) {
            const int tm_end = std::min(mc + MC / 16, tiles_m);
            for (int tm = mc; tm < tm_end; ++tm)
              for (int tn = nc; tn < tn_end; ++tn) {
                const double* a = A_packed + ((long)tm * K + k_start + kc) * 16;
                const double* b = B_packed + ((long)tn * K + k_start + kc) * 32;
                double* c = Cl + (long)tm * 16 * N + tn * 32;
                ukernel_16x32(kb_len, a, b, c, N, !is_first);
              }
          }
        }
      }
    }
  }
  // gemm_crsp.cpp:285-310: for every i < M, j < N sum the partials in
  // thread order 0..T-1, then D = alpha*sum + gamma (fused multiply-add, as
  // svmla).  Outside the 16*floor(M/16) x 32*floor(N/32) panels no thread
  // wrote a partial, so the sum is zero and D = alpha * 0 + gamma there.
#pragma omp parallel for num_threads(nthreads)
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      double sum = 0.0;
      for (int t = 0; t < nthreads; ++t) sum += C_local[t][(long)i * N + j];
      D[(long)i * N + j] = std::fma(sum, alpha, gamma);
    }
}

// gemm_opt.cpp:240-337: contiguous M-tiles per thread, K-blocking outermost,
// D written in place, alpha/gamma applied on the last KC block (:283-287).
void oracle_gemm_opt_packed(int M, int N, int K, const double* A_packed,
                            const double* B_packed, double* D, double alpha,
                            double gamma, int nthreads) {
  const int KC = 256, MC = 64, NC = 128;  // gemm_opt.cpp:37-47
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  const int tiles_m = M / 16, tiles_n = N / 32;
  const int per = (tiles_m + nthreads - 1) / nthreads;
#pragma omp parallel num_threads(nthreads)
  {
    const int tid = omp_get_thread_num();
    const int tm_start = tid * per;
    const int tm_end = std::min(tm_start + per, tiles_m);
    for (int kc = 0; tm_start < tm_end && kc < K; kc += KC) {
      const int k_len = std::min(KC, K - kc);
      const bool is_first = (kc == 0), is_last = (kc + KC >= K);
      for (int nc = 0; nc < tiles_n; nc += NC / 32) {
        const int nc_end = std::min(nc + NC / 32, tiles_n);
        for (int mc = tm_start; mc < tm_end; mc += MC / 16) {
          const int mc_end = std::min(mc + MC / 16, tm_end);
          for (int tn = nc; tn < nc_end; ++tn)
            for (int tm = mc; tm < mc_end; ++tm) {
              const double* a = A_packed + ((long)tm * K + kc) * 16;
              const double* b = B_packed + ((long)tn * K + kc) * 32;
              double* c = D + (long)tm * 16 * N + tn * 32;
              ukernel_16x32(k_len, a, b, c, N, !is_first);
              if (is_last)
                for (int i = 0; i < 16; ++i)
                  for (int j = 0; j < 32; ++j)
                    c[(long)i * N + j] = c[(long)i * N + j] * alpha + gamma;
            }
        }
      }
    }
  }
}

// gemm_crsp.cpp:472-483: naive reference, A column-major, B row-major,
// D row-major, D = alpha*sum + gamma, sum accumulated in k order.
void oracle_gemm_naive(int M, int N, int K, const double* A_colmajor,
                       const double* B_rowmajor, double* D, double alpha,
                       double gamma) {
#pragma omp parallel for
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      double sum = 0;
      for (int k = 0; k < K; ++k)
        sum += A_colmajor[i + (long)k * M] * B_rowmajor[(long)k * N + j];
      D[(long)i * N + j] = alpha * sum + gamma;
    }
}

// ---------------------------------------------------------------------------
// SYRK on the 16-packed layout P16(M,K): A[k,m] = A_packed[(m/16)*K*16 +
// k*16 + m%16]; C is M x M column-major, entries m1 <= m2 defined
// (= row-major lower triangle).
// ---------------------------------------------------------------------------

// syrk_crsp.cpp:355-373: naive triangle, C[m1 + m2*M] = alpha*sum + beta*C.
void oracle_syrk_reference(int K, int M, const double* A_packed, double* C,
                           double alpha, double beta) {
#pragma omp parallel for schedule(dynamic)
  for (int m1 = 0; m1 < M; ++m1)
    for (int m2 = m1; m2 < M; ++m2) {
      double sum = 0;
      for (int k = 0; k < K; ++k)
        sum += A_packed[(long)(m1 / 16) * K * 16 + (long)k * 16 + m1 % 16] *
               A_packed[(long)(m2 / 16) * K * 16 + (long)k * 16 + m2 % 16];
      C[m1 + (long)m2 * M] = alpha * sum + beta * C[m1 + (long)m2 * M];
    }
}

// sme_syrk.hpp:111-149 / syrk_crsp.cpp:246-348: dynamic schedule over tile
// pairs tm1 <= tm2, 16x16 blocks accumulated per KC=256 block and stored with
// a running beta.  Every pair's block is stored whole (sme_syrk.hpp:57-109,
// syrk_crsp.cpp:137-244), so the diagonal pairs tm1 == tm2 also define the
// entries m1 > m2 of their 16 x 16 block.  Two deliberate departures, both
// documented:
//  * Q3: the reference applies alpha only to the last KC block
//    (sme_syrk.hpp:123-133); this restatement computes the intended
//    alpha*A^T*A + beta*C (identical when alpha == 1, the only value the
//    reference tests, syrk_crsp.cpp:500).
//  * Q2: when M % 16 != 0 the reference stores whole 16x16 blocks past the
//    matrix (sme_syrk.hpp:66-108); here stores are clipped to m < M.
void oracle_syrk_tiles(int K, int M, const double* A_packed, double* C,
                       double alpha, double beta, int nthreads) {
  const int KC = 256;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  const int tiles_m = (M + 15) / 16;
#pragma omp parallel for schedule(dynamic) collapse(2) num_threads(nthreads)
  for (int tm1 = 0; tm1 < tiles_m; ++tm1)
    for (int tm2 = 0; tm2 < tiles_m; ++tm2) {
      if (tm1 > tm2) continue;
      const double* a1 = A_packed + (long)tm1 * K * 16;
      const double* a2 = A_packed + (long)tm2 * K * 16;
      double acc[16][16];
      for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j) acc[i][j] = 0.0;
      // One FMA chain per element across all KC blocks; summing per block and
      // then adding would change only the rounding order.
      for (int kc = 0; kc < K; kc += KC) {
        const int kl = std::min(KC, K - kc);
        for (int k = kc; k < kc + kl; ++k) {
          const double* x = a1 + (long)k * 16;
          const double* y = a2 + (long)k * 16;
          for (int i = 0; i < 16; ++i)
            for (int j = 0; j < 16; ++j) acc[i][j] = std::fma(x[i], y[j], acc[i][j]);
        }
      }
      for (int j = 0; j < 16; ++j) {
        const int col = tm2 * 16 + j;
        if (col >= M) continue;
        for (int i = 0; i < 16; ++i) {
          const int row = tm1 * 16 + i;
          if (row >= M) continue;  // whole 16 x 16 block, clipped to M
          double* c = C + row + (long)col * M;
          *c = (beta == 0.0) ? alpha * acc[i][j] : alpha * acc[i][j] + beta * *c;
        }
      }
    }
}

// ---------------------------------------------------------------------------
// Rotation Psi_new = Psi * Q (sme_trp_3d.hpp:1-16, :194-215).
// A grid-packed [Nd/16][nb][16], Q row-major [nb][nb],
// D band-packed [nb/16][Nd][16]; no alpha/gamma.
// ---------------------------------------------------------------------------

// trp_3d_test.cpp:56-70: naive reference.
void oracle_trp_reference(int Nd, int nb, const double* A_grid, const double* B,
                          double* D_band) {
#pragma omp parallel for
  for (int g = 0; g < Nd; ++g)
    for (int bp = 0; bp < nb; ++bp) {
      double sum = 0;
      for (int b = 0; b < nb; ++b)
        sum += A_grid[(long)(g / 16) * nb * 16 + (long)b * 16 + g % 16] *
               B[(long)b * nb + bp];
      D_band[(long)(bp / 16) * Nd * 16 + (long)g * 16 + bp % 16] = sum;
    }
}

// sme_trp_3d.hpp:194-215: parallel over (NC, MC) blocks; each 16-grid x
// 32-band microtile accumulates over all input bands (KC-blocked in ZA,
// one FMA chain per element) and is stored band-packed.  The reference's
// 32-wide band loop needs nb % 32 == 0; here a ragged last band block is
// clipped instead.
void oracle_trp_3d(int Nd, int nb, const double* A_grid, const double* B,
                   double* D_band, int nthreads) {
  const int TRP_MC = 64, TRP_NC = 128;  // sme_trp_3d.hpp:36-46
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  const int g_tiles = Nd / 16;
  const int mc_blocks = (g_tiles + TRP_MC - 1) / TRP_MC;
  const int nc_blocks = (nb + TRP_NC - 1) / TRP_NC;
#pragma omp parallel for collapse(2) schedule(dynamic) num_threads(nthreads)
  for (int ncb = 0; ncb < nc_blocks; ++ncb)
    for (int mcb = 0; mcb < mc_blocks; ++mcb) {
      const int nc0 = ncb * TRP_NC, nc1 = std::min(nc0 + TRP_NC, nb);
      const int mc0 = mcb * TRP_MC, mc1 = std::min(mc0 + TRP_MC, g_tiles);
      for (int bp0 = nc0; bp0 < nc1; bp0 += 32) {
        const int bpw = std::min(32, nc1 - bp0);
        for (int gt = mc0; gt < mc1; ++gt) {
          double acc[16][32];
          for (int i = 0; i < 16; ++i)
            for (int j = 0; j < 32; ++j) acc[i][j] = 0.0;
          const double* a = A_grid + (long)gt * nb * 16;
          for (int b = 0; b < nb; ++b) {
            const double* ab = a + (long)b * 16;
            const double* qb = B + (long)b * nb + bp0;
            for (int i = 0; i < 16; ++i)
              for (int j = 0; j < bpw; ++j) acc[i][j] = std::fma(ab[i], qb[j], acc[i][j]);
          }
          for (int i = 0; i < 16; ++i) {
            const int g = gt * 16 + i;
            for (int j = 0; j < bpw; ++j) {
              const int bp = bp0 + j;
              D_band[(long)(bp / 16) * Nd * 16 + (long)g * 16 + bp % 16] = acc[i][j];
            }
          }
        }
      }
    }
}

// ---------------------------------------------------------------------------
// General-layout checkers for the library's superset interface.  An operand
// is addressed by (outer index o, other index r):
//   layout 0 ROW_MAJOR : for A (o=m, r=k): m*ld + k   B (r=k, o=n): k*ld + n
//   layout 1 COL_MAJOR : A: m + k*ld                  B: k + n*ld
//   layout 2 PACK16    : (o/16)*ld*16 + r*16 + o%16   (ld = extent of r)
//   layout 3 PACK32    : (o/32)*ld*32 + r*32 + o%32
// D (m, n) follows the B convention with r = m, o = n:
//   ROW_MAJOR m*ld + n, COL_MAJOR m + n*ld, PACKP (n/P)*ld*P + m*P + n%P.
// ---------------------------------------------------------------------------
static inline long addr_a(int layout, long m, long k, long ld) {
  switch (layout) {
    case 0: return m * ld + k;
    case 1: return m + k * ld;
    case 2: return (m / 16) * ld * 16 + k * 16 + m % 16;
    default: return (m / 32) * ld * 32 + k * 32 + m % 32;
  }
}
static inline long addr_b(int layout, long k, long n, long ld) {
  switch (layout) {
    case 0: return k * ld + n;
    case 1: return k + n * ld;
    case 2: return (n / 16) * ld * 16 + k * 16 + n % 16;
    default: return (n / 32) * ld * 32 + k * 32 + n % 32;
  }
}

// Error-free transformations for double-double accumulation (Dekker/Knuth
// TwoSum and an FMA-based TwoProduct; Ogita, Rump & Oishi 2005, Dot2).
static inline void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
static inline void dd_fma_acc(double x, double y, double& hi, double& lo) {
  double p = x * y;
  double pe = std::fma(x, y, -p);
  double s, e;
  two_sum(hi, p, s, e);
  hi = s;
  lo += e + pe;
}

// D = alpha*op(A)*op(B) + beta*D + gamma, accumulated in double-double
// (dd != 0) or in plain FP64 FMA chains (dd == 0); the final combination is
// rounded once from double-double when dd != 0.
void oracle_dgemm(int la, int lb, int ld_layout, int M, int N, int K,
                  double alpha, const double* A, long lda, const double* B,
                  long ldb, double beta, double* D, long ldd, double gamma,
                  int dd) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      double hi = 0, lo = 0;
      for (int k = 0; k < K; ++k) {
        double x = A[addr_a(la, i, k, lda)], y = B[addr_b(lb, k, j, ldb)];
        if (dd) dd_fma_acc(x, y, hi, lo);
        else hi = std::fma(x, y, hi);
      }
      long di = addr_b(ld_layout, i, j, ldd);
      double s = hi + lo;
      double out = alpha * s;
      if (beta != 0.0) out += beta * D[di];
      D[di] = out + gamma;
    }
}

// C = alpha*A^T*A + beta*C on the lower triangle (row-major, i >= j) of the
// N x N output, A is K x N addressed with the B convention above.  Entries
// above the diagonal are left untouched.
void oracle_dsyrk(int la, int N, int K, double alpha, const double* A, long lda,
                  double beta, double* C, long ldc, int dd) {
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < N; ++i)
    for (int j = 0; j <= i; ++j) {
      double hi = 0, lo = 0;
      for (int k = 0; k < K; ++k) {
        double x = A[addr_b(la, k, i, lda)], y = A[addr_b(la, k, j, lda)];
        if (dd) dd_fma_acc(x, y, hi, lo);
        else hi = std::fma(x, y, hi);
      }
      double out = alpha * (hi + lo);
      if (beta != 0.0) out += beta * C[(long)i * ldc + j];
      C[(long)i * ldc + j] = out;
    }
}

// ---------------------------------------------------------------------------
// Non-local projector application (include/nloc_packed_opt2.hpp:229-447,
// 16-band-packed psi -> 32-band-packed result; :453-655, -> 16-band-packed),
// the sequential branch, restated step by step.  Projector p has nrow[p]
// grid indices index_all[roff[p] ..], chi_p column-major nrow x ncol at
// chi_all[choff[p] ..] and gamma_p at gamma_all[coff[p] ..].
//   1. gather (:55-79): filtered[r + b nrow] = psi16[(b/16 K + index[r]) 16
//      + b%16];
//   2. forward GEMM (:145-178 with beta = 1 on a zeroed chi_vector):
//      X[(coff + c) + b total_ncol] = 1 * X + dv * sum_r chi[r + c nrow]
//      filtered[r + b nrow], r ascending;
//   3. scale (:329-353): X rows of projector p times gamma_p, only when p's
//      atom_index differs from the previous projector's (starting from the
//      first projector's atom_index - 1);
//   4. backward GEMM (:185-224, beta = 0): temp[r + b nrow] = sum_c
//      chi[r + c nrow] * X[(coff + c) + b total_ncol], c ascending, each term
//      added to the running sum;
//   5. scatter (:88-137): result[(b/T K + index[r]) T + b%T] += temp[r + b
//      nrow], projectors in order, T = 32 or 16.
// Projectors with nrow = 0 or ncol = 0 are skipped (:259, :398).
void oracle_nloc_apply(int K, int nb, double dv, int num_proj,
                       const int* nrow, const int* ncol, const int* atom_index,
                       const int* index_all, const double* chi_all,
                       const double* gamma_all, const double* psi16,
                       double* result, int out_tile) {
  if (num_proj <= 0) return;
  std::vector<long> roff(num_proj + 1, 0), coff(num_proj + 1, 0),
      choff(num_proj + 1, 0);
  for (int p = 0; p < num_proj; ++p) {
    roff[p + 1] = roff[p] + nrow[p];
    coff[p + 1] = coff[p] + ncol[p];
    choff[p + 1] = choff[p] + (long)nrow[p] * ncol[p];
  }
  const long total_ncol = coff[num_proj];
  if (total_ncol == 0) return;
  std::vector<double> X((size_t)total_ncol * nb, 0.0);
  for (int p = 0; p < num_proj; ++p) {
    const int nr = nrow[p], nc = ncol[p];
    if (nr == 0 || nc == 0) continue;
    const int* idx = index_all + roff[p];
    const double* chi = chi_all + choff[p];
    std::vector<double> filtered((size_t)nr * nb);
    for (int b = 0; b < nb; ++b)
      for (int r = 0; r < nr; ++r)
        filtered[r + (size_t)b * nr] =
            psi16[((long)(b / 16) * K + idx[r]) * 16 + b % 16];
    for (int b = 0; b < nb; ++b)
      for (int c = 0; c < nc; ++c) {
        double sum = 0.0;
        for (int r = 0; r < nr; ++r)
          sum += chi[r + (size_t)c * nr] * filtered[r + (size_t)b * nr];
        double& x = X[(coff[p] + c) + (size_t)b * total_ncol];
        x = 1.0 * x + dv * sum;
      }
  }
  int atom = atom_index[0] - 1;
  for (int p = 0; p < num_proj; ++p) {
    if (atom == atom_index[p]) continue;
    atom = atom_index[p];
    const double* gamma = gamma_all + coff[p];
    for (int b = 0; b < nb; ++b)
      for (int c = 0; c < ncol[p]; ++c)
        X[(coff[p] + c) + (size_t)b * total_ncol] *= gamma[c];
  }
  for (int p = 0; p < num_proj; ++p) {
    const int nr = nrow[p], nc = ncol[p];
    if (nr == 0 || nc == 0) continue;
    const int* idx = index_all + roff[p];
    const double* chi = chi_all + choff[p];
    std::vector<double> temp((size_t)nr * nb, 0.0);
    for (int b = 0; b < nb; ++b)
      for (int c = 0; c < nc; ++c) {
        const double bv = 1.0 * X[(coff[p] + c) + (size_t)b * total_ncol];
        for (int r = 0; r < nr; ++r)
          temp[r + (size_t)b * nr] += chi[r + (size_t)c * nr] * bv;
      }
    for (int b = 0; b < nb; ++b)
      for (int r = 0; r < nr; ++r)
        result[((long)(b / out_tile) * K + idx[r]) * out_tile + b % out_tile] +=
            temp[r + (size_t)b * nr];
  }
}

// ---------------------------------------------------------------------------
// Real-space stencil on the packed extended grid
// (include/stencil_extended_packed.hpp, include/halo_packed.hpp).
// Local grid [ceil(nb/16)][K][16], K = ni nj nk, g = i + j ni + k ni nj;
// extended grid [ceil(nb/16)][K_ex][16], (ni+2F)(nj+2F)(nk+2F), interior at
// offset (F, F, F).
// ---------------------------------------------------------------------------

// halo_packed.hpp:86-123 (copy_interior_to_extended_packed), then the halo
// of a single periodic domain (every ghost point holds the interior point at
// the wrapped coordinates; the reference's single-rank test fills the x
// faces this way, halo_packed_test.cpp:84-150) or zero (boundary 0).  Whole
// 16-band rows move, band lanes past nb included, as in the reference.
void oracle_halo_extend(int ni, int nj, int nk, int nb, int F,
                        const double* local16, double* ext16, int periodic) {
  const long K = (long)ni * nj * nk;
  const int nie = ni + 2 * F, nje = nj + 2 * F, nke = nk + 2 * F;
  const long Kex = (long)nie * nje * nke;
  const int tiles = (nb + 15) / 16;
  for (int t = 0; t < tiles; ++t)
    for (int ke = 0; ke < nke; ++ke)
      for (int je = 0; je < nje; ++je)
        for (int ie = 0; ie < nie; ++ie) {
          const long gex = ie + (long)je * nie + (long)ke * nie * nje;
          const int i = ie - F, j = je - F, k = ke - F;
          const bool inside = i >= 0 && i < ni && j >= 0 && j < nj && k >= 0 &&
                              k < nk;
          for (int bl = 0; bl < 16; ++bl) {
            double v = 0.0;
            if (inside || periodic) {
              const int ii = ((i % ni) + ni) % ni, jj = ((j % nj) + nj) % nj,
                        kk = ((k % nk) + nk) % nk;
              const long g = ii + (long)jj * ni + (long)kk * ni * nj;
              v = local16[(t * K + g) * 16 + bl];
            }
            ext16[(t * Kex + gex) * 16 + bl] = v;
          }
        }
}

// stencil_extended_packed.hpp:199-253 (the scalar version, any FDn):
// out[g, b] = psi_ex[g_ex, b] (A[g] + c0), then + cx[r] (psi_ex[g_ex + r] +
// psi_ex[g_ex - r]) for r = 1..F, then the y terms (offsets r ni_ex), then
// the z terms (r ni_ex nj_ex); out packed by out_tile (32: :47-197,
// 16: :276-), bands b < nb only.
void oracle_stencil_apply(int ni, int nj, int nk, int nb, int F,
                          const double* ext16, const double* A_diag,
                          const double* cx, const double* cy, const double* cz,
                          double c0, double* out, int out_tile) {
  const long K = (long)ni * nj * nk;
  const long nie = ni + 2 * F, nje = nj + 2 * F;
  const long Kex = nie * nje * (nk + 2 * F);
  const long sj = nie, sk = nie * nje;
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) {
        const long g = i + (long)j * ni + (long)k * ni * nj;
        const long gex = (i + F) + (j + F) * sj + (k + F) * sk;
        const double a = A_diag[g] + c0;
        for (int b = 0; b < nb; ++b) {
          const double* s = ext16 + (long)(b / 16) * Kex * 16 + b % 16;
          double res = s[gex * 16] * a;
          for (int r = 1; r <= F; ++r)
            res += cx[r] * (s[(gex + r) * 16] + s[(gex - r) * 16]);
          for (int r = 1; r <= F; ++r)
            res += cy[r] * (s[(gex + r * sj) * 16] + s[(gex - r * sj) * 16]);
          for (int r = 1; r <= F; ++r)
            res += cz[r] * (s[(gex + r * sk) * 16] + s[(gex - r * sk) * 16]);
          out[((long)(b / out_tile) * K + g) * out_tile + b % out_tile] = res;
        }
      }
}

// ---------------------------------------------------------------------------
// Split form (Ozaki scheme II), exact integer product and its conversion to
// FP64 (the step oracle/ozaki.py's exact_product + to_double restates with
// Python integers), for sizes where Python integers are too slow.  xa (M x K)
// and xb (N x K) hold the scaled integers rint(x 2^sigma), |x| < 2^55, row
// per free index.  The sum of K products is exact in 128-bit arithmetic:
// every partial sum is bounded by ||xa_i|| ||xb_j|| (Cauchy-Schwarz on the
// prefix), which the scaling keeps below 2^110.  out[i N + j] =
// ldexp(fl(d / (m_a m_b)), -(e_a + e_b - 6)) with d = fl(fl(hi) 2^64 +
// fl(lo)), hi = C >> 64 (signed), lo = C mod 2^64, and the scale codes
// sa_i = 8 e_a + m_a - 8 (s = 2^(e - 3) m): the conversion and unscaling
// the library's CRT pass performs.  No counterpart in the reference (its
// path is FP64 only).
// ---------------------------------------------------------------------------
void oracle_ozaki_exact(int M, int N, int K, const int64_t* xa,
                        const int64_t* xb, const int* sa, const int* sb,
                        double* out) {
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      const int64_t* a = xa + (long)i * K;
      const int64_t* b = xb + (long)j * K;
      __int128 acc = 0;
      for (int k = 0; k < K; ++k) acc += (__int128)a[k] * b[k];
      const int64_t hi = (int64_t)(acc >> 64);
      const uint64_t lo = (uint64_t)acc;
      const double d = (double)hi * 18446744073709551616.0 + (double)lo;
      const int ma = 8 + (sa[i] & 7), mb = 8 + (sb[j] & 7);
      out[(long)i * N + j] =
          std::ldexp(d / (double)(ma * mb), -((sa[i] >> 3) + (sb[j] >> 3) - 6));
    }
}

}  // extern "C"
