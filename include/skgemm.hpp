// skgemm.hpp - header-only C++ facade with the reference's class names and
// call signatures, implemented on the C ABI in skgemm.h (no HIP types, no
// HIP headers: a host program links libskgemm.so and includes this file).
//
// Reference surface kept (shriramch/stream-k-gemm):
//   gemm<T>                 gemm_crsp.cpp:57-431, gemm_opt.cpp:55-426
//   sme_syrk<T>             include/sme_syrk.hpp:22-150, syrk_crsp.cpp:51-349
//   sme_gemm_trp_3d<T>      include/sme_trp_3d.hpp:48-220
//   sme_pack<T>             include/sme_pack.hpp:30-352
//   Hamiltonian::nloc_project_vectors_packed(_16to16)
//                           include/nloc_packed_opt2.hpp:229-655
//   Stencil_method::Packed  include/stencil_extended_packed.hpp:47-402,
//                           include/halo_packed.hpp:86-123
//   hbm_init/hbm_alloc/...  include/hbm_alloc.hpp:14-50
// Only T = double is provided (the reference instantiates double only).
//
// Conventions, as in the reference: host pointers, void returns, same
// layouts (A column-major / B row-major / D row-major for gemm::compute_test;
// PA [M/16][K][16] and PB [N/32][K][32] for the packed entry points; P16 for
// sme_syrk; grid-packed in / band-packed out for the rotation).  Each call
// copies its operands to the GPU, runs the Stream-K kernel and copies the
// result back.  A failing call cannot report through a void signature, so
// it prints the library's error text and aborts.
//
// Differences: sme_syrk applies alpha to the whole product (the reference
// applies it to the last KC block only, include/sme_syrk.hpp:123-133; the
// two agree for alpha = 1 or K <= 256) and never writes outside the M x M
// matrix (include/sme_syrk.hpp:66-108 does when M % 16 != 0).  The gemm
// entry points keep the reference's truncation to 16*floor(M/16) rows and
// 32*floor(N/32) columns (gemm_crsp.cpp:197-198,375-376) with gamma on the
// remaining entries; the C ABI itself computes every row and column.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "skgemm.h"

namespace skgemm_facade {

// One context per host thread (contexts are not shared across threads).
inline skgemm_ctx* ctx() {
  static thread_local skgemm_ctx* c = nullptr;
  if (c == nullptr) {
    const char* dev = std::getenv("SKGEMM_DEVICE");
    if (skgemm_create(dev ? std::atoi(dev) : 0, &c) != SKGEMM_OK ||
        // sme_syrk writes whole 16 x 16 diagonal blocks, as the reference's
        // tile store does (include/sme_syrk.hpp:57-109).
        skgemm_set_syrk_store(c, SKGEMM_SYRK_DIAG_BLOCKS) != SKGEMM_OK) {
      std::fprintf(stderr, "skgemm: %s\n", skgemm_last_error());
      std::abort();
    }
  }
  return c;
}

inline void check(int rc, const char* what) {
  if (rc != SKGEMM_OK) {
    std::fprintf(stderr, "skgemm: %s failed: %s\n", what, skgemm_last_error());
    std::abort();
  }
}

}  // namespace skgemm_facade

// ---- include/hbm_alloc.hpp:14-50 ------------------------------------------
// Host allocations, as in the reference without USE_HBM; operands are moved
// to HBM by each call.
inline void hbm_init() {}
template <typename T>
inline T* hbm_alloc(size_t count) {
  return new T[count];
}
template <typename T>
inline T* hbm_calloc(size_t count) {
  return new T[count]();
}
template <typename T>
inline void hbm_free(T* ptr) {
  delete[] ptr;
}

// ---- gemm_crsp.cpp / gemm_opt.cpp -------------------------------------------
template <typename T>
class gemm;

template <>
class gemm<double> {
 public:
  using T = double;
  static constexpr int ZA_TILE_M = 16;  // PA panel width (gemm_crsp.cpp:62)
  static constexpr int ZA_TILE_N = 32;  // PB panel width (gemm_crsp.cpp:63)

  // Rows >= Mt and columns >= Nt of the row-major M x N D: the reference's
  // reduction sums zero partials there (gemm_crsp.cpp:285-310), giving
  // alpha * 0 + gamma (one exact product, so fused or not it is the same).
  static void edges(int M, int N, int Mt, int Nt, T* D, T alpha, T gamma) {
    const T v = alpha * 0.0 + gamma;
    for (int i = 0; i < M; ++i)
      for (int j = (i < Mt ? Nt : 0); j < N; ++j) D[(size_t)i * N + j] = v;
  }

  // gemm_crsp.cpp:261-311.  C_local is the reference's per-thread workspace;
  // the GPU path keeps its workspace inside the library, so it is unused.
  static void compute_streamk_packed_core(const int M, const int N,
                                          const int K, const T* A_packed,
                                          const T* B_packed, T* D,
                                          const T alpha, const T gamma,
                                          T** /*C_local*/) {
    compute_streamk_packed(M, N, K, A_packed, B_packed, D, alpha, gamma);
  }

  // gemm_crsp.cpp:316-338: D = alpha * A * B + gamma on PA / PB.  PA holds
  // floor(M/16) panels and PB floor(N/32) (gemm_crsp.cpp:375-376), so the
  // product covers the first 16*floor(M/16) rows and 32*floor(N/32)
  // columns; the reference's reduction still writes every element of D
  // (gemm_crsp.cpp:285-310), with a zero sum outside those panels, i.e.
  // alpha * 0 + gamma there.
  static void compute_streamk_packed(const int M, const int N, const int K,
                                     const T* A_packed, const T* B_packed,
                                     T* D, const T alpha, const T gamma) {
    const int Mt = (M / ZA_TILE_M) * ZA_TILE_M, Nt = (N / ZA_TILE_N) * ZA_TILE_N;
    if (Mt > 0 && Nt > 0)
      skgemm_facade::check(
          skgemm_dgemm_host(skgemm_facade::ctx(), SKGEMM_PACK16, SKGEMM_PACK32,
                            SKGEMM_ROW_MAJOR, Mt, Nt, K, alpha, A_packed,
                            K > 0 ? K : 1, B_packed, K > 0 ? K : 1, 0.0, D, N,
                            gamma),
          "gemm::compute_streamk_packed");
    edges(M, N, Mt, Nt, D, alpha, gamma);
  }

  // gemm_opt.cpp:314-337 (same contract, M-tile decomposition there).
  static void compute_packed(const int M, const int N, const int K,
                             const T* A_packed, const T* B_packed, T* D,
                             const T alpha, const T gamma) {
    compute_streamk_packed(M, N, K, A_packed, B_packed, D, alpha, gamma);
  }

  // gemm_crsp.cpp:342-353: column-major A -> [M/16][K][16] (truncating).
  static void pack_A(const int M, const int K, const T* A_colmajor,
                     T* A_packed) {
    const int tiles_m = M / ZA_TILE_M;
    for (int tm = 0; tm < tiles_m; ++tm)
      for (int k = 0; k < K; ++k)
        std::memcpy(A_packed + ((size_t)tm * K + k) * ZA_TILE_M,
                    A_colmajor + (size_t)tm * ZA_TILE_M + (size_t)k * M,
                    ZA_TILE_M * sizeof(T));
  }

  // gemm_crsp.cpp:355-366: row-major B -> [N/32][K][32] (truncating).
  static void pack_B(const int N, const int K, const T* B_rowmajor,
                     T* B_packed) {
    const int tiles_n = N / ZA_TILE_N;
    for (int tn = 0; tn < tiles_n; ++tn)
      for (int k = 0; k < K; ++k)
        std::memcpy(B_packed + ((size_t)tn * K + k) * ZA_TILE_N,
                    B_rowmajor + (size_t)k * N + (size_t)tn * ZA_TILE_N,
                    ZA_TILE_N * sizeof(T));
  }

  // gemm_crsp.cpp:369-386: A column-major M x K, B row-major K x N, D
  // row-major M x N, D = alpha * A * B + gamma.  The reference packs A and B
  // (truncating to whole panels) and runs compute_streamk_packed, so the
  // same rows and columns come out: the product on the first
  // 16*floor(M/16) rows and 32*floor(N/32) columns, alpha * 0 + gamma on
  // the rest.  The GPU reads the unpacked layouts directly (no packing
  // pass); skgemm_dgemm_host computes every row and column for a caller
  // that wants them.
  static void compute_test(const int M, const int N, const int K,
                           const T* A_colmajor, const T* B_rowmajor, T* D,
                           const T alpha, const T gamma) {
    const int Mt = (M / ZA_TILE_M) * ZA_TILE_M, Nt = (N / ZA_TILE_N) * ZA_TILE_N;
    if (Mt > 0 && Nt > 0)
      skgemm_facade::check(
          skgemm_dgemm_host(skgemm_facade::ctx(), SKGEMM_COL_MAJOR,
                            SKGEMM_ROW_MAJOR, SKGEMM_ROW_MAJOR, Mt, Nt, K,
                            alpha, A_colmajor, M, B_rowmajor, N, 0.0, D, N,
                            gamma),
          "gemm::compute_test");
    edges(M, N, Mt, Nt, D, alpha, gamma);
  }

  // gemm_crsp.cpp:388-430: ns per call, operands resident on the device,
  // WITERS warm-up calls then ITERS timed calls (5 and 3 in the reference),
  // kernel time from HIP events.
  static double compute_benchmark(const int M, const int N, const int K,
                                  const T* A_colmajor, const T* B_rowmajor,
                                  T* D, const T alpha, const T gamma,
                                  int witers = 5, int iters = 3) {
    skgemm_ctx* c = skgemm_facade::ctx();
    void *dA, *dB, *dD;
    skgemm_facade::check(skgemm_dev_alloc(c, sizeof(T) * M * K, 0, &dA), "alloc");
    skgemm_facade::check(skgemm_dev_alloc(c, sizeof(T) * K * N, 0, &dB), "alloc");
    skgemm_facade::check(skgemm_dev_alloc(c, sizeof(T) * M * N, 1, &dD), "alloc");
    skgemm_memcpy_h2d(c, dA, A_colmajor, sizeof(T) * M * K);
    skgemm_memcpy_h2d(c, dB, B_rowmajor, sizeof(T) * K * N);
    // Same rows and columns as compute_test (whole panels; gamma on the
    // rest after the copy back).
    const int Mt = (M / ZA_TILE_M) * ZA_TILE_M, Nt = (N / ZA_TILE_N) * ZA_TILE_N;
    auto run = [&] {
      if (Mt == 0 || Nt == 0) return;
      skgemm_facade::check(
          skgemm_dgemm(c, SKGEMM_COL_MAJOR, SKGEMM_ROW_MAJOR, SKGEMM_ROW_MAJOR,
                       Mt, Nt, K, alpha, (const T*)dA, M, (const T*)dB, N, 0.0,
                       (T*)dD, N, gamma),
          "gemm::compute_benchmark");
    };
    for (int i = 0; i < witers; ++i) run();
    double total_ms = 0.0;
    for (int i = 0; i < iters; ++i) {
      run();
      double ms = 0.0;
      skgemm_last_kernel_ms(c, &ms);
      total_ms += ms;
    }
    skgemm_facade::check(skgemm_synchronize(c), "synchronize");
    skgemm_memcpy_d2h(c, D, dD, sizeof(T) * M * N);
    edges(M, N, Mt, Nt, D, alpha, gamma);
    skgemm_dev_free(c, dA);
    skgemm_dev_free(c, dB);
    skgemm_dev_free(c, dD);
    return total_ms * 1e6 / (iters > 0 ? iters : 1);
  }
};

// ---- include/sme_syrk.hpp / syrk_crsp.cpp -----------------------------------
template <typename T>
class sme_syrk;

template <>
class sme_syrk<double> {
 public:
  using T = double;
  static constexpr int TILE_M = 16;

  // include/sme_syrk.hpp:137-149: C = alpha * A^T * A + beta * C on the
  // upper triangle (m1 <= m2) of the M x M column-major C and on the whole
  // of every 16 x 16 diagonal block, as the reference's store writes them
  // (:57-109); A is P16 [ceil(M/16)][K][16].  Column-major upper ==
  // row-major lower, which is what the C ABI writes; the facade's context
  // selects SKGEMM_SYRK_DIAG_BLOCKS for the diagonal blocks.
  static void compute(const int K, const int M, const T* A_packed, T* C,
                      const T alpha, const T beta) {
    skgemm_facade::check(
        skgemm_dsyrk_host(skgemm_facade::ctx(), SKGEMM_PACK16, M, K, alpha,
                          A_packed, K > 0 ? K : 1, beta, C, M > 0 ? M : 1),
        "sme_syrk::compute");
  }
};

// ---- include/sme_trp_3d.hpp ---------------------------------------------------
template <typename T>
class sme_gemm_trp_3d;

template <>
class sme_gemm_trp_3d<double> {
 public:
  using T = double;
  // include/sme_trp_3d.hpp:194-215: D[g, bp] = sum_b A[g, b] * B[b, bp] with
  // A grid-packed [Nd/16][nb][16], B row-major [nb][nb], D band-packed
  // [nb/16][Nd][16]; no alpha / gamma.
  static void compute(const int Nd, const int nb, const T* A, const T* B,
                      T* D) {
    skgemm_facade::check(
        skgemm_dgemm_host(skgemm_facade::ctx(), SKGEMM_PACK16,
                          SKGEMM_ROW_MAJOR, SKGEMM_PACK16, Nd, nb, nb, 1.0, A,
                          nb > 0 ? nb : 1, B, nb > 0 ? nb : 1, 0.0, D,
                          Nd > 0 ? Nd : 1, 0.0),
        "sme_gemm_trp_3d::compute");
  }
};
template <typename T>
using sme_gemm_trp = sme_gemm_trp_3d<T>;

// ---- include/sme_pack.hpp ----------------------------------------------------
namespace skgemm_facade {

// Host-pointer layout conversion on the device: copies src (n_src doubles)
// and dst (n_dst doubles, so elements the conversion does not write keep
// their values, as in the reference) to HBM, runs `op`, copies dst back.
template <typename Op>
inline void convert_host(const double* src, size_t n_src, double* dst,
                         size_t n_dst, Op op, const char* what) {
  skgemm_ctx* c = ctx();
  void *ds, *dd;
  check(skgemm_dev_alloc(c, n_src * 8 + 8, 0, &ds), "alloc");
  check(skgemm_dev_alloc(c, n_dst * 8 + 8, 0, &dd), "alloc");
  if (n_src) skgemm_memcpy_h2d(c, ds, src, n_src * 8);
  if (n_dst) skgemm_memcpy_h2d(c, dd, dst, n_dst * 8);
  check(op(c, (const double*)ds, (double*)dd), what);
  check(skgemm_synchronize(c), "synchronize");
  if (n_dst) skgemm_memcpy_d2h(c, dst, dd, n_dst * 8);
  skgemm_dev_free(c, ds);
  skgemm_dev_free(c, dd);
}

}  // namespace skgemm_facade

template <typename T>
class sme_pack;

// Layouts (include/sme_pack.hpp:12-28): column-major [Nd, nb] (g + b*Nd;
// the reference's row-major [nb][Nd] is the same memory), grid-packed
// [Nd/16][nb][16] ((g/16)*nb*16 + b*16 + g%16), band-packed [nb/16][Nd][16]
// ((b/16)*Nd*16 + g*16 + b%16).  As in the reference, only whole 16-row
// (16-band) panels are converted: Nd/16 grid tiles, nb/16 band tiles.
template <>
class sme_pack<double> {
 public:
  using T = double;
  static constexpr int PACK_G = 16;

  // include/sme_pack.hpp:85-95: column-major [Nd, nb] -> grid-packed.
  // skgemm_pack addresses element (r, o) = (band b, grid g); src[g + b*Nd]
  // is r*ld + o with ld = Nd, i.e. ROW_MAJOR in that convention.
  static void pack_c16(const int Nd, const int nb, const T* src, T* dst) {
    const int g16 = (Nd / 16) * 16;
    skgemm_facade::convert_host(
        src, (size_t)Nd * nb, dst, (size_t)g16 * nb,
        [&](skgemm_ctx* c, const double* s, double* d) {
          return skgemm_pack(c, SKGEMM_ROW_MAJOR, s, Nd > 0 ? Nd : 1,
                             SKGEMM_PACK16, d, nb > 0 ? nb : 1, g16, nb);
        },
        "sme_pack::pack_c16");
  }
  // include/sme_pack.hpp:97-107: row-major [nb][Nd] holds the same memory.
  static void pack_r16(const int Nd, const int nb, const T* src, T* dst) {
    pack_c16(Nd, nb, src, dst);
  }
  // include/sme_pack.hpp:130-140: grid-packed -> column-major.
  static void unpack_c16(const int Nd, const int nb, const T* src, T* dst) {
    const int g16 = (Nd / 16) * 16;
    skgemm_facade::convert_host(
        src, (size_t)g16 * nb, dst, (size_t)Nd * nb,
        [&](skgemm_ctx* c, const double* s, double* d) {
          return skgemm_pack(c, SKGEMM_PACK16, s, nb > 0 ? nb : 1,
                             SKGEMM_ROW_MAJOR, d, Nd > 0 ? Nd : 1, g16, nb);
        },
        "sme_pack::unpack_c16");
  }
  static void unpack_r16(const int Nd, const int nb, const T* src, T* dst) {
    unpack_c16(Nd, nb, src, dst);
  }
  // include/sme_pack.hpp:232-244: band-packed -> grid-packed.
  static void band_to_grid_tile(const int Nd, const int nb, const T* src,
                                T* dst) {
    const int g16 = (Nd / 16) * 16, b16 = (nb / 16) * 16;
    skgemm_facade::convert_host(
        src, (size_t)b16 * Nd, dst, (size_t)g16 * nb,
        [&](skgemm_ctx* c, const double* s, double* d) {
          return skgemm_transpose(c, SKGEMM_PACK16, s, Nd > 0 ? Nd : 1,
                                  SKGEMM_PACK16, d, nb > 0 ? nb : 1, b16, g16);
        },
        "sme_pack::band_to_grid_tile");
  }
  // include/sme_pack.hpp:300-312: grid-packed -> band-packed.
  static void grid_to_band_tile(const int Nd, const int nb, const T* src,
                                T* dst) {
    const int g16 = (Nd / 16) * 16, b16 = (nb / 16) * 16;
    skgemm_facade::convert_host(
        src, (size_t)g16 * nb, dst, (size_t)b16 * Nd,
        [&](skgemm_ctx* c, const double* s, double* d) {
          return skgemm_transpose(c, SKGEMM_PACK16, s, nb > 0 ? nb : 1,
                                  SKGEMM_PACK16, d, Nd > 0 ? Nd : 1, g16, b16);
        },
        "sme_pack::grid_to_band_tile");
  }
};

// include/sme_pack.hpp:316-352 (C wrappers).
inline void linalg_pack_c16_f64(const int Nd, const int nb, const double* src,
                                double* dst) {
  sme_pack<double>::pack_c16(Nd, nb, src, dst);
}
inline void linalg_pack_r16_f64(const int Nd, const int nb, const double* src,
                                double* dst) {
  sme_pack<double>::pack_r16(Nd, nb, src, dst);
}
inline void linalg_unpack_c16_f64(const int Nd, const int nb,
                                  const double* src, double* dst) {
  sme_pack<double>::unpack_c16(Nd, nb, src, dst);
}
inline void linalg_unpack_r16_f64(const int Nd, const int nb,
                                  const double* src, double* dst) {
  sme_pack<double>::unpack_r16(Nd, nb, src, dst);
}

// ---- include/nloc_packed_opt2.hpp:229-655 -----------------------------------
// Hamiltonian::nloc_project_vectors_packed(_16to16) over the caller's own
// projector structures, read the way the reference reads them (a
// `nloc_projectors` vector whose elements have nrow, ncol, atom_index,
// chi.data (column-major nrow x ncol), gamma.data and index[], and an
// `offsets` vector), with host arrays: psi 16-band-packed
// [ceil(nb/16)][K][16], result [ceil(nb/T)][K][T], T = 32 or 16.  Each call
// uploads the projector set and the arrays, runs skgemm_nloc_apply and
// copies the result back.  The communicator is not used (one GPU holds the
// whole grid; the reference's MPI_Allreduce over ranks of one has no effect).
namespace skgemm_facade {

template <typename V>
inline void nloc_apply_host(double* result, int tile, const double* psi16,
                            const V& Vnloc, unsigned nb, unsigned K,
                            double dv) {
  const size_t np = Vnloc.nloc_projectors.size();
  if (np == 0 || nb == 0) return;
  std::vector<int> nrow(np), ncol(np), atom(np), index;
  std::vector<double> chi, gamma;
  for (size_t p = 0; p < np; ++p) {
    const auto& pr = Vnloc.nloc_projectors[p];
    nrow[p] = (int)pr.nrow;
    ncol[p] = (int)pr.ncol;
    atom[p] = (int)pr.atom_index;
    for (unsigned r = 0; r < pr.nrow; ++r) index.push_back((int)pr.index[r]);
    chi.insert(chi.end(), pr.chi.data, pr.chi.data + (size_t)pr.nrow * pr.ncol);
    gamma.insert(gamma.end(), pr.gamma.data, pr.gamma.data + pr.ncol);
  }
  skgemm_ctx* c = ctx();
  skgemm_nloc* h = nullptr;
  check(skgemm_nloc_create(c, (int)K, (int)np, nrow.data(), ncol.data(),
                           atom.data(), index.data(), chi.data(), gamma.data(),
                           &h),
        "nloc_project_vectors_packed");
  const size_t n16 = (size_t)((nb + 15) / 16) * K * 16;
  const size_t nt = (size_t)((nb + tile - 1) / tile) * K * tile;
  void *dp, *dr;
  check(skgemm_dev_alloc(c, n16 * 8, 0, &dp), "alloc");
  check(skgemm_dev_alloc(c, nt * 8, 0, &dr), "alloc");
  skgemm_memcpy_h2d(c, dp, psi16, n16 * 8);
  skgemm_memcpy_h2d(c, dr, result, nt * 8);
  check(skgemm_nloc_apply(h, (int)nb, dv, (const double*)dp, (double*)dr, tile),
        "nloc_project_vectors_packed");
  check(skgemm_synchronize(c), "synchronize");
  skgemm_memcpy_d2h(c, result, dr, nt * 8);
  skgemm_dev_free(c, dp);
  skgemm_dev_free(c, dr);
  skgemm_nloc_destroy(h);
}

}  // namespace skgemm_facade

namespace Hamiltonian {

template <typename T, typename V, typename Comm>
void nloc_project_vectors_packed(T* result_32packed, const T* psi_16packed,
                                 const V& Vnloc, unsigned nb, unsigned K, T dv,
                                 const Comm&) {
  static_assert(sizeof(T) == 8, "double only");
  skgemm_facade::nloc_apply_host(result_32packed, 32, psi_16packed, Vnloc, nb,
                                 K, dv);
}

template <typename T, typename V, typename Comm>
void nloc_project_vectors_packed_16to16(T* result_16packed,
                                        const T* psi_16packed, const V& Vnloc,
                                        unsigned nb, unsigned K, T dv,
                                        const Comm&) {
  static_assert(sizeof(T) == 8, "double only");
  skgemm_facade::nloc_apply_host(result_16packed, 16, psi_16packed, Vnloc, nb,
                                 K, dv);
}

}  // namespace Hamiltonian

// ---- include/stencil_extended_packed.hpp, include/halo_packed.hpp ----------
// The reference's stencil and interior copy with host arrays: extended
// psi [ceil(nb/16)][K_ex][16], output [ceil(nb/T)][K][T], A_diag [K].  Each
// call copies its operands to the GPU and the result back (elements the
// kernel does not write keep their values).
namespace skgemm_facade {

inline void stencil_host(const double* ext, double* out, int tile,
                         const double* A, unsigned ni, unsigned nj,
                         unsigned nk, unsigned nb, int FDn, const double* cx,
                         const double* cy, const double* cz, double c0) {
  if (nb == 0) return;
  const size_t K = (size_t)ni * nj * nk;
  const size_t Kex = (size_t)(ni + 2 * FDn) * (nj + 2 * FDn) * (nk + 2 * FDn);
  const size_t ne = (size_t)((nb + 15) / 16) * Kex * 16;
  const size_t no = (size_t)((nb + tile - 1) / tile) * K * tile;
  skgemm_ctx* c = ctx();
  void *de, *da, *dout;
  check(skgemm_dev_alloc(c, ne * 8, 0, &de), "alloc");
  check(skgemm_dev_alloc(c, K * 8, 0, &da), "alloc");
  check(skgemm_dev_alloc(c, no * 8, 0, &dout), "alloc");
  skgemm_memcpy_h2d(c, de, ext, ne * 8);
  skgemm_memcpy_h2d(c, da, A, K * 8);
  skgemm_memcpy_h2d(c, dout, out, no * 8);
  check(skgemm_stencil_apply(c, (int)ni, (int)nj, (int)nk, (int)nb, FDn,
                             (const double*)de, (const double*)da, cx, cy, cz,
                             c0, (double*)dout, tile),
        "calc_laplacian_extended_packed");
  check(skgemm_synchronize(c), "synchronize");
  skgemm_memcpy_d2h(c, out, dout, no * 8);
  skgemm_dev_free(c, de);
  skgemm_dev_free(c, da);
  skgemm_dev_free(c, dout);
}

}  // namespace skgemm_facade

namespace Stencil_method {
namespace Packed {

// include/stencil_extended_packed.hpp:47-197 (any FDn 1..8 here).
inline void calc_laplacian_extended_packed_16to32(
    const double* extended_packed, double* output_32packed,
    const double* A_diag_packed, unsigned ni, unsigned nj, unsigned nk,
    unsigned nb, int FDn, const double* coef_x, const double* coef_y,
    const double* coef_z, double coef_0) {
  skgemm_facade::stencil_host(extended_packed, output_32packed, 32,
                              A_diag_packed, ni, nj, nk, nb, FDn, coef_x,
                              coef_y, coef_z, coef_0);
}
// include/stencil_extended_packed.hpp:199-253: the same operator.
inline void calc_laplacian_extended_packed_16to32_scalar(
    const double* extended_packed, double* output_32packed,
    const double* A_diag_packed, unsigned ni, unsigned nj, unsigned nk,
    unsigned nb, int FDn, const double* coef_x, const double* coef_y,
    const double* coef_z, double coef_0) {
  calc_laplacian_extended_packed_16to32(extended_packed, output_32packed,
                                        A_diag_packed, ni, nj, nk, nb, FDn,
                                        coef_x, coef_y, coef_z, coef_0);
}
// include/stencil_extended_packed.hpp:276-402.
inline void calc_laplacian_extended_packed_16to16(
    const double* extended_packed, double* output_16packed,
    const double* A_diag_packed, unsigned ni, unsigned nj, unsigned nk,
    unsigned nb, int FDn, const double* coef_x, const double* coef_y,
    const double* coef_z, double coef_0) {
  skgemm_facade::stencil_host(extended_packed, output_16packed, 16,
                              A_diag_packed, ni, nj, nk, nb, FDn, coef_x,
                              coef_y, coef_z, coef_0);
}

// include/halo_packed.hpp:86-123: interior copy, ghost points untouched.
template <typename T>
void copy_interior_to_extended_packed(const T* local_packed,
                                      T* extended_packed, unsigned ni,
                                      unsigned nj, unsigned nk, unsigned nb,
                                      int FDn) {
  static_assert(sizeof(T) == 8, "double only");
  if (nb == 0) return;
  const size_t K = (size_t)ni * nj * nk;
  const size_t Kex = (size_t)(ni + 2 * FDn) * (nj + 2 * FDn) * (nk + 2 * FDn);
  const size_t t = (nb + 15) / 16;
  skgemm_facade::convert_host(
      (const double*)local_packed, t * K * 16, (double*)extended_packed,
      t * Kex * 16,
      [&](skgemm_ctx* c, const double* s, double* d) {
        return skgemm_halo_extend(c, (int)ni, (int)nj, (int)nk, (int)nb, FDn,
                                  s, d, SKGEMM_HALO_KEEP);
      },
      "copy_interior_to_extended_packed");
}

}  // namespace Packed
}  // namespace Stencil_method
