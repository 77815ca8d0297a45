// Host program written against the C++ facade exactly as a caller of the
// reference classes would be (gemm<double>, sme_syrk<double>,
// sme_gemm_trp_3d<double>, sme_pack<double>, linalg_pack_*,
// Hamiltonian::nloc_project_vectors_packed, Stencil_method::Packed): it checks
// each call against
// closed forms or naive loops computed here and exits non-zero on a
// mismatch.  Built with g++ against libskgemm.so (no HIP headers needed).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "skgemm.hpp"

static int failures = 0;

static void expect(bool ok, const char* what, double v) {
  std::printf("%-44s %s (%.3e)\n", what, ok ? "ok" : "FAIL", v);
  if (!ok) ++failures;
}

// Projector structures shaped like the reference's (only the members the
// facade reads).
struct NlocArray2 {
  double* data = nullptr;
};
struct NlocArray1 {
  double* data = nullptr;
};
struct NlocProj {
  unsigned nrow = 0, ncol = 0;
  int atom_index = 0;
  NlocArray2 chi;
  NlocArray1 gamma;
  std::vector<unsigned> index;
  std::vector<double> chi_store, gamma_store;
};
struct NlocV {
  std::vector<NlocProj> nloc_projectors;
  std::vector<unsigned> offsets;
};

int main() {
  // gemm_crsp.cpp:451-460 inputs; closed form
  // D_ij = alpha (1e-5 (I+J) S1 + 1e-4 I J K + 1e-6 S2) + gamma.
  {
    const int M = 32, N = 64, K = 4096;
    const double alpha = 2.0, gamma = 0.5;
    std::vector<double> A((size_t)M * K), B((size_t)K * N), D((size_t)M * N);
    for (int k = 0; k < K; ++k)
      for (int i = 0; i < M; ++i) A[i + (size_t)k * M] = 0.01 * (i + 1) + 0.001 * (k + 1);
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < N; ++j) B[(size_t)k * N + j] = 0.001 * (k + 1) + 0.01 * (j + 1);
    gemm<double>::compute_test(M, N, K, A.data(), B.data(), D.data(), alpha, gamma);
    const long double S1 = (long double)K * (K + 1) / 2,
                      S2 = (long double)K * (K + 1) * (2 * K + 1) / 6;
    double worst = 0.0;
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < N; ++j) {
        const long double I = i + 1, J = j + 1;
        const long double want =
            alpha * (1e-5L * (I + J) * S1 + 1e-4L * I * J * K + 1e-6L * S2) + gamma;
        worst = std::fmax(worst, (double)std::fabs((D[(size_t)i * N + j] - want) / want));
      }
    expect(worst < 1e-13, "gemm::compute_test vs closed form", worst);

    std::vector<double> PA((size_t)M * K), PB((size_t)K * N), D2((size_t)M * N);
    gemm<double>::pack_A(M, K, A.data(), PA.data());
    gemm<double>::pack_B(N, K, B.data(), PB.data());
    gemm<double>::compute_streamk_packed(M, N, K, PA.data(), PB.data(), D2.data(),
                                         alpha, gamma);
    double diff = 0.0;
    for (size_t e = 0; e < D.size(); ++e)
      diff = std::fmax(diff, std::fabs(D2[e] - D[e]) / std::fabs(D[e]));
    expect(diff < 1e-13, "gemm::compute_streamk_packed vs compute_test", diff);
    const double ns = gemm<double>::compute_benchmark(M, N, K, A.data(), B.data(),
                                                      D2.data(), alpha, gamma);
    expect(ns > 0.0, "gemm::compute_benchmark ns per call", ns);
  }
  // Ragged M = 20, N = 40 with buffers of the reference's sizes: PA holds
  // floor(M/16) = 1 panel, PB floor(N/32) = 1 (gemm_crsp.cpp:375-376).  The
  // product covers rows < 16 and columns < 32; every other entry is
  // alpha * 0 + gamma, as the reference's reduction writes it
  // (gemm_crsp.cpp:285-310).
  {
    const int M = 20, N = 40, K = 300;
    const double alpha = -1.5, gamma = 0.75;
    std::vector<double> A((size_t)M * K), B((size_t)K * N);
    for (int k = 0; k < K; ++k)
      for (int i = 0; i < M; ++i) A[i + (size_t)k * M] = 0.01 * (i + 1) + 0.001 * (k + 1);
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < N; ++j) B[(size_t)k * N + j] = 0.001 * (k + 1) + 0.01 * (j + 1);
    std::vector<double> PA((size_t)(M / 16) * K * 16), PB((size_t)(N / 32) * K * 32);
    gemm<double>::pack_A(M, K, A.data(), PA.data());
    gemm<double>::pack_B(N, K, B.data(), PB.data());
    const long double S1 = (long double)K * (K + 1) / 2,
                      S2 = (long double)K * (K + 1) * (2 * K + 1) / 6;
    for (int form = 0; form < 3; ++form) {
      std::vector<double> D((size_t)M * N, -99.0);
      if (form == 0)
        gemm<double>::compute_streamk_packed(M, N, K, PA.data(), PB.data(), D.data(),
                                             alpha, gamma);
      else if (form == 1)
        gemm<double>::compute_test(M, N, K, A.data(), B.data(), D.data(), alpha, gamma);
      else
        gemm<double>::compute_benchmark(M, N, K, A.data(), B.data(), D.data(), alpha,
                                        gamma, 1, 1);
      double worst = 0.0, edge = 0.0;
      for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j) {
          const double got = D[(size_t)i * N + j];
          if (i < 16 && j < 32) {
            const long double I = i + 1, J = j + 1;
            const long double want =
                alpha * (1e-5L * (I + J) * S1 + 1e-4L * I * J * K + 1e-6L * S2) + gamma;
            worst = std::fmax(worst, (double)std::fabs((got - want) / want));
          } else {
            edge = std::fmax(edge, std::fabs(got - gamma));
          }
        }
      static const char* names[3] = {"compute_streamk_packed", "compute_test",
                                     "compute_benchmark"};
      char what[96];
      std::snprintf(what, sizeof what, "gemm::%s M=20 N=40 panels", names[form]);
      expect(worst < 1e-13, what, worst);
      std::snprintf(what, sizeof what, "gemm::%s M=20 N=40 gamma edges", names[form]);
      expect(edge == 0.0, what, edge);
    }
  }
  // sme_syrk<double>::compute: P16 input, C = alpha A^T A + beta C on the
  // column-major upper triangle and on the whole of every 16 x 16 diagonal
  // block (include/sme_syrk.hpp:57-109); other entries below it untouched.
  {
    const int K = 3000, M = 70, t = (M + 15) / 16;
    std::vector<double> Ad((size_t)K * M), P((size_t)t * K * 16, 0.0);
    unsigned s = 12345;
    for (auto& x : Ad) {
      s = s * 1103515245u + 12345u;
      x = ((s >> 8) & 0xFFFF) / 65536.0 - 0.5;
    }
    for (int m = 0; m < M; ++m)
      for (int k = 0; k < K; ++k) P[((size_t)(m / 16) * K + k) * 16 + m % 16] = Ad[(size_t)k * M + m];
    std::vector<double> C((size_t)M * M, 0.25);
    sme_syrk<double>::compute(K, M, P.data(), C.data(), 1.5, 2.0);
    double worst = 0.0, untouched = 0.0;
    for (int m2 = 0; m2 < M; ++m2)
      for (int m1 = 0; m1 < M; ++m1) {
        const double got = C[m1 + (size_t)m2 * M];
        if (m1 <= m2 || m1 / 16 == m2 / 16) {
          long double acc = 0;
          for (int k = 0; k < K; ++k) acc += (long double)Ad[(size_t)k * M + m1] * Ad[(size_t)k * M + m2];
          const double want = (double)(1.5L * acc + 0.5L);
          worst = std::fmax(worst, std::fabs(got - want) / std::fmax(1.0, std::fabs(want)));
        } else {
          untouched = std::fmax(untouched, std::fabs(got - 0.25));
        }
      }
    expect(worst < 1e-13, "sme_syrk::compute vs naive (long double)", worst);
    expect(untouched == 0.0, "sme_syrk::compute leaves m1 > m2 off the blocks", untouched);
  }
  // sme_gemm_trp_3d<double>::compute with Q = I: the band-packed output
  // holds the grid-packed input, re-laid out.
  {
    const int Nd = 256, nb = 48;
    std::vector<double> col((size_t)Nd * nb), grid((size_t)Nd * nb), Q((size_t)nb * nb, 0.0),
        D((size_t)Nd * nb);
    for (size_t e = 0; e < col.size(); ++e) col[e] = std::sin(0.001 * (double)e);
    linalg_pack_c16_f64(Nd, nb, col.data(), grid.data());
    for (int b = 0; b < nb; ++b) Q[(size_t)b * nb + b] = 1.0;
    sme_gemm_trp_3d<double>::compute(Nd, nb, grid.data(), Q.data(), D.data());
    double worst = 0.0;
    for (int g = 0; g < Nd; ++g)
      for (int b = 0; b < nb; ++b)
        worst = std::fmax(worst, std::fabs(D[((size_t)(b / 16) * Nd + g) * 16 + b % 16] -
                                           col[g + (size_t)b * Nd]));
    expect(worst == 0.0, "sme_gemm_trp_3d::compute with Q = I", worst);
    std::vector<double> back((size_t)Nd * nb);
    linalg_unpack_c16_f64(Nd, nb, grid.data(), back.data());
    double rt = 0.0;
    for (size_t e = 0; e < col.size(); ++e) rt = std::fmax(rt, std::fabs(back[e] - col[e]));
    expect(rt == 0.0, "linalg_pack/unpack_c16_f64 round trip", rt);
    // the rotation's band-packed output converted back to grid-packed
    std::vector<double> grid2((size_t)Nd * nb, -1.0);
    sme_pack<double>::band_to_grid_tile(Nd, nb, D.data(), grid2.data());
    double bg = 0.0;
    for (size_t e = 0; e < grid.size(); ++e) bg = std::fmax(bg, std::fabs(grid2[e] - grid[e]));
    expect(bg == 0.0, "sme_pack::band_to_grid_tile of the rotation output", bg);
  }
  // pack.cpp:190-258: value[g, b] = g + 1000 b through band -> grid -> band,
  // and the r16 wrappers (same memory as c16).
  {
    const int Nd = 80, nb = 48;
    std::vector<double> band((size_t)Nd * nb), grid((size_t)Nd * nb, 0.0),
        band2((size_t)Nd * nb, 0.0);
    for (int g = 0; g < Nd; ++g)
      for (int b = 0; b < nb; ++b)
        band[((size_t)(b / 16) * Nd + g) * 16 + b % 16] = g + b * 1000.0;
    sme_pack<double>::band_to_grid_tile(Nd, nb, band.data(), grid.data());
    sme_pack<double>::grid_to_band_tile(Nd, nb, grid.data(), band2.data());
    double bad = 0.0;
    for (int g = 0; g < Nd; ++g)
      for (int b = 0; b < nb; ++b)
        bad = std::fmax(bad, std::fabs(grid[((size_t)(g / 16) * nb + b) * 16 + g % 16] -
                                       (g + b * 1000.0)));
    for (size_t e = 0; e < band.size(); ++e) bad = std::fmax(bad, std::fabs(band2[e] - band[e]));
    expect(bad == 0.0, "sme_pack band <-> grid (pack.cpp pattern)", bad);
    std::vector<double> rows((size_t)Nd * nb), packed((size_t)Nd * nb), back((size_t)Nd * nb);
    for (size_t e = 0; e < rows.size(); ++e) rows[e] = std::cos(0.01 * (double)e);
    linalg_pack_r16_f64(Nd, nb, rows.data(), packed.data());
    linalg_unpack_r16_f64(Nd, nb, packed.data(), back.data());
    double rt = 0.0;
    for (int g = 0; g < Nd; ++g)
      for (int b = 0; b < nb; ++b)
        rt = std::fmax(rt, std::fabs(packed[((size_t)(g / 16) * nb + b) * 16 + g % 16] -
                                     rows[(size_t)b * Nd + g]));
    for (size_t e = 0; e < rows.size(); ++e) rt = std::fmax(rt, std::fabs(back[e] - rows[e]));
    expect(rt == 0.0, "linalg_pack/unpack_r16_f64", rt);
  }
  // Hamiltonian::nloc_project_vectors_packed on projector structures shaped
  // like the reference's (nloc_packed_opt2_test.cpp:38-78), against the
  // unpacked formulation computed here with long double sums.
  {
    const unsigned K = 500, nb = 48;
    NlocV V;
    std::vector<unsigned> perm(K);
    for (int p = 0; p < 3; ++p) {
      NlocProj pr;
      pr.nrow = 60 + 10 * p;
      pr.ncol = 4 + p;
      pr.atom_index = p == 2 ? 2 : 1;  // projectors 0 and 1 share an atom
      pr.chi_store.resize((size_t)pr.nrow * pr.ncol);
      pr.gamma_store.resize(pr.ncol);
      for (size_t e = 0; e < pr.chi_store.size(); ++e)
        pr.chi_store[e] = std::sin(0.37 * (double)e + p);
      for (unsigned c = 0; c < pr.ncol; ++c) pr.gamma_store[c] = 0.5 + 0.1 * c;
      for (unsigned g = 0; g < K; ++g) perm[g] = (g * 37u + 11u * (unsigned)p) % K;
      pr.index.assign(perm.begin(), perm.begin() + pr.nrow);
      pr.chi.data = pr.chi_store.data();
      pr.gamma.data = pr.gamma_store.data();
      V.nloc_projectors.push_back(pr);
    }
    for (auto& pr : V.nloc_projectors) {  // re-point after the vector settled
      pr.chi.data = pr.chi_store.data();
      pr.gamma.data = pr.gamma_store.data();
    }
    std::vector<double> psi((size_t)K * nb), psi16(((nb + 15) / 16) * K * 16, 0.0),
        res32(((nb + 31) / 32) * K * 32, 0.0);
    for (size_t e = 0; e < psi.size(); ++e) psi[e] = std::cos(0.013 * (double)e);
    for (unsigned b = 0; b < nb; ++b)
      for (unsigned g = 0; g < K; ++g)
        psi16[((size_t)(b / 16) * K + g) * 16 + b % 16] = psi[g + (size_t)b * K];
    Hamiltonian::nloc_project_vectors_packed<double>(res32.data(), psi16.data(),
                                                     V, nb, K, 0.01, 0);
    std::vector<long double> want((size_t)K * nb, 0.0L);
    int atom = V.nloc_projectors[0].atom_index - 1;
    for (const auto& pr : V.nloc_projectors) {
      const bool scale = atom != pr.atom_index;
      atom = pr.atom_index;
      for (unsigned b = 0; b < nb; ++b) {
        std::vector<long double> x(pr.ncol, 0.0L);
        for (unsigned c = 0; c < pr.ncol; ++c) {
          for (unsigned r = 0; r < pr.nrow; ++r)
            x[c] += (long double)pr.chi.data[r + (size_t)c * pr.nrow] *
                    psi[pr.index[r] + (size_t)b * K];
          x[c] *= 0.01L;
          if (scale) x[c] *= pr.gamma.data[c];
        }
        for (unsigned r = 0; r < pr.nrow; ++r)
          for (unsigned c = 0; c < pr.ncol; ++c)
            want[pr.index[r] + (size_t)b * K] +=
                (long double)pr.chi.data[r + (size_t)c * pr.nrow] * x[c];
      }
    }
    double worst = 0.0;
    for (unsigned b = 0; b < nb; ++b)
      for (unsigned g = 0; g < K; ++g)
        worst = std::fmax(worst, std::fabs(res32[((size_t)(b / 32) * K + g) * 32 + b % 32] -
                                           (double)want[g + (size_t)b * K]));
    expect(worst < 1e-12, "Hamiltonian::nloc_project_vectors_packed", worst);
  }
  // Stencil_method::Packed: interior copy plus the 12th-order stencil of
  // stencil_extended_test.cpp:151-153 against a naive loop on the extended
  // grid (ghost points set here).
  {
    const unsigned ni = 12, nj = 10, nk = 8, nb = 48;
    const int F = 6;
    const unsigned nie = ni + 2 * F, nje = nj + 2 * F, nke = nk + 2 * F;
    const size_t K = (size_t)ni * nj * nk, Kex = (size_t)nie * nje * nke;
    std::vector<double> local(3 * K * 16), ext(3 * Kex * 16), A(K),
        out(2 * K * 32, 0.0);
    for (size_t e = 0; e < local.size(); ++e) local[e] = std::sin(0.011 * (double)e);
    for (size_t e = 0; e < ext.size(); ++e) ext[e] = std::cos(0.007 * (double)e);
    for (size_t g = 0; g < K; ++g) A[g] = 0.1 * (double)(g % 7);
    Stencil_method::Packed::copy_interior_to_extended_packed(local.data(), ext.data(),
                                                             ni, nj, nk, nb, F);
    double bad = 0.0;
    for (unsigned t = 0; t < 3; ++t)
      for (unsigned k = 0; k < nke; ++k)
        for (unsigned j = 0; j < nje; ++j)
          for (unsigned i = 0; i < nie; ++i) {
            const size_t ge = i + (size_t)j * nie + (size_t)k * nie * nje;
            const bool in = i >= (unsigned)F && i < ni + F && j >= (unsigned)F &&
                            j < nj + F && k >= (unsigned)F && k < nk + F;
            for (unsigned bl = 0; bl < 16; ++bl) {
              const double got = ext[(t * Kex + ge) * 16 + bl];
              const double want =
                  in ? local[(t * K + (i - F) + (size_t)(j - F) * ni +
                              (size_t)(k - F) * ni * nj) * 16 + bl]
                     : std::cos(0.007 * (double)((t * Kex + ge) * 16 + bl));
              bad = std::fmax(bad, std::fabs(got - want));
            }
          }
    expect(bad == 0.0, "Stencil_method::Packed::copy_interior_to_extended_packed", bad);
    const double c[7] = {0, -1.0 / 12, 2.0 / 3, -2.5, 8.0 / 3, -1.0 / 4, 1.0 / 30};
    Stencil_method::Packed::calc_laplacian_extended_packed_16to32(
        ext.data(), out.data(), A.data(), ni, nj, nk, nb, F, c, c, c, 0.5);
    double worst = 0.0;
    for (unsigned b = 0; b < nb; ++b)
      for (unsigned k = 0; k < nk; ++k)
        for (unsigned j = 0; j < nj; ++j)
          for (unsigned i = 0; i < ni; ++i) {
            const size_t g = i + (size_t)j * ni + (size_t)k * ni * nj;
            const size_t ge = (i + F) + (size_t)(j + F) * nie + (size_t)(k + F) * nie * nje;
            auto s = [&](size_t x) { return ext[((size_t)(b / 16) * Kex + x) * 16 + b % 16]; };
            long double r = (long double)s(ge) * (A[g] + 0.5);
            for (int q = 1; q <= F; ++q)
              r += c[q] * ((long double)s(ge + q) + s(ge - q) + s(ge + q * nie) +
                           s(ge - q * nie) + s(ge + (size_t)q * nie * nje) +
                           s(ge - (size_t)q * nie * nje));
            worst = std::fmax(worst, std::fabs(out[((size_t)(b / 32) * K + g) * 32 + b % 32] -
                                               (double)r));
          }
    expect(worst < 1e-12, "Stencil_method::Packed::calc_laplacian_extended_packed_16to32",
           worst);
  }
  std::printf("%s\n", failures ? "FAILED" : "all facade checks passed");
  return failures ? 1 : 0;
}
