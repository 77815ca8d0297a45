"""Pins the oracle (CPU restatement of the reference path) against the
reference's own known-answer inputs, closed forms, exact arithmetic and
invariants.  CPU only.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a transposed operand, a wrong index, a wrong alpha/beta placement)
fails it; test_oracle_mutations plants such mistakes and shows they are
caught.
"""
import os

import numpy as np
import pytest

from conftest import rel_fro

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GEMM_CASES = [(32, 64, 128, 1.0, 0.0), (32, 64, 4096, 1.0, 0.0),
              (32, 64, 4096, 2.0, 0.5), (64, 128, 2048, 1.0, 0.0)]
SYRK_CASES = [(64, 32), (256, 64), (512, 96), (1024, 128), (2048, 256)]


def unpack_p16(ap, K, M):
    """P16 [ceil(M/16)][K][16] -> dense K x M."""
    t = (M + 15) // 16
    return ap.reshape(t, K, 16).transpose(1, 0, 2).reshape(K, t * 16)[:, :M]


# ---- GEMM (gemm_crsp.cpp / gemm_opt.cpp) ----------------------------------

@pytest.mark.parametrize("case", range(4))
def test_gemm_kats_vs_golden_closed_form(orc, case):
    """gemm_crsp.cpp:555-565 cases against the exact closed form."""
    g = np.load(os.path.join(GOLDEN, "gemm_kat.npz"))
    M, N, K = (int(x) for x in g["case%d_shape" % case])
    alpha, gamma = (float(x) for x in g["case%d_coef" % case])
    D_exact = g["case%d_D" % case].ravel()
    A, B = orc.gemm_test_inputs(M, N, K)
    pa, pb = orc.pack_A(M, K, A), orc.pack_B(N, K, B)
    for nth in (1, 3, 8):
        D = orc.gemm_crsp(M, N, K, pa, pb, alpha, gamma, nthreads=nth)
        assert rel_fro(D, D_exact) < 1e-14
    D = orc.gemm_opt(M, N, K, pa, pb, alpha, gamma, nthreads=4)
    assert rel_fro(D, D_exact) < 1e-14
    D = orc.gemm_naive(M, N, K, A, B, alpha, gamma)
    assert rel_fro(D, D_exact) < 1e-14
    # elementwise, as the reference checks (gemm_crsp.cpp:487-497)
    assert np.max(np.abs(D - D_exact) / (np.abs(D_exact) + 1e-10)) < 1e-12


def test_gemm_random_all_layouts_vs_numpy(orc, rng):
    """General-layout checker vs numpy, all A/B/D layouts, ragged shape."""
    M, N, K = 37, 45, 29
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    ref = A @ B
    for la in range(4):
        Ab, lda = encode_a(A, la)
        for lb in range(4):
            Bb, ldb = encode_b(B, lb)
            for ld_ in range(4):
                D, ldd = encode_d(np.zeros((M, N)), ld_)
                orc.dgemm(la, lb, ld_, M, N, K, 1.0, Ab, lda, Bb, ldb, 0.0, D,
                          ldd, 0.0)
                assert rel_fro(decode_d(D, ld_, M, N, ldd), ref) < 1e-15


def test_dd_exact_on_integers(orc, rng):
    """Double-double accumulation is exact where the true result fits."""
    M, N, K = 9, 11, 1000
    A = rng.integers(-2**20, 2**20, (M, K)).astype(np.float64)
    B = rng.integers(-2**20, 2**20, (K, N)).astype(np.float64)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    D = np.zeros(M * N)
    orc.dgemm(0, 0, 0, M, N, K, 1.0, np.ascontiguousarray(A.ravel()), K,
              np.ascontiguousarray(B.ravel()), N, 0.0, D, N, 0.0, dd=True)
    assert np.array_equal(D.reshape(M, N), exact)


def test_dd_recovers_cancellation(orc):
    """[1e16, 1, -1e16] . [1, 1, 1] = 1: lost by plain FP64, kept by dd."""
    A = np.array([1e16, 1.0, -1e16])
    B = np.ones(3)
    D = np.zeros(1)
    orc.dgemm(0, 0, 0, 1, 1, 3, 1.0, A, 3, B, 1, 0.0, D, 1, 0.0, dd=True)
    assert D[0] == 1.0
    D0 = np.zeros(1)
    orc.dgemm(0, 0, 0, 1, 1, 3, 1.0, A, 3, B, 1, 0.0, D0, 1, 0.0, dd=False)
    assert D0[0] == 0.0


def test_gemm_beta_gamma_placement(orc, rng):
    M, N, K = 8, 6, 5
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    C0 = rng.uniform(-1, 1, (M, N))
    D = C0.copy().ravel()
    orc.dgemm(0, 0, 0, M, N, K, 0.5, A.ravel().copy(), K, B.ravel().copy(), N,
              -2.0, D, N, 0.25)
    assert rel_fro(D, 0.5 * A @ B - 2.0 * C0 + 0.25) < 1e-15


def closed_form(M, N, K):
    """A[i,k] = 0.01(i+1) + 0.001(k+1), B[k,j] = 0.001(k+1) + 0.01(j+1)
    (gemm_crsp.cpp:451-460): (AB)_ij = 1e-5 (I+J) S1 + 1e-4 I J K + 1e-6 S2
    with I = i+1, J = j+1, S1 = sum k, S2 = sum k^2 over k = 1..K."""
    I = np.arange(1, M + 1, dtype=np.float64)[:, None]
    J = np.arange(1, N + 1, dtype=np.float64)[None, :]
    S1 = K * (K + 1) / 2.0
    S2 = K * (K + 1) * (2 * K + 1) / 6.0
    return 1e-5 * (I + J) * S1 + 1e-4 * I * J * K + 1e-6 * S2


def test_crsp_partial_panels_gamma(orc):
    """Quirk Q1 in gemm_crsp: the product covers 16*floor(M/16) rows and
    32*floor(N/32) columns, but the reduction (gemm_crsp.cpp:285-310) loops
    over every i < M, j < N with zero partials outside the panels (calloc'd
    workspace, :326), so those entries get alpha * 0 + gamma.  Inside the
    panels the closed form holds."""
    M, N, K = 40, 70, 16
    alpha, gamma = 2.0, 0.375
    A, B = orc.gemm_test_inputs(M, N, K)
    D = np.full(M * N, -7.0)
    orc.gemm_crsp(M, N, K, orc.pack_A(M, K, A), orc.pack_B(N, K, B), alpha,
                  gamma, nthreads=2, D=D)
    D = D.reshape(M, N)
    assert np.all(D[32:, :] == gamma) and np.all(D[:32, 64:] == gamma)
    want = alpha * closed_form(32, 64, K) + gamma
    assert np.max(np.abs(D[:32, :64] - want) / np.abs(want)) < 1e-14


def test_opt_partial_panels_untouched(orc):
    """gemm_opt writes D in place per tile (gemm_opt.cpp:275-287): entries
    outside its 16*floor(M/16) x 32*floor(N/32) tiles are never written."""
    M, N, K = 40, 70, 16
    A, B = orc.gemm_test_inputs(M, N, K)
    D = np.full(M * N, -7.0)
    orc.gemm_opt(M, N, K, orc.pack_A(M, K, A), orc.pack_B(N, K, B), 1.0, 0.0,
                 nthreads=2, D=D)
    D = D.reshape(M, N)
    assert np.all(D[32:, :] == -7.0) and np.all(D[:32, 64:] == -7.0)
    assert np.all(D[:32, :64] != -7.0)


# ---- SYRK (syrk_crsp.cpp / include/sme_syrk.hpp) ----------------------------

def test_syrk_glibc_fixture(orc):
    """The first reference SYRK case reproduces the committed glibc inputs
    and the exact integer A^T A."""
    g = np.load(os.path.join(GOLDEN, "syrk_glibc.npz"))
    K, M, seed = (int(x) for x in g["shape"])
    ap = orc.init_A_packed(K, M, seed)
    assert np.array_equal(np.rint(ap * 100).astype(np.uint8), g["a_packed_int"])
    exact = g["ata_int"] / 1e4
    C = orc.syrk_reference(K, M, ap, 1.0, 0.0)
    C = C.reshape(M, M).T  # column-major -> [m1, m2]
    iu = np.triu_indices(M)
    assert np.max(np.abs(C[iu] - exact[iu])) < 1e-12 * np.max(exact)


@pytest.mark.parametrize("t", range(5))
def test_syrk_reference_cases(orc, t):
    """syrk_crsp.cpp:474-493: five cases, seeds 42..46, exact integer check
    of syrk_reference and of the tile-pair restatement."""
    K, M = SYRK_CASES[t]
    ap = orc.init_A_packed(K, M, 42 + t)
    ints = np.rint(unpack_p16(ap, K, M) * 100).astype(np.int64)
    exact = (ints.T @ ints) / 1e4
    iu = np.triu_indices(M)
    blk = diag_block_lower(M)
    for tiles, C in ((False, orc.syrk_reference(K, M, ap, 1.0, 0.0)),
                     (True, orc.syrk_tiles(K, M, ap, 1.0, 0.0, nthreads=4))):
        Cm = C.reshape(M, M).T
        err = np.abs(Cm[iu] - exact[iu]) / np.maximum(1.0, np.abs(exact[iu]))
        assert err.max() < 1e-13  # reference tolerance is 1e-9
        if tiles:
            # the tile store writes whole diagonal 16 x 16 blocks
            # (sme_syrk.hpp:57-109): m1 > m2 there holds (A^T A)[m1, m2]
            e2 = np.abs(Cm[blk] - exact[blk]) / np.maximum(1.0, np.abs(exact[blk]))
            assert e2.max() < 1e-13
            assert np.all(Cm[~blk & ~np.triu(np.ones((M, M), bool))] == 0.0)
        else:
            assert np.all(np.tril(Cm, -1) == 0.0)  # syrk_reference: upper only


def diag_block_lower(M):
    """Mask [m1, m2] of the entries m1 > m2 inside the 16 x 16 diagonal
    blocks (m1 // 16 == m2 // 16)."""
    i = np.arange(M)
    return (i[:, None] > i[None, :]) & (i[:, None] // 16 == i[None, :] // 16)


def test_syrk_alpha_beta_and_ragged(orc, rng):
    """alpha / beta on the whole product, M % 16 != 0 (the reference's store
    overruns the matrix here, quirk Q2; the restatement clips)."""
    K, M = 300, 37
    A = rng.uniform(-1, 1, (K, M))
    t = (M + 15) // 16
    ap = np.zeros((t, K, 16))
    for m in range(M):
        ap[m // 16, :, m % 16] = A[:, m]
    C0 = rng.uniform(-1, 1, (M, M))           # C0[m1, m2]
    C = C0.ravel(order="F").copy()            # column-major storage
    C = orc.syrk_tiles(K, M, ap.ravel(), 0.75, -1.5, C=C, nthreads=3)
    Cm = C.reshape(M, M).T                    # back to [m1, m2]
    iu = np.triu_indices(M)
    ref = 0.75 * (A.T @ A) - 1.5 * C0
    assert np.max(np.abs(Cm[iu] - ref[iu])) < 1e-13
    blk = diag_block_lower(M)                 # m1 > m2 in a diagonal block
    assert np.max(np.abs(Cm[blk] - ref[blk])) < 1e-13
    rest = ~blk & ~np.triu(np.ones((M, M), bool))
    assert np.array_equal(Cm[rest], C0[rest])  # the rest of the lower part


def test_syrk_orthonormal_columns_give_identity(orc, rng):
    K, M = 4096, 48
    Q, _ = np.linalg.qr(rng.standard_normal((K, M)))
    ap = np.zeros((M // 16, K, 16))
    for m in range(M):
        ap[m // 16, :, m % 16] = Q[:, m]
    C = orc.syrk_tiles(K, M, ap.ravel(), 1.0, 0.0).reshape(M, M).T
    iu = np.triu_indices(M)
    assert np.max(np.abs(C[iu] - np.eye(M)[iu])) < 1e-13


def test_syrk_power_of_two_scaling_is_exact(orc, rng):
    """Ms(A diag(2^e)) = diag(2^e) Ms(A) diag(2^e), bit for bit."""
    K, M = 512, 32
    A = rng.uniform(-1, 1, (K, M))
    e = rng.integers(-20, 20, M)
    As = A * np.ldexp(1.0, e)[None, :]

    def packed(X):
        ap = np.zeros((M // 16, K, 16))
        for m in range(M):
            ap[m // 16, :, m % 16] = X[:, m]
        return ap.ravel()

    C = orc.syrk_tiles(K, M, packed(A), 1.0, 0.0, nthreads=1).reshape(M, M).T
    Cs = orc.syrk_tiles(K, M, packed(As), 1.0, 0.0, nthreads=1).reshape(M, M).T
    scale = np.ldexp(1.0, e[:, None] + e[None, :])
    iu = np.triu_indices(M)
    assert np.array_equal(Cs[iu], (C * scale)[iu])


def test_dsyrk_general_vs_numpy(orc, rng):
    N, K = 41, 77
    A = rng.uniform(-1, 1, (K, N))
    ref = A.T @ A
    il = np.tril_indices(N)
    for la in range(4):
        Ab, lda = encode_b(A, la)
        C = np.full(N * N, 3.0)
        orc.dsyrk(la, N, K, 1.0, Ab, lda, 0.0, C, N)
        Cm = C.reshape(N, N)
        assert rel_fro(Cm[il], ref[il]) < 1e-15
        assert np.all(Cm[np.triu_indices(N, 1)] == 3.0)


# ---- rotation (include/sme_trp_3d.hpp, trp_3d_test.cpp) -------------------

@pytest.mark.parametrize("Nd,nb", [(64, 32), (256, 64), (512, 128)])
def test_trp_closed_form(orc, Nd, nb):
    """trp_3d_test.cpp:125-138 inputs; D[g,bp] = 1e-5 (G+P) S1 +
    1e-6 G P nb + 1e-4 S2 with G = g+1, P = bp+1."""
    A, B = orc.trp_test_inputs(Nd, nb)
    S1 = nb * (nb + 1) / 2
    S2 = nb * (nb + 1) * (2 * nb + 1) / 6
    G = np.arange(1, Nd + 1)[:, None]
    P = np.arange(1, nb + 1)[None, :]
    Dx = 1e-5 * (G + P) * S1 + 1e-6 * G * P * nb + 1e-4 * S2
    for D in (orc.trp_reference(Nd, nb, A, B), orc.trp_3d(Nd, nb, A, B, 3)):
        Dd = D.reshape(nb // 16, Nd, 16).transpose(1, 0, 2).reshape(Nd, nb)
        assert rel_fro(Dd, Dx) < 1e-14


def test_trp_identity_and_permutation(orc, rng):
    Nd, nb = 128, 48
    X = rng.uniform(-1, 1, (Nd, nb))
    Ag = X.reshape(Nd // 16, 16, nb).transpose(0, 2, 1).ravel().copy()
    perm = rng.permutation(nb)
    Q = np.eye(nb)[:, perm].ravel().copy()  # D = X[:, perm]
    for Qm, expect in ((np.eye(nb).ravel().copy(), X), (Q, X[:, perm])):
        D = orc.trp_3d(Nd, nb, Ag, Qm, 2)
        Dd = D.reshape(nb // 16, Nd, 16).transpose(1, 0, 2).reshape(Nd, nb)
        assert np.array_equal(Dd, expect)


# ---- packing (pack.cpp / include/sme_pack.hpp) ------------------------------

def test_pack_roundtrip_and_layout(orc, rng):
    Nd, nb = 96, 40
    src = rng.uniform(-1, 1, Nd * nb)  # column-major [Nd, nb]
    p = orc.pack_c16(Nd, nb, src)
    g, b = 37, 11
    assert p[(g // 16) * nb * 16 + b * 16 + g % 16] == src[g + b * Nd]
    assert np.array_equal(orc.unpack_c16(Nd, nb, p), src)
    M, K = 48, 20
    A = rng.uniform(-1, 1, M * K)
    pa = orc.pack_A(M, K, A)
    assert pa[(2 * K + 5) * 16 + 3] == A[(2 * 16 + 3) + 5 * M]
    N = 64
    B = rng.uniform(-1, 1, K * N)
    pb = orc.pack_B(N, K, B)
    assert pb[(1 * K + 7) * 32 + 9] == B[7 * N + 32 + 9]


def band_pack(X):
    """Column-major dense Nd x nb -> band-packed [nb/16][Nd][16] by numpy
    reshapes (element (g, b) at (b/16, g, b%16))."""
    Nd, nb = X.shape
    return X.reshape(Nd, nb // 16, 16).transpose(1, 0, 2).ravel().copy()


def grid_pack(X):
    """Dense Nd x nb -> grid-packed [Nd/16][nb][16] (element (g, b) at
    (g/16, b, g%16))."""
    Nd, nb = X.shape
    return X.reshape(Nd // 16, 16, nb).transpose(0, 2, 1).ravel().copy()


def test_band_grid_convert_known_answer(orc):
    # pack.cpp:190-258: value[g, b] = g + 1000 b, band -> grid -> band.
    Nd, nb = 64, 48
    band = np.zeros((nb // 16) * Nd * 16)
    for g in range(Nd):
        for b in range(nb):
            band[(b // 16) * Nd * 16 + g * 16 + b % 16] = g + b * 1000
    grid = orc.band_to_grid(Nd, nb, band)
    for g in range(Nd):
        for b in range(nb):
            assert grid[(g // 16) * nb * 16 + b * 16 + g % 16] == g + b * 1000
    assert np.array_equal(orc.grid_to_band(Nd, nb, grid), band)


def test_band_grid_convert_matches_reshapes(orc, rng):
    Nd, nb = 96, 80
    X = rng.uniform(-1, 1, (Nd, nb))
    band, grid = band_pack(X), grid_pack(X)
    assert np.array_equal(orc.band_to_grid(Nd, nb, band), grid)
    assert np.array_equal(orc.grid_to_band(Nd, nb, grid), band)
    # grid-packed is pack_c16 of the column-major matrix
    assert np.array_equal(orc.pack_c16(Nd, nb, X.ravel(order="F").copy()), grid)
    # a swapped pair of indices is caught
    Y = X.copy()
    Y[[3, 4]] = Y[[4, 3]]
    assert not np.array_equal(orc.band_to_grid(Nd, nb, band_pack(Y)), grid)


# ---- non-local projectors (include/nloc_packed_opt2.hpp) ---------------------

NLOC_CONFIGS = [(100, 16, 20, 5, 1), (200, 32, 50, 8, 2), (500, 48, 100, 12, 3),
                (1000, 64, 200, 16, 2), (2000, 128, 300, 20, 4)]


def nloc_projectors(rng, K, nrow, ncol, nproj, atoms=None):
    """Random projectors as in nloc_packed_opt2_test.cpp:239-266: chi and
    gamma uniform, nrow distinct grid indices per projector (projectors may
    share grid points)."""
    atoms = list(range(1, nproj + 1)) if atoms is None else atoms
    nrows = [nrow] * nproj
    ncols = [ncol] * nproj
    idx = np.concatenate([rng.permutation(K)[:nrow] for _ in range(nproj)])
    chi = np.concatenate([rng.uniform(-1, 1, nrow * ncol) for _ in range(nproj)])
    gam = np.concatenate([rng.uniform(-1, 1, ncol) * 0.5 for _ in range(nproj)])
    return nrows, ncols, atoms, idx, chi, gam


def nloc_unpacked(K, nb, dv, nrows, ncols, atoms, idx, chi, gam, psi, result,
                  scale_all=False):
    """The reference test's unpacked formulation (nloc_packed_opt2_test.cpp:
    84-165) with numpy products: psi and result column-major K x nb."""
    P = psi.reshape(nb, K).T
    R = result.reshape(nb, K).T.copy()
    coff = np.concatenate([[0], np.cumsum(ncols)])
    roff = np.concatenate([[0], np.cumsum(nrows)])
    choff = np.concatenate([[0], np.cumsum(np.array(nrows) * np.array(ncols))])
    X = np.zeros((coff[-1], nb))
    chis = []
    for p in range(len(nrows)):
        C = chi[choff[p]:choff[p + 1]].reshape(ncols[p], nrows[p]).T
        chis.append(C)
        X[coff[p]:coff[p + 1]] = dv * (C.T @ P[idx[roff[p]:roff[p + 1]]])
    atom = atoms[0] - 1
    for p in range(len(nrows)):
        if scale_all or atom != atoms[p]:
            atom = atoms[p]
            X[coff[p]:coff[p + 1]] *= gam[coff[p]:coff[p + 1], None]
    for p in range(len(nrows)):
        R[idx[roff[p]:roff[p + 1]]] += chis[p] @ X[coff[p]:coff[p + 1]]
    return R.T.ravel().copy()


def band_pack_cm(v, K, nb, T):
    """Column-major K x nb -> T-band-packed [ceil(nb/T)][K][T] (zero pad)."""
    t = (nb + T - 1) // T
    out = np.zeros((t, K, T))
    M = v.reshape(nb, K).T
    for b in range(nb):
        out[b // T, :, b % T] = M[:, b]
    return out.ravel()


def band_unpack_cm(v, K, nb, T):
    t = (nb + T - 1) // T
    A = v.reshape(t, K, T)
    return np.stack([A[b // T, :, b % T] for b in range(nb)]).ravel()


@pytest.mark.parametrize("cfg", NLOC_CONFIGS)
@pytest.mark.parametrize("tile", [32, 16])
def test_nloc_matches_unpacked_formulation(orc, cfg, tile):
    K, nb, nrow, ncol, nproj = cfg
    rng = np.random.default_rng(K + nb)
    nrows, ncols, atoms, idx, chi, gam = nloc_projectors(rng, K, nrow, ncol, nproj)
    psi = rng.uniform(-1, 1, K * nb)
    res0 = rng.uniform(-1, 1, K * nb)
    dv = 0.01
    want = nloc_unpacked(K, nb, dv, nrows, ncols, atoms, idx, chi, gam, psi, res0)
    out = band_pack_cm(res0, K, nb, tile)
    orc.nloc_apply(K, nb, dv, nrows, ncols, atoms, idx, chi, gam,
                   band_pack_cm(psi, K, nb, 16), out, tile)
    got = band_unpack_cm(out, K, nb, tile)
    assert np.max(np.abs(got - want)) < 1e-12


def test_nloc_closed_form_and_gamma_once_per_atom(orc):
    """One column of ones: every touched point gets gamma dv sum(psi[idx]);
    two projectors of one atom: only the first is scaled."""
    K, nb = 64, 16
    psi = np.arange(K * nb, dtype=np.float64) % 7          # small integers
    idx = np.array([3, 9, 17, 40])
    for atoms, scales in (([5, 6], (2.0, 3.0)), ([5, 5], (2.0, 1.0))):
        res = np.zeros(K * 32)                              # 32-packed, nb = 16
        orc.nloc_apply(K, nb, 0.5, [4, 4], [1, 1], atoms,
                       np.concatenate([idx, idx]), np.ones(8), [2.0, 3.0],
                       band_pack_cm(psi, K, nb, 16), res, 32)
        got = band_unpack_cm(res, K, nb, 32).reshape(nb, K).T
        P = psi.reshape(nb, K).T
        s = 0.5 * P[idx].sum(axis=0)                        # dv sum, per band
        want = np.zeros((K, nb))
        want[idx] = (scales[0] + scales[1]) * s
        assert np.array_equal(got, want)


def test_nloc_planted_mistake_is_caught(orc):
    """Scaling every projector (ignoring the once-per-atom rule) differs."""
    K, nb = 300, 32
    rng = np.random.default_rng(9)
    nrows, ncols, atoms, idx, chi, gam = nloc_projectors(rng, K, 40, 6, 3,
                                                         atoms=[1, 1, 2])
    psi = rng.uniform(-1, 1, K * nb)
    res0 = np.zeros(K * nb)
    good = nloc_unpacked(K, nb, 0.1, nrows, ncols, atoms, idx, chi, gam, psi, res0)
    bad = nloc_unpacked(K, nb, 0.1, nrows, ncols, atoms, idx, chi, gam, psi, res0,
                        scale_all=True)
    out = band_pack_cm(res0, K, nb, 32)
    orc.nloc_apply(K, nb, 0.1, nrows, ncols, atoms, idx, chi, gam,
                   band_pack_cm(psi, K, nb, 16), out, 32)
    got = band_unpack_cm(out, K, nb, 32)
    assert np.max(np.abs(got - good)) < 1e-12
    assert np.max(np.abs(got - bad)) > 1e-3


# ---- real-space stencil (include/stencil_extended_packed.hpp, halo_packed.hpp)

# stencil_extended_test.cpp:151-153: 12th-order coefficients used by the test
STENCIL_COEF = np.array([0, -1.0 / 12, 2.0 / 3, -2.5, 8.0 / 3, -1.0 / 4, 1.0 / 30])


def grid_pack16(V, nb):
    """[nb][nk][nj][ni] -> 16-band-packed [ceil(nb/16)][K][16] (zero pad)."""
    t = (nb + 15) // 16
    out = np.zeros((t, V[0].size, 16))
    for b in range(nb):
        out[b // 16, :, b % 16] = V[b].ravel()
    return out.ravel()


def grid_unpack(v, nb, K, T):
    t = (nb + T - 1) // T
    A = v.reshape(t, K, T)
    return np.stack([A[b // T, :, b % T] for b in range(nb)])


def stencil_numpy(Sx, A, F, cx, cy, cz, c0):
    """The reference test's naive stencil (stencil_extended_test.cpp:18-52)
    with numpy slices: Sx [nb][nk+2F][nj+2F][ni+2F] extended, A [nk][nj][ni]."""
    nb = Sx.shape[0]
    nk, nj, ni = A.shape
    C = Sx[:, F:F + nk, F:F + nj, F:F + ni]
    out = C * (A + c0)
    for r in range(1, F + 1):
        out = out + cx[r] * (Sx[:, F:F + nk, F:F + nj, F + r:F + r + ni] +
                             Sx[:, F:F + nk, F:F + nj, F - r:F - r + ni])
    for r in range(1, F + 1):
        out = out + cy[r] * (Sx[:, F:F + nk, F + r:F + r + nj, F:F + ni] +
                             Sx[:, F:F + nk, F - r:F - r + nj, F:F + ni])
    for r in range(1, F + 1):
        out = out + cz[r] * (Sx[:, F + r:F + r + nk, F:F + nj, F:F + ni] +
                             Sx[:, F - r:F - r + nk, F:F + nj, F:F + ni])
    return out


@pytest.mark.parametrize("shape", [(12, 10, 8, 48, 6), (7, 5, 9, 40, 2),
                                   (4, 4, 4, 16, 1)])
@pytest.mark.parametrize("tile", [32, 16])
def test_stencil_matches_unpacked_formulation(orc, rng, shape, tile):
    ni, nj, nk, nb, F = shape
    Sx = rng.uniform(-1, 1, (nb, nk + 2 * F, nj + 2 * F, ni + 2 * F))
    A = rng.uniform(-1, 1, (nk, nj, ni))
    c = STENCIL_COEF if F == 6 else rng.uniform(-1, 1, F + 1)
    ext = grid_pack16(Sx, nb)
    out = orc.stencil_apply(ni, nj, nk, nb, F, ext, A.ravel(), c, c * 0.5,
                            c * 0.25, 1.5, tile)
    want = stencil_numpy(Sx, A, F, c, c * 0.5, c * 0.25, 1.5)
    got = grid_unpack(out, nb, ni * nj * nk, tile)
    assert np.max(np.abs(got - want.reshape(nb, -1))) < 1e-12


def test_stencil_closed_forms(orc):
    """Constant input: (A + c0 + 2 sum(cx + cy + cz)) psi; input linear in i:
    the symmetric x terms give 2 i sum(cx); all exact in binary fractions."""
    ni, nj, nk, nb, F = 6, 5, 4, 16, 3
    c = np.array([0, 0.5, -0.25, 0.125])
    A = np.full((nk, nj, ni), 2.0)
    Sx = np.full((nb, nk + 2 * F, nj + 2 * F, ni + 2 * F), 3.0)
    out = grid_unpack(orc.stencil_apply(ni, nj, nk, nb, F, grid_pack16(Sx, nb),
                                        A.ravel(), c, c, c, 1.0, 16),
                      nb, ni * nj * nk, 16)
    assert np.all(out == 3.0 * (2.0 + 1.0 + 3 * 2 * c.sum()))
    ie = np.arange(ni + 2 * F, dtype=np.float64)
    Sx = np.broadcast_to(ie, (nb, nk + 2 * F, nj + 2 * F, ni + 2 * F)).copy()
    out = grid_unpack(orc.stencil_apply(ni, nj, nk, nb, F, grid_pack16(Sx, nb),
                                        A.ravel(), c, 0 * c, 0 * c, 0.0, 32),
                      nb, ni * nj * nk, 32).reshape(nb, nk, nj, ni)
    i_ex = np.arange(ni) + F
    assert np.all(out == (i_ex * 2.0 + 2 * i_ex * c.sum())[None, None, None, :])


@pytest.mark.parametrize("periodic", [True, False])
def test_halo_extend(orc, rng, periodic):
    """Interior copy (halo_packed.hpp:86-123) and the single-domain halo:
    np.pad(mode='wrap') for a periodic domain, zeros otherwise; the
    reference's spot checks (halo_packed_test.cpp:177-196)."""
    ni, nj, nk, nb, F = 6, 5, 4, 20, 2
    L = rng.uniform(-1, 1, (nb, nk, nj, ni))
    ext = orc.halo_extend(ni, nj, nk, nb, F, grid_pack16(L, nb), periodic)
    mode = "wrap" if periodic else "constant"
    want = np.pad(L, ((0, 0), (F, F), (F, F), (F, F)), mode=mode)
    K_ex = (ni + 2 * F) * (nj + 2 * F) * (nk + 2 * F)
    assert np.array_equal(grid_unpack(ext, nb, K_ex, 16), want.reshape(nb, -1))
    if periodic:
        nie, nje = ni + 2 * F, nj + 2 * F
        e = ext.reshape(-1, K_ex, 16)
        assert e[0, F * nie + F * nie * nje, 0] == L[0, 0, 0, ni - F]
        assert e[0, (ni + F) + F * nie + F * nie * nje, 0] == L[0, 0, 0, 0]


def test_stencil_planted_mistake_is_caught(orc, rng):
    """Dropping the outermost x term (r = F) is seen."""
    ni, nj, nk, nb, F = 8, 6, 5, 16, 3
    Sx = rng.uniform(-1, 1, (nb, nk + 2 * F, nj + 2 * F, ni + 2 * F))
    A = rng.uniform(-1, 1, (nk, nj, ni))
    c = rng.uniform(-1, 1, F + 1)
    got = grid_unpack(orc.stencil_apply(ni, nj, nk, nb, F, grid_pack16(Sx, nb),
                                        A.ravel(), c, c, c, 0.0, 16),
                      nb, ni * nj * nk, 16)
    bad_c = c.copy()
    bad_c[F] = 0.0
    bad = stencil_numpy(Sx, A, F, bad_c, c, c, 0.0).reshape(nb, -1)
    assert np.max(np.abs(got - bad)) > 1e-3


# ---- planted mistakes are caught -----------------------------------------------

def test_oracle_mutations_are_detected(orc, rng):
    """Transposing B, dropping the last k, or flipping gamma's sign in the
    checker's output must fail the checks above."""
    g = np.load(os.path.join(GOLDEN, "gemm_kat.npz"))
    M, N, K = (int(x) for x in g["case2_shape"])
    alpha, gamma = (float(x) for x in g["case2_coef"])
    D_exact = g["case2_D"].ravel()
    A, B = orc.gemm_test_inputs(M, N, K)
    good = orc.gemm_naive(M, N, K, A, B, alpha, gamma)
    assert rel_fro(good, D_exact) < 1e-14
    dropped = orc.gemm_naive(M, N, K - 1, A, B[:(K - 1) * N].copy(), alpha, gamma)
    assert rel_fro(dropped, D_exact) > 1e-6
    flipped = orc.gemm_naive(M, N, K, A, B, alpha, -gamma)
    assert rel_fro(flipped, D_exact) > 1e-6
    # transposed operand on random data (the KAT closed form is symmetric)
    X = rng.uniform(-1, 1, (16, 16))
    Y = rng.uniform(-1, 1, (16, 16))
    D = np.zeros(256)
    orc.dgemm(0, 1, 0, 16, 16, 16, 1.0, X.ravel().copy(), 16, Y.ravel().copy(),
              16, 0.0, D, 16, 0.0)  # B read column-major = Y^T
    assert rel_fro(D, X @ Y) > 1e-3


# ---- layout helpers (shared with the GPU tests) ----------------------------------

def encode_a(A, layout, pad=3):
    """Dense M x K -> flat buffer in `layout` (A convention) and its ld."""
    M, K = A.shape
    if layout == 0:
        ld = K + pad
        buf = np.zeros((M, ld)); buf[:, :K] = A
        return buf.ravel(), ld
    if layout == 1:
        ld = M + pad
        buf = np.zeros((K, ld)); buf[:, :M] = A.T
        return buf.ravel(), ld
    P = 16 if layout == 2 else 32
    t = (M + P - 1) // P
    ld = K + pad
    buf = np.zeros((t, ld, P))
    for m in range(M):
        buf[m // P, :K, m % P] = A[m]
    return buf.ravel(), ld


def encode_b(B, layout, pad=3):
    """Dense K x N -> flat buffer in `layout` (B convention) and its ld."""
    K, N = B.shape
    if layout == 0:
        ld = N + pad
        buf = np.zeros((K, ld)); buf[:, :N] = B
        return buf.ravel(), ld
    if layout == 1:
        ld = K + pad
        buf = np.zeros((N, ld)); buf[:, :K] = B.T
        return buf.ravel(), ld
    P = 16 if layout == 2 else 32
    t = (N + P - 1) // P
    ld = K + pad
    buf = np.zeros((t, ld, P))
    for n in range(N):
        buf[n // P, :K, n % P] = B[:, n]
    return buf.ravel(), ld


def encode_d(D, layout, pad=2):
    M, N = D.shape
    if layout == 0:
        ld = N + pad
        buf = np.zeros((M, ld)); buf[:, :N] = D
        return buf.ravel(), ld
    if layout == 1:
        ld = M + pad
        buf = np.zeros((N, ld)); buf[:, :M] = D.T
        return buf.ravel(), ld
    P = 16 if layout == 2 else 32
    t = (N + P - 1) // P
    ld = M + pad
    buf = np.zeros((t, ld, P))
    for n in range(N):
        buf[n // P, :M, n % P] = D[:, n]
    return buf.ravel(), ld


def decode_d(buf, layout, M, N, ld):
    if layout == 0:
        return buf.reshape(M, ld)[:, :N]
    if layout == 1:
        return buf.reshape(N, ld)[:, :M].T
    P = 16 if layout == 2 else 32
    t = (N + P - 1) // P
    b = buf.reshape(t, ld, P)
    out = np.empty((M, N))
    for n in range(N):
        out[:, n] = b[n // P, :M, n % P]
    return out
