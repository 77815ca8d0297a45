"""GPU parity of the FP64 Stream-K GEMM against the oracle, through the C ABI.

Bar: relative Frobenius error <= 1e-13 against the double-double oracle
(plain FP64 contract); exact where the math makes it exact.
"""
import os

import numpy as np
import pytest

from conftest import rel_fro
from test_oracle import decode_d, encode_a, encode_b, encode_d

pytestmark = pytest.mark.gpu
TOL = 1e-13
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dev(skg, ctx, a):
    return skg.DeviceArray.from_host(ctx, a)


@pytest.mark.parametrize("case", range(4))
def test_reference_kats(skg, ctx, orc, case):
    """gemm_crsp.cpp:555-565 through the compute_test layouts (A column-major,
    B row-major, D row-major, D = alpha*A*B + gamma) vs the exact closed form."""
    g = np.load(os.path.join(GOLDEN, "gemm_kat.npz"))
    M, N, K = (int(x) for x in g["case%d_shape" % case])
    alpha, gamma = (float(x) for x in g["case%d_coef" % case])
    A, B = orc.gemm_test_inputs(M, N, K)
    D = np.zeros(M * N)
    print("case", case, M, N, K, flush=True)
    ctx.set_tuning(-1, 0)
    ctx.dgemm_host(skg.COL_MAJOR, skg.ROW_MAJOR, skg.ROW_MAJOR, M, N, K, alpha,
                   A, M, B, N, 0.0, D, N, gamma)
    exact = g["case%d_D" % case].ravel()
    assert rel_fro(D, exact) < TOL
    assert np.max(np.abs(D - exact) / (np.abs(exact) + 1e-10)) < 1e-12


@pytest.mark.parametrize("case", range(4))
def test_packed_entry_matches_crsp_restatement(skg, ctx, orc, case):
    """compute_streamk_packed: PA [M/16][K][16], PB [N/32][K][32]."""
    g = np.load(os.path.join(GOLDEN, "gemm_kat.npz"))
    M, N, K = (int(x) for x in g["case%d_shape" % case])
    alpha, gamma = (float(x) for x in g["case%d_coef" % case])
    A, B = orc.gemm_test_inputs(M, N, K)
    pa, pb = orc.pack_A(M, K, A), orc.pack_B(N, K, B)
    ref = orc.gemm_crsp(M, N, K, pa, pb, alpha, gamma, nthreads=4)
    D = np.zeros(M * N)
    ctx.dgemm_host(skg.PACK16, skg.PACK32, skg.ROW_MAJOR, M, N, K, alpha, pa, K,
                   pb, K, 0.0, D, N, gamma)
    assert rel_fro(D, ref) < TOL


def test_config1_gemm_1024(skg, ctx, orc):
    """c1: M=N=K=1024, row-major A and B uniform(-1,1) seeds 42 / 43."""
    M = N = K = 1024
    A = orc.fill_uniform(M * K, 42)
    B = orc.fill_uniform(K * N, 43)
    ref = np.zeros(M * N)
    orc.dgemm(0, 0, 0, M, N, K, 1.0, A, K, B, N, 0.0, ref, N, 0.0, dd=True)
    for tile in [-1] + list(range(skg.num_tiles())):
        ctx.set_tuning(tile, 0)
        D = np.zeros(M * N)
        ctx.dgemm_host(skg.ROW_MAJOR, skg.ROW_MAJOR, skg.ROW_MAJOR, M, N, K, 1.0,
                       A, K, B, N, 0.0, D, N, 0.0)
        err = rel_fro(D, ref)
        print("tile", tile, "err", err, flush=True)
        assert err < TOL
    ctx.set_tuning(-1, 0)


def test_all_layouts_ragged(skg, ctx, orc, rng):
    """Every A/B/D layout (row, col, PACK16, PACK32) with padded lds on a
    shape that is a multiple of no tile."""
    M, N, K = 77, 45, 133
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    ref = A @ B
    ctx.set_tuning(-1, 0)
    for la in range(4):
        Ab, lda = encode_a(A, la)
        for lb in range(4):
            Bb, ldb = encode_b(B, lb)
            for ld_ in range(4):
                Db, ldd = encode_d(np.zeros((M, N)), ld_)
                ctx.dgemm_host(la, lb, ld_, M, N, K, 1.0, Ab, lda, Bb, ldb, 0.0,
                               Db, ldd, 0.0)
                got = decode_d(Db, ld_, M, N, ldd)
                chk = np.zeros(M * N)
                assert rel_fro(got, ref) < TOL, (la, lb, ld_)
                del chk


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (1, 300, 7), (300, 1, 7),
                                   (17, 33, 1), (65, 129, 2), (63, 64, 17),
                                   (16, 16, 4096), (200, 3, 999), (513, 257, 65)])
def test_edge_shapes(skg, ctx, orc, rng, M, N, K):
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    ref = np.zeros(M * N)
    orc.dgemm(0, 0, 0, M, N, K, 1.0, np.ascontiguousarray(A.ravel()), K,
              np.ascontiguousarray(B.ravel()), N, 0.0, ref, N, 0.0)
    D = np.zeros(M * N)
    ctx.dgemm_host(0, 0, 0, M, N, K, 1.0, np.ascontiguousarray(A.ravel()), K,
                   np.ascontiguousarray(B.ravel()), N, 0.0, D, N, 0.0)
    assert rel_fro(D, ref) < TOL


def test_zero_sizes_and_k0(skg, ctx, rng):
    D = np.full(12, 5.0)
    ctx.dgemm_host(0, 0, 0, 0, 4, 3, 1.0, np.zeros(12), 3, np.zeros(12), 4, 0.0,
                   D, 4, 0.0)
    assert np.all(D == 5.0)  # M == 0: nothing written
    C0 = rng.uniform(-1, 1, 12)
    D = C0.copy()
    ctx.dgemm_host(0, 0, 0, 3, 4, 0, 2.0, np.zeros(1), 1, np.zeros(1), 4, 0.5,
                   D, 4, 0.25)
    assert np.allclose(D, 0.5 * C0 + 0.25, rtol=0, atol=1e-15)  # K == 0


def test_alpha_beta_gamma(skg, ctx, orc, rng):
    M, N, K = 150, 90, 300
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    C0 = rng.uniform(-1, 1, (M, N))
    for kchunks in (1, 4):
        ctx.set_kchunks(kchunks)
        for alpha, beta, gamma in [(2.0, 0.0, 0.5), (-1.0, 1.0, 0.0),
                                   (0.5, -2.0, 0.25), (0.0, 3.0, 1.0)]:
            D = C0.copy().ravel()
            ctx.dgemm_host(0, 0, 0, M, N, K, alpha,
                           np.ascontiguousarray(A.ravel()), K,
                           np.ascontiguousarray(B.ravel()), N, beta, D, N,
                           gamma)
            ref = alpha * (A @ B) + beta * C0 + gamma
            assert rel_fro(D, ref) < TOL
    ctx.set_kchunks(0)


def test_beta_zero_never_reads_d(skg, ctx, rng):
    M, N, K = 64, 64, 64
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    D = np.full(M * N, np.nan)
    ctx.dgemm_host(0, 0, 0, M, N, K, 1.0, np.ascontiguousarray(A.ravel()), K,
                   np.ascontiguousarray(B.ravel()), N, 0.0, D, N, 0.0)
    assert np.all(np.isfinite(D)) and rel_fro(D, A @ B) < TOL


def test_streamk_partition_invariance(skg, ctx, orc, rng):
    """Every tile configuration and grid size (data-parallel, split tiles,
    many peers per tile) agrees with the oracle."""
    M, N, K = 300, 200, 1500
    A = rng.uniform(-1, 1, (M, K)).ravel()
    B = rng.uniform(-1, 1, (K, N)).ravel()
    ref = np.zeros(M * N)
    orc.dgemm(0, 0, 0, M, N, K, 1.0, A, K, B, N, 0.0, ref, N, 0.0, dd=True)
    results = []
    for tile in range(skg.num_tiles()):
        for grid, kchunks in ((1, 1), (2, 1), (3, 2), (7, 1), (64, 3), (0, 1),
                              (0, 8), (16, 8)):
            ctx.set_tuning(tile, grid)
            ctx.set_kchunks(kchunks)
            D = np.zeros(M * N)
            ctx.dgemm_host(0, 0, 0, M, N, K, 1.0, A, K, B, N, 0.0, D, N, 0.0)
            err = rel_fro(D, ref)
            assert err < TOL, (tile, grid, kchunks, err)
            results.append(D)
    ctx.set_tuning(-1, 0)
    ctx.set_kchunks(0)
    for D in results[1:]:
        assert rel_fro(D, results[0]) < 1e-14


def test_repeated_calls_reuse_workspace(skg, ctx, rng):
    """Fix-up flags are consumed and cleared by every launch: back-to-back
    split-tile launches of different shapes stay correct."""
    for it in range(6):
        M, N, K = [(256, 256, 4096), (100, 60, 5000), (512, 128, 777)][it % 3]
        A = rng.uniform(-1, 1, (M, K))
        B = rng.uniform(-1, 1, (K, N))
        ctx.set_tuning(1, 97)
        D = np.zeros(M * N)
        ctx.dgemm_host(0, 0, 0, M, N, K, 1.0, np.ascontiguousarray(A.ravel()),
                       K, np.ascontiguousarray(B.ravel()), N, 0.0, D, N, 0.0)
        assert rel_fro(D, A @ B) < TOL
    ctx.set_tuning(-1, 0)


@pytest.mark.parametrize("Nd,nb", [(64, 32), (256, 64), (512, 128),
                                   (1024, 256), (2048, 128), (4096, 128)])
def test_rotation_trp3d_cases(skg, ctx, orc, Nd, nb):
    """trp_3d_test.cpp:237-238: grid-packed A, row-major Q, band-packed D."""
    A, Q = orc.trp_test_inputs(Nd, nb)
    ref = orc.trp_reference(Nd, nb, A, Q)
    D = np.zeros(((nb + 15) // 16) * Nd * 16)
    ctx.dgemm_host(skg.PACK16, skg.ROW_MAJOR, skg.PACK16, Nd, nb, nb, 1.0, A, nb,
                   Q, nb, 0.0, D, Nd, 0.0)
    assert rel_fro(D, ref) < TOL


def test_config2_rotation_full_size(skg, ctx, orc):
    """c2 rotation Psi (32768 x 520, normal(0, 1/sqrt(Nd))) * Q (520 x 520
    uniform(-1,1)), column-major Psi.  Checked on 256 sampled rows against
    the double-double oracle, and Q = I reproduces Psi bit for bit."""
    Nd, Ns = 32768, 520
    psi = skg.DeviceArray(ctx, Nd * Ns)
    q = skg.DeviceArray(ctx, Ns * Ns)
    out = skg.DeviceArray(ctx, Nd * Ns)
    ctx.fill_normal(psi.ptr, Nd * Ns, 1, 1.0 / np.sqrt(Nd))
    ctx.fill_uniform(q.ptr, Ns * Ns, 2, -1.0, 1.0)
    ctx.set_tuning(-1, 0)
    ctx.dgemm(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, Nd, Ns, Ns, 1.0,
              psi.ptr, Nd, q.ptr, Ns, 0.0, out.ptr, Nd, 0.0)
    ctx.synchronize()
    P = psi.download().reshape(Ns, Nd).T      # [g, b]
    Q = q.download().reshape(Ns, Ns)
    R = out.download().reshape(Ns, Nd).T
    rows = np.random.default_rng(7).choice(Nd, 256, replace=False)
    sub = np.ascontiguousarray(P[rows])
    ref = np.zeros(256 * Ns)
    orc.dgemm(0, 0, 0, 256, Ns, Ns, 1.0, sub.ravel(), Ns, Q.ravel().copy(), Ns,
              0.0, ref, Ns, 0.0, dd=True)
    assert rel_fro(R[rows], ref) < TOL
    eye = skg.DeviceArray.from_host(ctx, np.eye(Ns).ravel())
    ctx.dgemm(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, Nd, Ns, Ns, 1.0,
              psi.ptr, Nd, eye.ptr, Ns, 0.0, out.ptr, Nd, 0.0)
    ctx.synchronize()
    assert np.array_equal(out.download(), psi.download())


def test_device_pack_matches_reference_layouts(skg, ctx, orc, rng):
    """skgemm_pack: linalg_pack_c16_f64 and gemm::pack_A / pack_B layouts."""
    Nd, nb = 96, 40
    src = rng.uniform(-1, 1, Nd * nb)
    d_src = dev(skg, ctx, src)
    d_dst = skg.DeviceArray(ctx, (Nd // 16) * nb * 16, zero=True)
    ctx.pack(skg.ROW_MAJOR, d_src.ptr, Nd, skg.PACK16, d_dst.ptr, nb, Nd, nb)
    ctx.synchronize()
    assert np.array_equal(d_dst.download(), orc.pack_c16(Nd, nb, src))
    back = skg.DeviceArray(ctx, Nd * nb, zero=True)
    ctx.pack(skg.PACK16, d_dst.ptr, nb, skg.ROW_MAJOR, back.ptr, Nd, Nd, nb)
    ctx.synchronize()
    assert np.array_equal(back.download(), src)
    M, K = 48, 20
    A = rng.uniform(-1, 1, M * K)                 # column-major M x K
    dA = dev(skg, ctx, A)
    dpa = skg.DeviceArray(ctx, M * K, zero=True)
    # pack_A: element (r = k, o = m); source m + k*M is ROW_MAJOR (B conv.)
    ctx.pack(skg.ROW_MAJOR, dA.ptr, M, skg.PACK16, dpa.ptr, K, M, K)
    ctx.synchronize()
    assert np.array_equal(dpa.download(), orc.pack_A(M, K, A))
    N = 64
    Bm = rng.uniform(-1, 1, K * N)                # row-major K x N
    dB = dev(skg, ctx, Bm)
    dpb = skg.DeviceArray(ctx, K * N, zero=True)
    ctx.pack(skg.ROW_MAJOR, dB.ptr, N, skg.PACK32, dpb.ptr, K, N, K)
    ctx.synchronize()
    assert np.array_equal(dpb.download(), orc.pack_B(N, K, Bm))


def test_autotune_save_load_and_use(skg, ctx, orc, rng, tmp_path):
    """The autotuner's plan is recorded, persisted, reloaded by a fresh
    context, and the tuned launch is still correct."""
    M, N, K = 700, 130, 900
    path = str(tmp_path / "tune.txt")
    best = ctx.autotune(0, M, N, K, la=skg.ROW_MAJOR, lb=skg.ROW_MAJOR,
                        cache_path=path)
    assert best["us"] > 0 and 0 <= best["tile"] < skg.num_tiles()
    plan = ctx.query_plan(0, M, N, K, la=skg.ROW_MAJOR, lb=skg.ROW_MAJOR)
    assert (plan["tile"], plan["grid"], plan["kchunks"]) == (
        best["tile"], best["grid"], best["kchunks"])
    with skg.Context(0) as c2:
        c2.load_tuning(path)
        p2 = c2.query_plan(0, M, N, K, la=skg.ROW_MAJOR, lb=skg.ROW_MAJOR)
        assert (p2["tile"], p2["grid"], p2["kchunks"]) == (
            plan["tile"], plan["grid"], plan["kchunks"])
        A = rng.uniform(-1, 1, (M, K))
        B = rng.uniform(-1, 1, (K, N))
        D = np.zeros(M * N)
        c2.dgemm_host(0, 0, 0, M, N, K, 1.0, np.ascontiguousarray(A.ravel()),
                      K, np.ascontiguousarray(B.ravel()), N, 0.0, D, N, 0.0)
        assert rel_fro(D, A @ B) < TOL


# ---- projection GEMM hp = Psi^T (H Psi) (sub_gemm.sh:26-30) -------------------

def test_projection_packed_layouts(skg, ctx, orc, rng):
    """hp = Psi^T HPsi with Psi band-packed 16 (PA [Ns/16][Nd][16]) and HPsi
    band-packed 32 (PB [Ns/32][Nd][32], the H*Psi output layout of
    include/nloc_packed_opt2.hpp:1-8), Ns not a multiple of 16 or 32."""
    Ns, Nd = 72, 12288
    P = rng.standard_normal((Nd, Ns)) / np.sqrt(Nd)
    H = rng.standard_normal((Nd, Ns)) / np.sqrt(Nd)
    pa, lda = encode_a(np.ascontiguousarray(P.T), 2, 0)
    pb, ldb = encode_b(H, 3, 0)
    hp = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    dA, dB = dev(skg, ctx, pa), dev(skg, ctx, pb)
    ctx.dgemm(skg.PACK16, skg.PACK32, skg.ROW_MAJOR, Ns, Ns, Nd, 1.0, dA.ptr,
              lda, dB.ptr, ldb, 0.0, hp.ptr, Ns, 0.0)
    ctx.synchronize()
    ref = np.zeros(Ns * Ns)
    # A = Psi^T: element (m, k) = P[k, m] at m*Nd + k (ROW_MAJOR, ld = Nd)
    orc.dgemm(0, 0, 0, Ns, Ns, Nd, 1.0, P.ravel(order="F").copy(), Nd,
              H.ravel().copy(), Ns, 0.0, ref, Ns, 0.0, dd=True)
    assert rel_fro(hp.download(), ref) < TOL


def test_projection_full_size(skg, ctx, orc):
    """Ns = 616, Nd = 175616 (the reference's projection sweep shape): a
    16 x 16 sample of hp against the double-double oracle, and with
    HPsi = Psi (repacked on the device) hp is the SYRK of Psi."""
    Ns, Nd = 616, 175616
    tA, tB = (Ns + 15) // 16, (Ns + 31) // 32
    psi = skg.DeviceArray(ctx, tA * Nd * 16)
    hpsi = skg.DeviceArray(ctx, tB * Nd * 32)
    ctx.fill_normal(psi.ptr, tA * Nd * 16, 11, 1.0 / np.sqrt(Nd))
    ctx.fill_normal(hpsi.ptr, tB * Nd * 32, 12, 1.0 / np.sqrt(Nd))
    hp = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    ctx.dgemm(skg.PACK16, skg.PACK32, skg.ROW_MAJOR, Ns, Ns, Nd, 1.0, psi.ptr,
              Nd, hpsi.ptr, Nd, 0.0, hp.ptr, Ns, 0.0)
    ctx.synchronize()
    got = hp.download().reshape(Ns, Ns)
    pa = psi.download().reshape(tA, Nd, 16)
    pb = hpsi.download().reshape(tB, Nd, 32)
    g = np.random.default_rng(5)
    rows = np.sort(g.choice(Ns, 16, replace=False))
    cols = np.sort(g.choice(Ns, 16, replace=False))
    A_sub = np.stack([pa[i // 16, :, i % 16] for i in rows])      # 16 x Nd
    B_sub = np.stack([pb[j // 32, :, j % 32] for j in cols], 1)   # Nd x 16
    ref = np.zeros(256)
    orc.dgemm(0, 0, 0, 16, 16, Nd, 1.0, A_sub.ravel().copy(), Nd,
              B_sub.ravel().copy(), 16, 0.0, ref, 16, 0.0, dd=True)
    assert rel_fro(got[np.ix_(rows, cols)], ref.reshape(16, 16)) < TOL
    # HPsi = Psi: repack P16 -> P32 on the device, then hp == SYRK(Psi)
    ctx.pack(skg.PACK16, psi.ptr, Nd, skg.PACK32, hpsi.ptr, Nd, Ns, Nd)
    ctx.dgemm(skg.PACK16, skg.PACK32, skg.ROW_MAJOR, Ns, Ns, Nd, 1.0, psi.ptr,
              Nd, hpsi.ptr, Nd, 0.0, hp.ptr, Ns, 0.0)
    ms = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    ctx.dsyrk(skg.PACK16, Ns, Nd, 1.0, psi.ptr, Nd, 0.0, ms.ptr, Ns)
    ctx.synchronize()
    full = hp.download().reshape(Ns, Ns)
    assert rel_fro(full, full.T) < TOL
    il = np.tril_indices(Ns)
    assert rel_fro(full[il], ms.download().reshape(Ns, Ns)[il]) < TOL


@pytest.mark.parametrize("pad", [2, 3])
def test_every_tile_every_layout(skg, ctx, orc, rng, pad):
    """Every tile configuration (register-staged and LDS-DMA) on every A / B
    layout, ragged extents; even leading dimensions give 16-byte aligned
    pairs (the DMA tiles run), odd ones take the register-staged sibling."""
    M, N, K = 78, 46, 134
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    ref = np.zeros(M * N)
    orc.dgemm(0, 0, 0, M, N, K, 1.0, np.ascontiguousarray(A.ravel()), K,
              np.ascontiguousarray(B.ravel()), N, 0.0, ref, N, 0.0, dd=True)
    for tile in range(skg.num_tiles()):
        ctx.set_tuning(tile, 0)
        for la in range(4):
            Ab, lda = encode_a(A, la, pad)
            for lb in range(4):
                Bb, ldb = encode_b(B, lb, pad)
                D = np.zeros(M * N)
                ctx.dgemm_host(la, lb, 0, M, N, K, 1.0, Ab, lda, Bb, ldb, 0.0,
                               D, N, 0.0)
                assert rel_fro(D, ref) < TOL, (tile, la, lb, pad)
    ctx.set_tuning(-1, 0)


STRIP_TILE = 11   # 128 x 128 LDS-DMA tile whose last tile column carries
                  # an N remainder of up to 16 columns
STRIP_TILE_2 = 13  # the same with two DMA stages


@pytest.mark.parametrize("tile", [STRIP_TILE, STRIP_TILE_2])
@pytest.mark.parametrize("M,N,K", [(256, 136, 64), (300, 144, 70),
                                   (128, 130, 520), (384, 520, 37),
                                   (256, 129, 33), (136, 256, 64),
                                   (520, 300, 37), (144, 128, 70)])
def test_strip_tiles(skg, ctx, rng, M, N, K, tile):
    """The strip configuration on every A / B layout, row-major,
    column-major (launched transposed, so a column strip becomes a row
    strip and back) and band-packed D, with alpha, beta and gamma, under
    the default grid, split tiles (Stream-K fix-up of the strip) and K
    chunks.  Column strips (N remainder), row strips (M remainder, N a
    multiple of 128 or its remainder above 16); odd N or K takes the
    register-staged sibling with plain tiling."""
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    D0 = rng.uniform(-1, 1, (M, N))
    alpha, beta, gamma = 0.75, 0.5, 0.25
    ref = alpha * (A @ B) + beta * D0 + gamma
    info = skg.tile_info(tile)
    assert (info["bm"], info["bn"], info["waves"]) == (128, 128, 8)
    try:
        for grid, kchunks in ((0, 1), (5, 1), (0, 3)):
            ctx.set_tuning(tile, grid)
            ctx.set_kchunks(kchunks)
            for la in range(3):
                Ab, lda = encode_a(A, la, 2)
                for lb in range(4):
                    Bb, ldb = encode_b(B, lb, 2)
                    for ld_ in (0, 1, 2):
                        Db, ldd = encode_d(D0, ld_)
                        ctx.dgemm_host(la, lb, ld_, M, N, K, alpha, Ab, lda,
                                       Bb, ldb, beta, Db, ldd, gamma)
                        got = decode_d(Db, ld_, M, N, ldd)
                        err = rel_fro(got, ref)
                        assert err < TOL, (grid, kchunks, la, lb, ld_, err)
    finally:
        ctx.set_tuning(-1, 0)
        ctx.set_kchunks(0)


def test_strip_plan_and_c2_rotation(skg, ctx, rng):
    """N = 520 under the strip tile: 4 tile columns (the last one 136 wide);
    the c2 rotation (grid-packed Psi 32768 x 520, row-major Q, band-packed
    out) agrees with numpy and bit for bit with the plain 128 x 128 tile
    (same MFMA sequence per 16 x 16 block, no split units)."""
    ctx.set_tuning(STRIP_TILE, 0)
    try:
        plan = ctx.query_plan(0, 32768, 520, 520, skg.ROW_MAJOR, skg.ROW_MAJOR)
        assert plan["tile"] == STRIP_TILE and plan["tiles"] == 256 * 4
        plan = ctx.query_plan(0, 32768, 600, 520, skg.ROW_MAJOR, skg.ROW_MAJOR)
        assert plan["tiles"] == 256 * 5            # remainder 88: no strip
    finally:
        ctx.set_tuning(-1, 0)
    Nd, nb = 32768, 520
    psi = rng.standard_normal((Nd, nb)) / np.sqrt(Nd)
    Q = rng.uniform(-1, 1, (nb, nb))
    ref = psi @ Q
    Ab, lda = encode_a(psi, 2, 0)
    out = {}
    for tile in (STRIP_TILE, 8):
        # 256 workgroups: 1024 (1280) tiles of 33 K steps split into whole
        # tiles, 4 (5) per workgroup, so no unit is split.
        ctx.set_tuning(tile, 256)
        try:
            Db = np.zeros(((nb + 15) // 16) * Nd * 16)
            ctx.dgemm_host(skg.PACK16, skg.ROW_MAJOR, skg.PACK16, Nd, nb, nb,
                           1.0, Ab, lda, np.ascontiguousarray(Q.ravel()), nb,
                           0.0, Db, Nd, 0.0)
        finally:
            ctx.set_tuning(-1, 0)
        out[tile] = decode_d(Db, 2, Nd, nb, Nd)
        assert rel_fro(out[tile], ref) < TOL
    assert np.array_equal(out[STRIP_TILE], out[8])


def test_strip_c2_rotation_column_major(skg, ctx, rng):
    """The c2 rotation as the step benchmark runs it: Psi 32768 x 520
    column-major, Q row-major, D column-major, launched transposed (Q^T
    Psi^T, a row strip of 8 on the last tile row) under the autotuner's
    choice of the strip tile."""
    Nd, nb = 32768, 520
    psi = rng.standard_normal((Nd, nb)) / np.sqrt(Nd)
    Q = rng.uniform(-1, 1, (nb, nb))
    ref = psi @ Q
    Ab, lda = encode_a(psi, 1, 0)
    for grid in (0, 100):
        ctx.set_tuning(STRIP_TILE, grid)
        try:
            D = np.zeros(Nd * nb)
            ctx.dgemm_host(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, Nd,
                           nb, nb, 1.0, Ab, lda,
                           np.ascontiguousarray(Q.ravel()), nb, 0.0, D, Nd,
                           0.0)
        finally:
            ctx.set_tuning(-1, 0)
        assert rel_fro(decode_d(D, 1, Nd, nb, Nd), ref) < TOL, grid


def test_banded_tile_order(skg, ctx, rng):
    """Short M, long N, no strip: tiles walk bands of ceil(tiles_n / 8)
    tile columns (here 2 x 17 tiles of 128, bands of 3, the last one
    partial), under the default grid, split tiles and K chunks."""
    M, N, K = 200, 2100, 300
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    ref = A @ B + 0.5
    try:
        for grid, kchunks in ((0, 1), (7, 1), (37, 1), (0, 3)):
            ctx.set_tuning(8, grid)
            ctx.set_kchunks(kchunks)
            D = np.zeros(M * N)
            ctx.dgemm_host(0, 0, 0, M, N, K, 1.0,
                           np.ascontiguousarray(A.ravel()), K,
                           np.ascontiguousarray(B.ravel()), N, 0.0, D, N, 0.5)
            assert rel_fro(D.reshape(M, N), ref) < TOL, (grid, kchunks)
    finally:
        ctx.set_tuning(-1, 0)
        ctx.set_kchunks(0)
