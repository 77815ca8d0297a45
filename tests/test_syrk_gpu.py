"""GPU parity of the FP64 Stream-K SYRK (lower-triangle tiles only) against
the oracle through the C ABI.  Bar: 1e-13 relative Frobenius on the lower
triangle.  Default store: entries above the diagonal are never written;
SYRK_DIAG_BLOCKS (the reference's store, include/sme_syrk.hpp:57-109) also
writes the whole of every 16 x 16 diagonal block."""
import numpy as np
import pytest

from conftest import rel_fro
from test_oracle import SYRK_CASES, encode_b, unpack_p16

pytestmark = pytest.mark.gpu
TOL = 1e-13


def diag_blocks_upper(N):
    """Mask [i, j] of row-major entries j > i inside the 16 x 16 diagonal
    blocks (what SYRK_DIAG_BLOCKS adds to the lower triangle)."""
    r = np.arange(N)
    return (r[None, :] > r[:, None]) & (r[None, :] // 16 == r[:, None] // 16)


@pytest.mark.parametrize("t", range(5))
def test_reference_cases_p16(skg, ctx, orc, t):
    """syrk_crsp.cpp:474-493 through sme_syrk's layouts: P16 input, M x M
    column-major upper triangle (== row-major lower) output, glibc inputs.
    Default store: the rest is untouched.  SYRK_DIAG_BLOCKS: the whole
    output of the oracle's restatement of the tile store (whole 16 x 16
    blocks of every tile pair tm1 <= tm2, sme_syrk.hpp:57-109), the
    strict lower part off the diagonal blocks untouched."""
    K, M = SYRK_CASES[t]
    ap = orc.init_A_packed(K, M, 42 + t)
    ref = orc.syrk_reference(K, M, ap, 1.0, 0.0)      # column-major upper
    C = np.full(M * M, 0.1)                            # syrk_crsp.cpp:499
    ctx.set_tuning(-1, 0)
    ctx.dsyrk_host(skg.PACK16, M, K, 1.0, ap, K, 0.0, C, M)
    iu = np.triu_indices(M)
    got = C.reshape(M, M).T[iu]                        # [m1, m2], m1 <= m2
    want = ref.reshape(M, M).T[iu]
    assert rel_fro(got, want) < TOL
    err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
    assert err.max() < 1e-12                           # reference: 1e-9
    assert np.all(C.reshape(M, M).T[np.tril_indices(M, -1)] == 0.1)
    # The reference's store: whole diagonal blocks.
    full = orc.syrk_tiles(K, M, ap, 1.0, 0.0, C=np.full(M * M, 0.1))
    Cb = np.full(M * M, 0.1)
    ctx.set_syrk_store(skg.SYRK_DIAG_BLOCKS)
    try:
        ctx.dsyrk_host(skg.PACK16, M, K, 1.0, ap, K, 0.0, Cb, M)
    finally:
        ctx.set_syrk_store(skg.SYRK_LOWER)
    written = full != 0.1
    assert rel_fro(Cb[written], full[written]) < TOL
    assert np.all(Cb[~written] == 0.1)
    blk = diag_blocks_upper(M).T                       # [m1, m2], m1 > m2
    assert written.reshape(M, M).T[blk].all()


def _diag_block_ref(A, C0, alpha, beta):
    N = A.shape[1]
    ref = alpha * (A.T @ A) + beta * C0
    keep = np.tril(np.ones((N, N), bool)) | diag_blocks_upper(N)
    return ref, keep


@pytest.mark.parametrize("N,K,tile,kchunks", [(137, 531, -1, 0),
                                              (137, 531, -1, 3),
                                              (200, 64, 12, 1),
                                              (130, 520, 12, 3),
                                              (136, 70, 11, 1),
                                              (64, 300, 1, 2),
                                              (9, 40, -1, 0)])
def test_diag_blocks_plain(skg, ctx, rng, N, K, tile, kchunks):
    """SYRK_DIAG_BLOCKS on the FP64 kernel: ragged N (the last block clipped
    to N), K chunks (the chunk-reduce pass), strip tiles with the corner
    block, alpha and beta; every other entry above the diagonal untouched."""
    A = rng.uniform(-1, 1, (K, N))
    C0 = rng.uniform(-1, 1, (N, N))
    ref, keep = _diag_block_ref(A, C0, 0.5, -1.25)
    ctx.set_syrk_store(skg.SYRK_DIAG_BLOCKS)
    try:
        for layout in (0, 1, 2):
            ctx.set_tuning(tile, 0)
            ctx.set_kchunks(kchunks)
            Ab, lda = encode_b(A, layout, 2)
            C = C0.copy().ravel()
            ctx.dsyrk_host(layout, N, K, 0.5, Ab, lda, -1.25, C, N)
            Cm = C.reshape(N, N)
            assert rel_fro(Cm[keep], ref[keep]) < TOL, layout
            assert np.array_equal(Cm[~keep], C0[~keep]), layout
    finally:
        ctx.set_syrk_store(skg.SYRK_LOWER)
        ctx.set_tuning(-1, 0)
        ctx.set_kchunks(0)


def test_diag_blocks_split_and_batched(skg, ctx, rng):
    """SYRK_DIAG_BLOCKS through the split form (CRT pass) and a batched
    launch (plain FP64 and split): the same entries as the plain call."""
    N, K = 150, 2000
    A = rng.uniform(-1, 1, (K, N))
    C0 = rng.uniform(-1, 1, (N, N))
    ref, keep = _diag_block_ref(A, C0, 1.0, 0.5)
    ctx.set_syrk_store(skg.SYRK_DIAG_BLOCKS)
    try:
        C = C0.copy().ravel()
        bound = ctx.dsyrk_ozaki_host(0, N, K, 1.0, np.ascontiguousarray(A.ravel()),
                                     N, 0.5, C, N)
        Cm = C.reshape(N, N)
        assert 0 < bound < 1e-12
        assert rel_fro(Cm[keep], ref[keep]) < TOL
        assert np.array_equal(Cm[~keep], C0[~keep])
        dA = skg.DeviceArray.from_host(ctx, A)
        for precision in (skg.FP64, skg.OZAKI_INT8):
            dC = skg.DeviceArray.from_host(ctx, C0)
            p = skg.Problem(kind=1, layout_a=0, layout_b=0, layout_d=0, M=N,
                            N=N, K=K, alpha=1.0, beta=0.5, gamma=0.0,
                            A=dA.ptr, lda=N, B=dA.ptr, ldb=N, D=dC.ptr, ldd=N)
            ctx.batched([p], precision=precision)
            ctx.synchronize()
            Cm = dC.download().reshape(N, N)
            assert rel_fro(Cm[keep], ref[keep]) < TOL, precision
            assert np.array_equal(Cm[~keep], C0[~keep]), precision
    finally:
        ctx.set_syrk_store(skg.SYRK_LOWER)


@pytest.mark.parametrize("layout", [0, 1, 2, 3])
def test_layouts_ragged_alpha_beta(skg, ctx, orc, rng, layout):
    N, K = 137, 531
    A = rng.uniform(-1, 1, (K, N))
    Ab, lda = encode_b(A, layout)
    C0 = rng.uniform(-1, 1, (N, N))
    C = C0.copy().ravel()
    ctx.set_kchunks(1 + layout)  # K chunking with beta != 0 and a mask
    ctx.dsyrk_host(layout, N, K, 0.5, Ab, lda, -1.25, C, N)
    ctx.set_kchunks(0)
    ref = 0.5 * (A.T @ A) - 1.25 * C0
    il = np.tril_indices(N)
    Cm = C.reshape(N, N)
    assert rel_fro(Cm[il], ref[il]) < TOL
    iu = np.triu_indices(N, 1)
    assert np.array_equal(Cm[iu], C0[iu])


@pytest.mark.parametrize("N,K", [(1, 1), (1, 1000), (2, 3), (16, 16),
                                 (63, 65), (64, 1), (65, 64), (129, 4097),
                                 (300, 17)])
def test_edge_shapes(skg, ctx, orc, rng, N, K):
    A = rng.uniform(-1, 1, (K, N))
    ref = np.zeros(N * N)
    orc.dsyrk(0, N, K, 1.0, np.ascontiguousarray(A.ravel()), N, 0.0, ref, N)
    C = np.zeros(N * N)
    ctx.dsyrk_host(0, N, K, 1.0, np.ascontiguousarray(A.ravel()), N, 0.0, C, N)
    il = np.tril_indices(N)
    assert rel_fro(C.reshape(N, N)[il], ref.reshape(N, N)[il]) < TOL


def test_partition_invariance(skg, ctx, orc, rng):
    N, K = 200, 3000
    A = rng.uniform(-1, 1, (K, N)).ravel()
    ref = np.zeros(N * N)
    orc.dsyrk(0, N, K, 1.0, A, N, 0.0, ref, N, dd=True)
    il = np.tril_indices(N)
    for tile in range(skg.num_tiles()):
        info = skg.tile_info(tile)
        if info["bm"] != info["bn"]:
            continue
        for grid, kchunks in ((1, 1), (2, 1), (5, 3), (33, 1), (0, 1),
                              (0, 8), (24, 8)):
            ctx.set_tuning(tile, grid)
            ctx.set_kchunks(kchunks)
            C = np.zeros(N * N)
            ctx.dsyrk_host(0, N, K, 1.0, A, N, 0.0, C, N)
            err = rel_fro(C.reshape(N, N)[il], ref.reshape(N, N)[il])
            assert err < TOL, (tile, grid, kchunks, err)
    ctx.set_tuning(-1, 0)
    ctx.set_kchunks(0)


def test_config2_syrk_full_size(skg, ctx, orc):
    """c2: Psi 32768 x 520 normal(0, 1/sqrt(Nd)) column-major.  Sampled rows
    of Ms against the double-double oracle; diagonal against column norms;
    agreement with the GEMM path Psi^T Psi."""
    Nd, Ns = 32768, 520
    psi = skg.DeviceArray(ctx, Nd * Ns)
    ms = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    ctx.fill_normal(psi.ptr, Nd * Ns, 1, 1.0 / np.sqrt(Nd))
    ctx.set_tuning(-1, 0)
    ctx.dsyrk(skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0, ms.ptr, Ns)
    full = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    # GEMM path: A = Psi^T (ROW_MAJOR in A convention: m*ld + k with ld = Nd)
    ctx.dgemm(skg.ROW_MAJOR, skg.COL_MAJOR, skg.ROW_MAJOR, Ns, Ns, Nd, 1.0,
              psi.ptr, Nd, psi.ptr, Nd, 0.0, full.ptr, Ns, 0.0)
    ctx.synchronize()
    P = psi.download().reshape(Ns, Nd)                 # row b = column b
    Ms = ms.download().reshape(Ns, Ns)
    Mf = full.download().reshape(Ns, Ns)
    rows = np.sort(np.random.default_rng(11).choice(Ns, 48, replace=False))
    sub = np.ascontiguousarray(P[rows])                # 48 x Nd
    ref = np.zeros(48 * Ns)
    orc.dgemm(0, 1, 0, 48, Ns, Nd, 1.0, sub.ravel(), Nd, P.ravel().copy(), Nd,
              0.0, ref, Ns, 0.0, dd=True)
    ref = ref.reshape(48, Ns)
    got = np.concatenate([Ms[r, :r + 1] for r in rows])
    want = np.concatenate([ref[i, :r + 1] for i, r in enumerate(rows)])
    assert rel_fro(got, want) < TOL
    norms = np.einsum("ij,ij->i", P, P)
    assert rel_fro(np.diag(Ms), norms) < 1e-14
    il = np.tril_indices(Ns)
    assert rel_fro(Ms[il], Mf[il]) < 1e-14


def test_orthonormal_psi_gives_identity(skg, ctx, rng):
    K, N = 20000, 96
    Q, _ = np.linalg.qr(rng.standard_normal((K, N)))
    Qc = np.asfortranarray(Q).ravel(order="F").copy()   # column-major
    C = np.zeros(N * N)
    ctx.dsyrk_host(skg.COL_MAJOR, N, K, 1.0, Qc, K, 0.0, C, N)
    il = np.tril_indices(N)
    assert np.max(np.abs(C.reshape(N, N)[il] - np.eye(N)[il])) < 1e-13


def test_power_of_two_scaling_exact(skg, ctx, rng):
    """Column scaling by 2^e scales Ms exactly (same partition, same
    summation order)."""
    K, N = 4096, 80
    A = rng.uniform(-1, 1, (K, N))
    e = rng.integers(-30, 30, N)
    As = A * np.ldexp(1.0, e)[None, :]
    ctx.set_tuning(1, 50)
    C = np.zeros(N * N)
    Cs = np.zeros(N * N)
    ctx.dsyrk_host(0, N, K, 1.0, np.ascontiguousarray(A.ravel()), N, 0.0, C, N)
    ctx.dsyrk_host(0, N, K, 1.0, np.ascontiguousarray(As.ravel()), N, 0.0, Cs, N)
    ctx.set_tuning(-1, 0)
    il = np.tril_indices(N)
    scale = np.ldexp(1.0, e[:, None] + e[None, :])
    assert np.array_equal(Cs.reshape(N, N)[il], (C.reshape(N, N) * scale)[il])


def test_every_square_tile_column_major(skg, ctx, orc, rng):
    """SYRK on column-major Psi (contiguous along K: the k-contiguous LDS
    image of the DMA tiles) with every square tile configuration, split and
    unsplit K."""
    N, K = 136, 2000
    P = rng.standard_normal((K, N))
    colmajor = P.ravel(order="F").copy()        # element (k, n) at k + n*K
    ref = np.zeros(N * N)
    orc.dsyrk(0, N, K, 1.0, P.ravel().copy(), N, 0.0, ref, N, dd=True)
    il = np.tril_indices(N)
    for tile in range(skg.num_tiles()):
        info = skg.tile_info(tile)
        if info["bm"] != info["bn"]:
            continue
        for grid, kchunks in ((0, 1), (0, 4), (37, 1)):
            ctx.set_tuning(tile, grid)
            ctx.set_kchunks(kchunks)
            C = np.zeros(N * N)
            ctx.dsyrk_host(1, N, K, 1.0, colmajor, K, 0.0, C, N)
            err = rel_fro(C.reshape(N, N)[il], ref.reshape(N, N)[il])
            assert err < TOL, (tile, grid, kchunks, err)
    ctx.set_tuning(-1, 0)
    ctx.set_kchunks(0)


@pytest.mark.parametrize("tile,N,K", [(12, 72, 100), (12, 80, 33),
                                      (12, 130, 520), (12, 200, 64),
                                      (11, 136, 70), (11, 144, 37),
                                      (14, 80, 33), (14, 200, 64),
                                      (13, 136, 70), (17, 72, 16),
                                      (17, 80, 33), (17, 130, 520),
                                      (17, 200, 64), (17, 200, 80)])
def test_strip_tiles(skg, ctx, rng, tile, N, K):
    """Strip tiles on a SYRK: N = T x BM + r (r <= 16) tiles a T x T
    triangle whose last tile row carries the r remainder rows; wave 0 of the
    last diagonal tile forms the r x r corner from the strip alone.  Every
    layout, alpha and beta, default grid, split tiles and K chunks; the
    upper triangle is never written.  (N = 200 has remainder 8 on 64-wide
    tiles; tile 12 is 64 x 64, tile 11 128 x 128, tiles 14 and 13 the same
    with two DMA stages, tile 17 as 12 with paired steps: K = 16, 33, 64 and
    80 give units of 1, 3, 4 and 5 steps.)"""
    A = rng.uniform(-1, 1, (K, N))
    C0 = rng.uniform(-1, 1, (N, N))
    ref = 0.5 * (A.T @ A) - 1.25 * C0
    il, iu = np.tril_indices(N), np.triu_indices(N, 1)
    try:
        for grid, kchunks in ((0, 1), (3, 1), (0, 3)):
            ctx.set_tuning(tile, grid)
            ctx.set_kchunks(kchunks)
            for layout in range(4):
                Ab, lda = encode_b(A, layout, 2)
                C = C0.copy().ravel()
                ctx.dsyrk_host(layout, N, K, 0.5, Ab, lda, -1.25, C, N)
                Cm = C.reshape(N, N)
                err = rel_fro(Cm[il], ref[il])
                assert err < TOL, (grid, kchunks, layout, err)
                assert np.array_equal(Cm[iu], C0[iu])
    finally:
        ctx.set_tuning(-1, 0)
        ctx.set_kchunks(0)


def test_strip_c2_syrk(skg, ctx, rng):
    """The c2 SYRK (Psi 32768 x 520 column-major) on the 64 x 64 strip tile:
    8 x 9 / 2 = 36 tiles (the plan reports them), the 8 remainder rows on
    the last tile row and the 8 x 8 corner."""
    Nd, N = 32768, 520
    ctx.set_tuning(12, 0)
    try:
        plan = ctx.query_plan(1, N, N, Nd, skg.COL_MAJOR)
        assert plan["tile"] == 12 and plan["tiles"] == 36
        psi = rng.standard_normal((Nd, N)) / np.sqrt(Nd)
        Ab, lda = encode_b(psi, 1, 0)
        C = np.zeros(N * N)
        ctx.dsyrk_host(skg.COL_MAJOR, N, Nd, 1.0, Ab, lda, 0.0, C, N)
    finally:
        ctx.set_tuning(-1, 0)
    ref = psi.T @ psi
    il = np.tril_indices(N)
    assert rel_fro(C.reshape(N, N)[il], ref[il]) < TOL
