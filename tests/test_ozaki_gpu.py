"""GPU parity of the split-precision (Ozaki scheme II, INT8 MFMA) GEMM and
SYRK.  Bars: bit-for-bit equality with the CPU restatement (oracle/ozaki.py:
exact integer product, no CRT) on the same inputs; relative Frobenius error
<= 1e-13 against the double-double oracle and <= the returned bound on the
c3-style inputs (columns scaled by 10^u, u ~ U(0, 6))."""
import os

import numpy as np
import pytest

from conftest import rel_fro
from oracle import ozaki
from test_oracle import decode_d, encode_a, encode_b, encode_d

pytestmark = pytest.mark.gpu


def c3_like(rng, K, N):
    P = rng.standard_normal((K, N)) / np.sqrt(K)
    return P * 10.0 ** rng.uniform(0, 6, N)[None, :]


@pytest.mark.parametrize("layout", [0, 1, 2, 3])
def test_syrk_bit_exact_vs_restatement(skg, ctx, rng, layout):
    K, N = 300, 70
    X = c3_like(rng, K, N)                     # K x N
    Xb, ld = encode_b(X, layout)
    C = np.full(N * N, 9.0)
    bound = ctx.dsyrk_ozaki_host(layout, N, K, 1.0, Xb, ld, 0.0, C, N)
    ref, _, _, _ = ozaki.product(X.T.copy(), X.T.copy(), 14)
    Cm = C.reshape(N, N)
    il = np.tril_indices(N)
    assert np.array_equal(Cm[il], ref[il])
    assert np.all(Cm[np.triu_indices(N, 1)] == 9.0)
    assert 0 < bound < 1e-13


def test_gemm_bit_exact_all_layouts(skg, ctx, rng):
    M, N, K = 50, 40, 200
    A = rng.standard_normal((M, K)) * 10.0 ** rng.uniform(0, 4, M)[:, None]
    B = rng.standard_normal((K, N)) * 10.0 ** rng.uniform(0, 4, N)[None, :]
    ref, _, _, _ = ozaki.product(A, B.T.copy(), 14)
    for la in range(4):
        Ab, lda = encode_a(A, la)
        for lb in range(4):
            Bb, ldb = encode_b(B, lb)
            for ld_ in (0, 1, 2):
                Db, ldd = encode_d(np.zeros((M, N)), ld_)
                ctx.dgemm_ozaki_host(la, lb, ld_, M, N, K, 1.0, Ab, lda, Bb, ldb, 0.0,
                                Db, ldd, 0.0, slices=14)
                got = decode_d(Db, ld_, M, N, ldd)
                assert np.array_equal(got, ref), (la, lb, ld_)


def test_partition_invariance_is_bitwise(skg, ctx, rng):
    """Integer partials: any grid gives the identical result."""
    K, N = 5000, 200
    X = c3_like(rng, K, N).ravel()
    outs = []
    for grid in (0, 1, 3, 64, 257):
        ctx.set_tuning(-1, grid)
        C = np.zeros(N * N)
        ctx.dsyrk_ozaki_host(0, N, K, 1.0, X, N, 0.0, C, N)
        outs.append(C)
    ctx.set_tuning(-1, 0)
    for C in outs[1:]:
        assert np.array_equal(C, outs[0])


def test_small_integer_inputs_exact(skg, ctx, rng):
    M, N, K = 33, 65, 129
    A = rng.integers(-1000, 1000, (M, K)).astype(np.float64)
    B = rng.integers(-1000, 1000, (K, N)).astype(np.float64)
    D = np.zeros(M * N)
    ctx.dgemm_ozaki_host(0, 0, 0, M, N, K, 1.0, np.ascontiguousarray(A.ravel()), K,
                    np.ascontiguousarray(B.ravel()), N, 0.0, D, N, 0.0)
    assert np.array_equal(D.reshape(M, N), A @ B)


def test_alpha_beta_gamma(skg, ctx, rng):
    M, N, K = 120, 90, 333
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    C0 = rng.uniform(-1, 1, (M, N))
    D = C0.copy().ravel()
    b = ctx.dgemm_ozaki_host(0, 0, 0, M, N, K, 0.5, np.ascontiguousarray(A.ravel()),
                        K, np.ascontiguousarray(B.ravel()), N, -2.0, D, N, 0.25)
    ref = 0.5 * (A @ B) - 2.0 * C0 + 0.25
    assert rel_fro(D, ref) < 1e-13 and b < 1e-13


@pytest.mark.parametrize("K,N", [(16384, 256), (2048, 520)])
def test_c3_style_accuracy_vs_double_double(skg, ctx, orc, rng, K, N):
    X = c3_like(rng, K, N)
    Xc = np.asfortranarray(X).ravel(order="F").copy()   # column-major
    C = np.zeros(N * N)
    bound = ctx.dsyrk_ozaki_host(skg.COL_MAJOR, N, K, 1.0, Xc, K, 0.0, C, N)
    ref = np.zeros(N * N)
    orc.dsyrk(0, N, K, 1.0, np.ascontiguousarray(X.ravel()), N, 0.0, ref, N,
              dd=True)
    il = np.tril_indices(N)
    err = rel_fro(C.reshape(N, N)[il], ref.reshape(N, N)[il])
    print("K", K, "N", N, "err", err, "bound", bound, flush=True)
    assert err < 1e-13 and err <= bound and bound < 1e-13


def test_rotation_split_accuracy(skg, ctx, orc, rng):
    """Psi * Q in the split form: rows of Psi carry the 10^u column scales."""
    Nd, Ns = 4096, 256
    P = c3_like(rng, Nd, Ns)
    Q = rng.uniform(-1, 1, (Ns, Ns))
    Pc = np.asfortranarray(P).ravel(order="F").copy()
    D = np.zeros(Nd * Ns)
    bound = ctx.dgemm_ozaki_host(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, Nd, Ns,
                            Ns, 1.0, Pc, Nd, Q.ravel().copy(), Ns, 0.0, D, Nd,
                            0.0)
    ref = np.zeros(Nd * Ns)
    orc.dgemm(1, 0, 1, Nd, Ns, Ns, 1.0, Pc, Nd, Q.ravel().copy(), Ns, 0.0, ref,
              Nd, 0.0, dd=True)
    err = rel_fro(D, ref)
    assert err < 1e-13 and err <= bound


def test_split_k_limit(skg, ctx):
    """K beyond 32 chunks of 131072 products (the CRT's chunk-sum range) is
    refused before any launch."""
    K = 32 * 131072 + 1
    a = skg.DeviceArray(ctx, 64)
    c = skg.DeviceArray(ctx, 64)
    with pytest.raises(skg.SkgemmError) as e:
        ctx.dsyrk_ozaki(skg.COL_MAJOR, 8, K, 1.0, a.ptr, K, 0.0, c.ptr, 8)
    assert e.value.code == skg.ERR_NOTSUP
    a.free()
    c.free()


@pytest.mark.parametrize("K,N,layout", [(140000, 40, 0), (175616, 64, 1),
                                         (262400, 24, 2)])
def test_long_k_in_chunks_bit_exact(skg, ctx, rng, K, N, layout):
    """K beyond the 131072 products one INT32 sum holds exactly: the plan
    cuts K into chunks and the CRT pass adds their residues modulo each
    modulus.  Bit for bit against the exact-integer restatement, for the
    chooser's plan and for plans that ask for fewer chunks than needed (the
    planner raises them); the bound as for short K."""
    X = c3_like(rng, K, N)
    Xb, ld = encode_b(X, layout)
    ref, _, _ = ozaki.product_rows(X.T.copy(), X.T.copy(), 14)
    il = np.tril_indices(N)
    try:
        # "11,1,0,32": the most chunks the CRT adds without a reduction
        # (their plain sum of signed residue bytes); "12,1,0,9": the
        # five-stage 256 x 256 tile; "10,1,0,5": the early-barrier one.
        for plan in (None, "6,1", "5,1,0,1", "1,0", "11,1,0,32", "12,1,0,9",
                     "10,1,0,5"):
            if plan is None:
                os.environ.pop("SKGEMM_OZ_PLAN", None)
            else:
                os.environ["SKGEMM_OZ_PLAN"] = plan
            print("K", K, "N", N, "plan", plan, flush=True)
            C = np.zeros(N * N)
            bound = ctx.dsyrk_ozaki_host(layout, N, K, 1.0, Xb, ld, 0.0, C, N)
            assert np.array_equal(C.reshape(N, N)[il], ref[il]), plan
            assert 0 < bound < 1e-12
    finally:
        os.environ.pop("SKGEMM_OZ_PLAN", None)


def test_reference_sweep_shape_split(skg, ctx, orc):
    """The reference's SYRK sweep shape (sub_syrk.sh:26-29: M = 616,
    K = Nd = 175616) in the split form on Psi normal(0, 1/sqrt(Nd)) with
    10^U(0,6) column scales: sampled rows against the double-double oracle
    and, bit for bit, against the exact-integer restatement; certifying
    mode reaches the 1e-13 bound."""
    Nd, Ns = 175616, 616
    psi = skg.DeviceArray(ctx, Nd * Ns)
    ctx.fill_normal(psi.ptr, Nd * Ns, 21, 1.0 / np.sqrt(Nd))
    ctx.scale_columns_pow10(psi.ptr, Nd, Ns, Nd, 22, 6.0)
    ms = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    bound = ctx.dsyrk_ozaki(skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0,
                            ms.ptr, Ns, slices=14)
    C = ms.download().reshape(Ns, Ns)
    bound_c = ctx.dsyrk_ozaki(skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0,
                              ms.ptr, Ns, slices=skg.SLICES_AUTO)
    Lc = ctx.split_moduli()
    Cc = ms.download().reshape(Ns, Ns)
    P = psi.download().reshape(Ns, Nd)
    psi.free()
    ms.free()
    rows = np.sort(np.random.default_rng(23).choice(Ns, 12, replace=False))
    sub = np.ascontiguousarray(P[rows])
    ref = np.zeros(len(rows) * Ns)
    orc.dgemm(0, 1, 0, len(rows), Ns, Nd, 1.0, sub.ravel(), Nd, P.ravel(), Nd,
              0.0, ref, Ns, 0.0, dd=True)
    ref = ref.reshape(len(rows), Ns)
    low = lambda X: np.concatenate([X[i, :r + 1] for i, r in enumerate(rows)])
    err = rel_fro(low(C[rows]), low(ref))
    err_c = rel_fro(low(Cc[rows]), low(ref))
    print("175616 x 616: err %.3e bound %.3e; certifying %d moduli err %.3e "
          "bound %.3e" % (err, bound, Lc, err_c, bound_c), flush=True)
    assert err < 1e-13 and err <= bound
    assert 8 <= Lc <= 16 and 0 < bound_c <= 1e-13 and err_c <= bound_c
    norms = np.einsum("ij,ij->i", P, P)
    for L, X in ((14, C), (Lc, Cc)):
        exact, _, _ = ozaki.product_rows(sub, P, L, norm2_a=norms[rows],
                                         norm2_b=norms)
        # columns whose scale code is stable under the host's other
        # summation order of the norms
        _, h = ozaki.tables(L)
        c0 = ozaki.scale_code(norms, h)
        ok = ((ozaki.scale_code(norms * (1 + 1e-12), h) == c0) &
              (ozaki.scale_code(norms * (1 - 1e-12), h) == c0))
        for i, r in enumerate(rows):
            if ok[r]:
                cols = np.arange(r + 1)[ok[:r + 1]]
                assert np.array_equal(X[r, cols], exact[i, cols]), (L, r)


def test_certifying_syrk(skg, ctx, orc, rng):
    """SKGEMM_SLICES_AUTO on a SYRK: the fewest moduli whose bound is at most
    1e-13, chosen from the column norms before the products; the result is
    bit for bit the split form at that count, within the bound of the
    double-double oracle; a count below it would not certify."""
    K, N = 16384, 256
    X = c3_like(rng, K, N)
    Xc = np.asfortranarray(X).ravel(order="F").copy()
    C = np.zeros(N * N)
    bound = ctx.dsyrk_ozaki_host(skg.COL_MAJOR, N, K, 1.0, Xc, K, 0.0, C, N,
                                 slices=skg.SLICES_AUTO)
    L = ctx.split_moduli()
    assert 8 <= L <= 16 and 0 < bound <= 1e-13
    ref, _, _ = ozaki.product_rows(X.T.copy(), X.T.copy(), L)
    il = np.tril_indices(N)
    assert np.array_equal(C.reshape(N, N)[il], ref[il])
    dd = np.zeros(N * N)
    orc.dsyrk(0, N, K, 1.0, np.ascontiguousarray(X.ravel()), N, 0.0, dd, N,
              dd=True)
    assert rel_fro(C.reshape(N, N)[il], dd.reshape(N, N)[il]) <= bound
    # one modulus fewer: its a-posteriori bound misses 1e-13
    C2 = np.zeros(N * N)
    b2 = ctx.dsyrk_ozaki_host(skg.COL_MAJOR, N, K, 1.0, Xc, K, 0.0, C2, N,
                              slices=L - 1)
    assert b2 > 1e-13


def test_certifying_gemm(skg, ctx, rng):
    """SKGEMM_SLICES_AUTO on a GEMM: an ordinary product certifies at 14
    moduli; a product that cancels to rounding level (B in the null space
    of A's rows) cannot be certified by 16 moduli and is computed by the
    plain FP64 path (bound -1, modulus count 0), from the original D when
    beta != 0 (D is restored before every further attempt)."""
    M, N, K = 64, 48, 512
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    D = np.zeros(M * N)
    b = ctx.dgemm_ozaki_host(0, 0, 0, M, N, K, 1.0, np.ascontiguousarray(A.ravel()),
                             K, np.ascontiguousarray(B.ravel()), N, 0.0, D, N,
                             slices=skg.SLICES_AUTO)
    assert ctx.split_moduli() == 14 and 0 < b <= 1e-13
    Q, _ = np.linalg.qr(A.T)                       # K x M, rows of A's span
    Bn = B - Q @ (Q.T @ B)                         # A Bn ~ 1e-15
    C0 = rng.uniform(-1, 1, (M, N))
    D1 = C0.copy().ravel()
    D2 = C0.copy().ravel()
    Ab = np.ascontiguousarray(A.ravel())
    Bb = np.ascontiguousarray(Bn.ravel())
    b = ctx.dgemm_ozaki_host(0, 0, 0, M, N, K, 1.5, Ab, K, Bb, N, 0.5, D1, N,
                             0.25, slices=skg.SLICES_AUTO)
    assert b == -1.0 and ctx.split_moduli() == 0
    ctx.dgemm_host(0, 0, 0, M, N, K, 1.5, Ab, K, Bb, N, 0.5, D2, N, 0.25)
    assert np.array_equal(D1, D2)


def test_stage_timing(skg, ctx, rng):
    """skgemm_timing_stages: three event pairs per split-form call (residue
    pass, INT8 products, CRT), one per call otherwise."""
    K, N = 8192, 512
    dA = skg.DeviceArray.from_host(ctx, c3_like(rng, K, N))
    dC = skg.DeviceArray(ctx, N * N, zero=True)
    ctx.timing_enable(32)
    try:
        ctx.timing_stages(True)
        for _ in range(3):
            ctx.dsyrk_ozaki(skg.ROW_MAJOR, N, K, 1.0, dA.ptr, N, 0.0, dC.ptr,
                            N, bound=False)
        t = ctx.timing_read()
        assert len(t) == 9 and np.all(t > 0)
        ctx.timing_stages(False)
        for _ in range(2):
            ctx.dsyrk_ozaki(skg.ROW_MAJOR, N, K, 1.0, dA.ptr, N, 0.0, dC.ptr,
                            N, bound=False)
        t2 = ctx.timing_read()
        assert len(t2) == 2
        assert abs(np.median(t2) - np.median(t.reshape(3, 3).sum(1))) < \
            0.5 * np.median(t2)
    finally:
        ctx.timing_stages(False)
        ctx.timing_enable(0)


def test_device_entry_points_reject_host_pointers(skg, ctx):
    """A host array handed to a device-pointer entry point is refused with
    SKGEMM_ERR_ARG before any kernel runs."""
    A = np.ones(64 * 64)
    C = np.zeros(64 * 64)
    for call in (lambda: ctx.dsyrk_ozaki(0, 64, 64, 1.0, A, 64, 0.0, C, 64),
                 lambda: ctx.dsyrk(0, 64, 64, 1.0, A, 64, 0.0, C, 64),
                 lambda: ctx.dgemm(0, 0, 0, 64, 64, 64, 1.0, A, 64, A, 64, 0.0,
                                   C, 64)):
        with pytest.raises(skg.SkgemmError) as e:
            call()
        assert e.value.code == skg.ERR_ARG


@pytest.mark.parametrize("L", [8, 12, 13, 14, 15, 16])
def test_moduli_counts_bit_exact(skg, ctx, rng, L):
    """Modulus counts: up to 14 moduli the two-limb residue split and the
    three-limb (37-bit) CRT, above it the three-limb residue split (|x| up
    to 2^60) and the four-limb (32-bit) CRT."""
    K, N = 257, 40
    X = c3_like(rng, K, N)
    C = np.zeros(N * N)
    ctx.dsyrk_ozaki_host(0, N, K, 1.0, np.ascontiguousarray(X.ravel()), N, 0.0,
                         C, N, slices=L)
    ref, _, _, _ = ozaki.product(X.T.copy(), X.T.copy(), L)
    il = np.tril_indices(N)
    assert np.array_equal(C.reshape(N, N)[il], ref[il])


def test_modulus_count_range_is_checked(skg, ctx):
    A = np.ones(16)
    C = np.zeros(16)
    for L in (7, 17):
        with pytest.raises(skg.SkgemmError) as e:
            ctx.dsyrk_ozaki_host(0, 4, 4, 1.0, A, 4, 0.0, C, 4, slices=L)
        assert e.value.code == skg.ERR_ARG


def test_every_plan_gives_identical_bits(skg, ctx, rng):
    """Data-parallel and Stream-K schedules of every INT8 tile configuration
    produce the same integers, hence the same FP64 result."""
    K, N = 1500, 600
    X = c3_like(rng, K, N).ravel()
    outs = []
    try:
        for plan in ("0,1", "1,1", "2,1", "3,1", "0,0", "1,0,77", "2,0",
                     "3,0,300", "0,1,0,3", "1,1,100,5", "3,1,0,7",
                     "5,1,0,3", "6,1", "5,0", "6,0,200", "4,1,0,2",
                     "7,1", "8,1,0,2", "9,1,0,3", "7,0", "9,0,150"):
            os.environ["SKGEMM_OZ_PLAN"] = plan
            C = np.zeros(N * N)
            ctx.dsyrk_ozaki_host(0, N, K, 1.0, X, N, 0.0, C, N)
            outs.append(C)
    finally:
        os.environ.pop("SKGEMM_OZ_PLAN", None)
    for C in outs[1:]:
        assert np.array_equal(C, outs[0])


def test_rotation_plans_identical(skg, ctx, rng):
    M, N, K = 2000, 300, 700
    A = rng.standard_normal((M, K))
    B = rng.uniform(-1, 1, (K, N))
    Ab = np.ascontiguousarray(A.ravel())
    Bb = np.ascontiguousarray(B.ravel())
    outs = []
    try:
        for plan in ("0,1", "1,1", "2,0", "3,0", "5,1", "6,1", "5,0", "7,1",
                     "8,1", "9,1", "9,0"):
            os.environ["SKGEMM_OZ_PLAN"] = plan
            D = np.zeros(M * N)
            ctx.dgemm_ozaki_host(0, 0, 0, M, N, K, 1.0, Ab, K, Bb, N, 0.0, D, N)
            outs.append(D)
    finally:
        os.environ.pop("SKGEMM_OZ_PLAN", None)
    for D in outs[1:]:
        assert np.array_equal(D, outs[0])
    assert rel_fro(outs[0], (A @ B).ravel()) < 1e-13


def test_data_parallel_syrk_vs_double_double(skg, ctx, orc, rng):
    """N = 1024 takes the data-parallel schedule (36 tiles of 128 per
    modulus)."""
    K, N = 2048, 1024
    X = c3_like(rng, K, N)
    Xc = np.asfortranarray(X).ravel(order="F").copy()
    C = np.zeros(N * N)
    bound = ctx.dsyrk_ozaki_host(skg.COL_MAJOR, N, K, 1.0, Xc, K, 0.0, C, N)
    ref = np.zeros(N * N)
    orc.dsyrk(0, N, K, 1.0, np.ascontiguousarray(X.ravel()), N, 0.0, ref, N,
              dd=True)
    il = np.tril_indices(N)
    err = rel_fro(C.reshape(N, N)[il], ref.reshape(N, N)[il])
    assert err < 1e-13 and err <= bound


@pytest.mark.parametrize("kind", ["nan", "inf", "tiny", "huge"])
def test_fallback_to_plain_fp64(skg, ctx, rng, kind):
    """A row whose scale cannot be represented (non-finite, or a squared norm
    outside [2^-1000, 2^1000]) sends the call to the plain FP64 path: the
    result is bit-identical to skgemm_dgemm and the bound reads -1."""
    M, N, K = 70, 50, 90
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    if kind == "nan":
        A[3, 5] = np.nan
    elif kind == "inf":
        B[7, 2] = np.inf
    elif kind == "tiny":
        A[10] *= 1e-170
    else:
        B[:, 4] *= 1e160
    C0 = rng.uniform(-1, 1, (M, N))
    D1 = C0.copy().ravel()
    D2 = C0.copy().ravel()
    Ab = np.ascontiguousarray(A.ravel())
    Bb = np.ascontiguousarray(B.ravel())
    b = ctx.dgemm_ozaki_host(0, 0, 0, M, N, K, 1.5, Ab, K, Bb, N, 0.5, D1, N,
                             0.25)
    ctx.dgemm_host(0, 0, 0, M, N, K, 1.5, Ab, K, Bb, N, 0.5, D2, N, 0.25)
    assert b == -1.0
    assert np.array_equal(D1, D2, equal_nan=True)


@pytest.mark.parametrize("K,N", [(20000, 300), (40000, 100)])
def test_pipelined_split_syrk_identical(skg, ctx, rng, K, N):
    """With SKGEMM_OZ_PIPE=1, a k-contiguous operand under a data-parallel
    plan takes the pipelined sequence: K is cut into chunks of whole 4096-wide residue segments
    (K = 20000: five chunks of one segment; K = 40000: five chunks of two),
    the residue pass of chunk c + 1 runs beside the INT8 products of chunk c
    on a second stream, and every chunk after the first writes its own
    residue planes.  Its integers, hence its result, equal those of the
    single-launch data-parallel plan (no SKGEMM_OZ_PIPE) and of Stream-K,
    and it stays within 1e-13 of the plain FP64 SYRK."""
    X = c3_like(rng, K, N)
    Xc = np.ascontiguousarray(X.T.ravel())   # column-major K x N, ld = K
    outs = []
    try:
        for plan, pipe in (("6,1", "1"), ("6,1,0,3", "1"), ("6,1", None),
                           ("1,0", "1"), (None, None)):
            for key, val in (("SKGEMM_OZ_PLAN", plan), ("SKGEMM_OZ_PIPE", pipe)):
                if val is None:
                    os.environ.pop(key, None)
                else:
                    os.environ[key] = val
            print("plan", plan, "pipe", pipe, "K", K, "N", N, flush=True)
            C = np.zeros(N * N)
            ctx.dsyrk_ozaki_host(skg.COL_MAJOR, N, K, 1.0, Xc, K, 0.0, C, N)
            outs.append(C)
    finally:
        os.environ.pop("SKGEMM_OZ_PLAN", None)
        os.environ.pop("SKGEMM_OZ_PIPE", None)
    for C in outs[1:]:
        assert np.array_equal(C, outs[0])
    ref = np.zeros(N * N)
    ctx.dsyrk_host(skg.COL_MAJOR, N, K, 1.0, Xc, K, 0.0, ref, N)
    il = np.tril_indices(N)
    assert rel_fro(outs[0].reshape(N, N)[il], ref.reshape(N, N)[il]) < 1e-13


@pytest.mark.parametrize("K,N", [(5000, 70), (700, 33), (4096, 300),
                                 (9000, 257)])
def test_fused_norm_split_identical(skg, ctx, rng, K, N):
    """A k-contiguous operand takes one kernel for norms, scales and
    residues; it reproduces the two-kernel pass (SKGEMM_OZ_FUSE=0) bit for
    bit: K with a remainder past the last 1024 block, K below 768, row
    counts that leave padding rows, SYRK and GEMM.  (The restatement
    checks above run through the fused pass for their k-contiguous
    layouts.)"""
    X = c3_like(rng, K, N)
    Xc = np.ascontiguousarray(X.T.ravel())   # column-major K x N, ld = K
    A = rng.standard_normal((N, K)) * 10.0 ** rng.uniform(0, 4, N)[:, None]
    B = rng.standard_normal((K, 40))
    outs = []
    try:
        for fuse in (None, "0"):
            if fuse is None:
                os.environ.pop("SKGEMM_OZ_FUSE", None)
            else:
                os.environ["SKGEMM_OZ_FUSE"] = fuse
            print("fuse", fuse, "K", K, "N", N, flush=True)
            C = np.zeros(N * N)
            b1 = ctx.dsyrk_ozaki_host(skg.COL_MAJOR, N, K, 1.0, Xc, K, 0.0, C,
                                      N)
            D = np.zeros(N * 40)
            b2 = ctx.dgemm_ozaki_host(0, 0, 0, N, 40, K, 1.0,
                                      np.ascontiguousarray(A.ravel()), K,
                                      np.ascontiguousarray(B.ravel()), 40,
                                      0.0, D, 40, 0.0)
            outs.append((C, D, b1, b2))
    finally:
        os.environ.pop("SKGEMM_OZ_FUSE", None)
    (C0, D0, b10, b20), (C1, D1, b11, b21) = outs
    assert np.array_equal(C0, C1) and np.array_equal(D0, D1)
    # the bounds' rounding-error sums are added in another order
    assert abs(b10 - b11) <= 1e-9 * b10 and abs(b20 - b21) <= 1e-9 * b20
