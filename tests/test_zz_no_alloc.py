"""skgemm_dsyrk_noalloc (the allocation-free SYRK, its scratch in a device
array of the library's) against skgemm_dsyrk through the C ABI, on a fresh
context: two calls in a row, the second on a smaller input, each within
1e-13 relative Frobenius of skgemm_dsyrk on the lower triangle with the
upper triangle untouched; a plan beyond the scratch limit is rejected with
SKGEMM_ERR_NOMEM and runs once its K chunks are dropped."""
import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu


def run_pair(skg, ctx, la, N, K, A_host, lda, alpha, beta, C0):
    dA = skg.DeviceArray.from_host(ctx, A_host)
    ref = skg.DeviceArray.from_host(ctx, C0)
    got = skg.DeviceArray.from_host(ctx, C0)
    ctx.dsyrk(la, N, K, alpha, dA.ptr, lda, beta, ref.ptr, N)
    ctx.dsyrk_noalloc(la, N, K, alpha, dA.ptr, lda, beta, got.ptr, N)
    ctx.synchronize()
    r = ref.download().reshape(N, N)
    g = got.download().reshape(N, N)
    for x in (dA, ref, got):
        x.free()
    il, iu = np.tril_indices(N), np.triu_indices(N, 1)
    return rel_fro(g[il], r[il]), np.array_equal(g[iu], C0.reshape(N, N)[iu])


def test_noalloc_two_calls_match_dsyrk(skg):
    ctx = skg.Context(0)
    try:
        rng = np.random.default_rng(5)
        # c2 overlap: Psi 32768 x 520 column-major (its plan has K chunks,
        # whose sums then live in the scratch array too)
        Nd, Ns = 32768, 520
        psi = (rng.standard_normal((Ns, Nd)) / np.sqrt(Nd)).ravel()
        plan = ctx.query_plan(1, Ns, Ns, Nd, skg.COL_MAJOR)
        print("c2 plan", plan, flush=True)
        C0 = rng.uniform(-1, 1, Ns * Ns)
        err, upper = run_pair(skg, ctx, skg.COL_MAJOR, Ns, Nd, psi, Nd, 1.0,
                              0.0, C0)
        assert err < 1e-13 and upper
        # a smaller, ragged second call with alpha and beta
        N, K = 137, 531
        A = rng.uniform(-1, 1, (K, N + 3))
        C0 = rng.uniform(-1, 1, N * N)
        err, upper = run_pair(skg, ctx, skg.ROW_MAJOR, N, K, A.ravel(), N + 3,
                              0.5, -1.25, C0)
        assert err < 1e-13 and upper
    finally:
        ctx.close()


def test_noalloc_rejects_beyond_limit(skg):
    ctx = skg.Context(0)
    try:
        rng = np.random.default_rng(6)
        N, K = 2048, 4096
        A = rng.uniform(-1, 1, (K, N)).ravel()
        dA = skg.DeviceArray.from_host(ctx, A)
        dC = skg.DeviceArray(ctx, N * N, zero=True)
        ctx.set_kchunks(16)    # 16 x 2048^2 doubles of chunk sums: 512 MiB
        with pytest.raises(skg.SkgemmError) as e:
            ctx.dsyrk_noalloc(skg.ROW_MAJOR, N, K, 1.0, dA.ptr, N, 0.0,
                              dC.ptr, N)
        assert e.value.code == skg.ERR_NOMEM and "limit" in str(e.value)
        ctx.set_kchunks(1)
        C0 = np.zeros(N * N)
        dA.free()
        dC.free()
        err, upper = run_pair(skg, ctx, skg.ROW_MAJOR, N, K, A, N, 1.0, 0.0,
                              C0)
        assert err < 1e-13 and upper
    finally:
        ctx.close()
