"""The wait-free (last-arriver) Stream-K fix-up on the GPU.  On the same
plan it gives the same bits as the waiting fix-up (the partials are added
in the same order): single launches with split tiles, strip tiles, K
chunks, edge tiles and SYRK diagonal blocks, and batched launches.  And it
stays exact when launches on two streams share the device, each with the
whole device's worth of workgroups, which the waiting mode must not do
(its owners wait for peers that need not be resident)."""
import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu

# (kind, M, N, K, tile, grid, kchunks): grids that do not divide the
# iteration space split tiles between workgroups.
CASES = [
    (0, 520, 520, 5000, 13, 7, 1),      # column strip, split tiles
    (0, 520, 520, 5000, 11, 37, 2),     # column strip, K chunks
    (0, 300, 200, 3000, 8, 11, 1),      # edge tiles
    (0, 1024, 1024, 1024, 8, 0, 1),     # the full default grid
    (0, 256, 256, 20000, 15, 5, 1),     # long K over few workgroups
    (1, 520, 520, 20000, 12, 13, 3),    # SYRK row strip and corner
    (1, 512, 512, 8192, 8, 9, 1),       # SYRK, diagonal tiles split
    (1, 1000, 1000, 4000, 10, 0, 4),    # SYRK edges, K chunks, full grid
]


def _run(skg, ctx, dA, dB, dD, kind, M, N, K):
    if kind == 0:
        ctx.dgemm(skg.ROW_MAJOR, skg.ROW_MAJOR, skg.ROW_MAJOR, M, N, K, 1.25,
                  dA.ptr, K, dB.ptr, N, 0.5, dD.ptr, N, 0.25)
    else:
        ctx.dsyrk(skg.COL_MAJOR, N, K, 1.0, dA.ptr, K, 0.0, dD.ptr, N)
    ctx.synchronize()
    return dD.download()


@pytest.mark.parametrize("case", CASES, ids=[str(c) for c in CASES])
def test_waitfree_bits_match_waiting(skg, ctx, orc, rng, case):
    kind, M, N, K, tile, grid, kc = case
    print("case", case, flush=True)
    A = rng.uniform(-1, 1, (M if kind == 0 else N) * K)
    B = rng.uniform(-1, 1, K * N)
    D0 = rng.uniform(-1, 1, M * N)
    dA = skg.DeviceArray.from_host(ctx, A)
    dB = skg.DeviceArray.from_host(ctx, B)
    dD = skg.DeviceArray(ctx, M * N)
    ctx.set_tuning(tile, grid)
    ctx.set_kchunks(kc)
    try:
        out = {}
        for wf in (False, True, False, True):
            ctx.set_waitfree(wf)
            dD.upload(D0)
            out.setdefault(wf, []).append(_run(skg, ctx, dA, dB, dD, kind,
                                               M, N, K))
    finally:
        ctx.set_waitfree(False)
        ctx.set_tuning(-1, 0)
        ctx.set_kchunks(0)
    assert np.array_equal(out[False][0], out[False][1])
    assert np.array_equal(out[True][0], out[True][1])
    assert np.array_equal(out[True][0], out[False][0])
    # and it is the right answer
    ref = D0.copy()
    if kind == 0:
        orc.dgemm(0, 0, 0, M, N, K, 1.25, A, K, B, N, 0.5, ref, N, 0.25,
                  dd=True)
        assert rel_fro(out[True][0], ref) < 1e-13
    else:
        full = A.reshape(N, K) @ A.reshape(N, K).T
        low = np.tril_indices(N)
        got = out[True][0].reshape(N, N)
        assert rel_fro(got[low], full[low]) < 1e-13


def test_waitfree_batched_bits_match_waiting(skg, ctx, rng):
    probs, bufs = [], []
    shapes = [(1, 264, 264, 9000), (0, 1000, 264, 264), (1, 128, 128, 30000),
              (0, 640, 520, 2000)]
    for kind, M, N, K in shapes:
        A = skg.DeviceArray.from_host(ctx, rng.uniform(-1, 1, (M if kind == 0
                                                               else N) * K))
        B = skg.DeviceArray.from_host(ctx, rng.uniform(-1, 1, K * N))
        D = skg.DeviceArray(ctx, M * N, zero=True)
        bufs.append((A, B, D))
        if kind == 0:
            probs.append(skg.Problem(0, skg.ROW_MAJOR, skg.ROW_MAJOR,
                                     skg.ROW_MAJOR, M, N, K, 1.0, 0.0, 0.0,
                                     A.ptr, K, B.ptr, N, D.ptr, N))
        else:
            probs.append(skg.Problem(1, skg.COL_MAJOR, skg.COL_MAJOR,
                                     skg.ROW_MAJOR, N, N, K, 1.0, 0.0, 0.0,
                                     A.ptr, K, A.ptr, K, D.ptr, N))
    res = {}
    try:
        for grid in (0, 19):
            ctx.set_tuning(-1, grid)
            for wf in (False, True):
                ctx.set_waitfree(wf)
                for graph in (0, 1):
                    ctx.batched(probs, use_graph=graph)
                    ctx.synchronize()
                    res[(grid, wf, graph)] = [b[2].download() for b in bufs]
    finally:
        ctx.set_waitfree(False)
        ctx.set_tuning(-1, 0)
    for grid in (0, 19):
        base = res[(grid, False, 0)]
        for key in ((grid, True, 0), (grid, False, 1), (grid, True, 1)):
            for a, b in zip(base, res[key]):
                assert np.array_equal(a, b), key


def test_waitfree_two_streams_share_the_device(skg, rng):
    """Two contexts (two streams), each launching with the whole device's
    worth of workgroups at once, many times over: every result is the one
    the context gives alone."""
    M = N = 1024
    K = 8192
    c1, c2 = skg.Context(0), skg.Context(0)
    try:
        A = rng.uniform(-1, 1, M * K)
        B = rng.uniform(-1, 1, K * N)
        arrs = []
        for c in (c1, c2):
            c.set_waitfree(True)
            c.set_tuning(8, 0)
            c.set_kchunks(1)
            arrs.append((skg.DeviceArray.from_host(c, A),
                         skg.DeviceArray.from_host(c, B),
                         skg.DeviceArray(c, M * N, zero=True),
                         skg.DeviceArray(c, N * N, zero=True)))
        plan = c1.query_plan(0, M, N, K, skg.ROW_MAJOR, skg.ROW_MAJOR)
        info = c1.device_info()
        assert plan["grid"] >= info["num_cus"]       # the whole device

        def launch(c, a):
            dA, dB, dD, dS = a
            c.dgemm(skg.ROW_MAJOR, skg.ROW_MAJOR, skg.ROW_MAJOR, M, N, K, 1.0,
                    dA.ptr, K, dB.ptr, N, 0.0, dD.ptr, N, 0.0)
            c.dsyrk(skg.COL_MAJOR, N, K, 1.0, dB.ptr, K, 0.0, dS.ptr, N)

        launch(c1, arrs[0])
        c1.synchronize()
        want_d, want_s = arrs[0][2].download(), arrs[0][3].download()
        for rep in range(10):
            launch(c1, arrs[0])
            launch(c2, arrs[1])
            c1.synchronize()
            c2.synchronize()
            for a in arrs:
                assert np.array_equal(a[2].download(), want_d), rep
                assert np.array_equal(a[3].download(), want_s), rep
    finally:
        c1.close()
        c2.close()
