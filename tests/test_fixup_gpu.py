"""The Stream-K fix-up's failure path on the GPU: an owner whose bounded
wait for a peer's partial runs out gives up instead of hanging, every
workgroup exits, the call reports SKGEMM_ERR_TIMEOUT at the next
synchronisation, and the context's next call is exact (the host zeroes the
status word and the flags on a timeout).  The condition is forced from the
host on a tile split over two workgroups: the hold hook keeps the peer's
ready flag down (a flag the host never lets be raised) and clamps the spin
limit to 64 sleep rounds, so the owner times out on every held launch; no
kernel is made to fault or hang."""
import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu


def test_spin_limit_timeout_then_exact(skg, ctx, orc, rng):
    M = N = 128
    K = 16384
    A = rng.uniform(-1, 1, (M, K))
    B = rng.uniform(-1, 1, (K, N))
    ref = np.zeros(M * N)
    orc.dgemm(0, 0, 0, M, N, K, 1.0, A.ravel().copy(), K, B.ravel().copy(), N,
              0.0, ref, N, 0.0, dd=True)
    dA = skg.DeviceArray.from_host(ctx, A)
    dB = skg.DeviceArray.from_host(ctx, B)
    dD = skg.DeviceArray(ctx, M * N, zero=True)
    # One 128 x 128 tile (4 LDS-DMA stages) split over two workgroups:
    # workgroup 0 owns the tile's first half of K and waits for workgroup
    # 1's partial of the second half.
    ctx.set_waitfree(False)
    ctx.set_tuning(8, 2)
    ctx.set_kchunks(1)
    timeouts = 0
    try:
        plan = ctx.query_plan(0, M, N, K, skg.ROW_MAJOR, skg.ROW_MAJOR)
        assert plan["tile"] == 8 and plan["grid"] == 2 and plan["tiles"] == 1
        for attempt in range(2):
            print("held launch", attempt, flush=True)
            ctx.set_fixup_hold(True)
            ctx.dgemm(skg.ROW_MAJOR, skg.ROW_MAJOR, skg.ROW_MAJOR, M, N, K,
                      1.0, dA.ptr, K, dB.ptr, N, 0.0, dD.ptr, N, 0.0)
            try:
                ctx.synchronize()
            except skg.SkgemmError as e:
                assert e.code == skg.ERR_TIMEOUT, e
                timeouts += 1
            ctx.set_fixup_hold(False)
            for _ in range(2):
                ctx.dgemm(skg.ROW_MAJOR, skg.ROW_MAJOR, skg.ROW_MAJOR, M, N,
                          K, 1.0, dA.ptr, K, dB.ptr, N, 0.0, dD.ptr, N, 0.0)
                ctx.synchronize()
                assert rel_fro(dD.download(), ref) < 1e-13
    finally:
        ctx.set_fixup_hold(False)
        ctx.set_tuning(-1, 0)
        ctx.set_kchunks(0)
    assert timeouts == 2, "a held launch did not report a timeout"
