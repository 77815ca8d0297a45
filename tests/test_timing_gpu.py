"""Per-launch timing through the C ABI (skgemm_timing_enable / _read,
skgemm_last_kernel_ms): one event pair per launch, from the ring while it
has room and from the context's own pair after that."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_ring_and_last_launch(skg, ctx):
    M = N = K = 256
    A = skg.DeviceArray(ctx, M * K)
    B = skg.DeviceArray(ctx, K * N)
    D = skg.DeviceArray(ctx, M * N, zero=True)
    ctx.fill_uniform(A.ptr, M * K, 1, -1, 1)
    ctx.fill_uniform(B.ptr, K * N, 2, -1, 1)

    def run():
        ctx.dgemm(skg.ROW_MAJOR, skg.ROW_MAJOR, skg.ROW_MAJOR, M, N, K, 1.0,
                  A.ptr, K, B.ptr, N, 0.0, D.ptr, N, 0.0)

    ctx.timing_enable(3)
    with pytest.raises(skg.SkgemmError):        # nothing timed since enable
        ctx.last_kernel_ms()
    run()
    ctx.synchronize()
    first = ctx.last_kernel_ms()                 # from the ring's first pair
    assert 0.0 < first < 1000.0
    for _ in range(4):                           # two past the ring's end
        run()
    ctx.synchronize()
    last = ctx.last_kernel_ms()                  # from the context's pair
    assert 0.0 < last < 1000.0
    ms = ctx.timing_read(10)
    assert len(ms) == 3 and np.all(np.asarray(ms) > 0.0)
    run()                                        # the ring restarts
    ctx.synchronize()
    assert len(ctx.timing_read(10)) == 1
    assert 0.0 < ctx.last_kernel_ms() < 1000.0
