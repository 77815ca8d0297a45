"""GPU tests of skgemm_batched: many GEMM / SYRK problems of uneven sizes in
one persistent launch (plain FP64) or one captured split-form sequence.
Bars: 1e-13 relative Frobenius against the double-double oracle for plain
FP64 (each problem); bit-identical to the single-problem split-form call
for the Ozaki path (its integer steps do not depend on the partition);
graph replays identical to eager launches."""
import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-13


def make_batch(skg, ctx, rng, specs, beta=0.0):
    """specs: list of (kind, M, N, K).  GEMM: A col-major M x K, B row-major
    K x N, D col-major M x N.  SYRK: A col-major K x N, C row-major N x N.
    Returns (problems, host inputs, device arrays)."""
    probs, host, keep = [], [], []
    for kind, M, N, K in specs:
        if kind == 0:
            A = rng.uniform(-1, 1, (M, K))
            B = rng.uniform(-1, 1, (K, N))
            D0 = rng.uniform(-1, 1, (M, N))
            dA = skg.DeviceArray.from_host(ctx, np.asfortranarray(A).ravel(order="F"))
            dB = skg.DeviceArray.from_host(ctx, np.ascontiguousarray(B).ravel())
            dD = skg.DeviceArray.from_host(ctx, np.asfortranarray(D0).ravel(order="F"))
            p = skg.Problem(kind=0, layout_a=skg.COL_MAJOR, layout_b=skg.ROW_MAJOR,
                            layout_d=skg.COL_MAJOR, M=M, N=N, K=K, alpha=1.5,
                            beta=beta, gamma=0.25, A=dA.ptr, lda=M, B=dB.ptr,
                            ldb=N, D=dD.ptr, ldd=M)
            host.append((A, B, D0))
        else:
            A = rng.standard_normal((K, N)) / np.sqrt(K)
            C0 = rng.uniform(-1, 1, (N, N))
            dA = skg.DeviceArray.from_host(ctx, np.asfortranarray(A).ravel(order="F"))
            dD = skg.DeviceArray.from_host(ctx, np.ascontiguousarray(C0).ravel())
            p = skg.Problem(kind=1, layout_a=skg.COL_MAJOR, layout_b=0,
                            layout_d=0, M=N, N=N, K=K, alpha=1.0, beta=beta,
                            gamma=0.0, A=dA.ptr, lda=K, B=None, ldb=0,
                            D=dD.ptr, ldd=N)
            dB = None
            host.append((A, None, C0))
        probs.append(p)
        keep.append((dA, dB, dD))
    return probs, host, keep


def reference(orc, spec, h, beta):
    kind, M, N, K = spec
    if kind == 0:
        A, B, D0 = h
        ref = np.zeros(M * N)
        orc.dgemm(0, 0, 0, M, N, K, 1.5, np.ascontiguousarray(A.ravel()), K,
                  np.ascontiguousarray(B.ravel()), N, 0.0, ref, N, 0.0, dd=True)
        return ref.reshape(M, N) + beta * D0 + 0.25
    A, _, C0 = h
    ref = np.zeros(N * N)
    orc.dsyrk(0, N, K, 1.0, np.ascontiguousarray(A.ravel()), N, 0.0, ref, N,
              dd=True)
    return ref.reshape(N, N) + beta * C0


def result(spec, dev):
    kind, M, N, K = spec
    if kind == 0:
        return dev[2].download().reshape(N, M).T
    return dev[2].download().reshape(N, N)


def check_all(orc, specs, host, keep, beta):
    for spec, h, dev in zip(specs, host, keep):
        got = result(spec, dev)
        want = reference(orc, spec, h, beta)
        if spec[0] == 1:
            il = np.tril_indices(spec[2])
            got, want = got[il], want[il]
        assert rel_fro(got, want) < TOL, spec


SPECS = [(0, 300, 136, 136), (1, 0, 136, 4096), (0, 1000, 520, 520),
         (1, 0, 520, 2000), (0, 64, 64, 1), (1, 0, 128, 13824),
         (0, 4096, 256, 256), (1, 0, 40, 100), (0, 17, 33, 5000)]


@pytest.mark.parametrize("use_graph", [0, 1])
def test_plain_batch_matches_oracle(skg, ctx, orc, rng, use_graph):
    probs, host, keep = make_batch(skg, ctx, rng, SPECS, beta=-0.5)
    ctx.batched(probs, precision=skg.FP64, use_graph=use_graph)
    ctx.synchronize()
    check_all(orc, SPECS, host, keep, -0.5)


def test_graph_replays_new_contents(skg, ctx, orc, rng):
    """One captured graph serves batches of different sizes and counts: the
    problem table is refreshed in device memory before every replay."""
    for specs in (SPECS, SPECS[::-1][:5], [(1, 0, 700, 3000)], SPECS):
        probs, host, keep = make_batch(skg, ctx, rng, specs)
        ctx.batched(probs, precision=skg.FP64, use_graph=1)
        ctx.synchronize()
        check_all(orc, specs, host, keep, 0.0)


def test_graph_replay_accumulates_like_eager(skg, ctx, rng):
    """beta = 1: three replays give D0 + 3 A B exactly as three eager
    launches do (same partition, same order), with host work in between."""
    specs = [(0, 500, 200, 300), (1, 0, 300, 2000)]
    probs, host, keep = make_batch(skg, ctx, rng, specs, beta=1.0)
    probs2, _, keep2 = make_batch(skg, ctx, np.random.default_rng(7), specs,
                                  beta=1.0)
    for (dA, dB, dD), (eA, eB, eD) in zip(keep, keep2):
        eA.upload(dA.download())
        if dB is not None:
            eB.upload(dB.download())
        eD.upload(dD.download())
    for _ in range(3):
        ctx.batched(probs, precision=skg.FP64, use_graph=1)
        ctx.synchronize()
        _ = np.linalg.norm(np.ones(1000))        # host work between replays
        ctx.batched(probs2, precision=skg.FP64, use_graph=0)
        ctx.synchronize()
    for a, b in zip(keep, keep2):
        assert np.array_equal(a[2].download(), b[2].download())


def test_split_batch_bit_identical_to_single_calls(skg, ctx, rng):
    specs = [(0, 300, 136, 136), (1, 0, 136, 4096), (1, 0, 520, 2000),
             (0, 2000, 256, 256)]
    probs, host, keep = make_batch(skg, ctx, rng, specs)
    singles = []
    for p, (kind, M, N, K), (dA, dB, dD) in zip(probs, specs, keep):
        out = skg.DeviceArray(ctx, dD.n)
        out.upload(dD.download())
        if kind == 0:
            ctx.dgemm_ozaki(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, M, N,
                            K, 1.5, dA.ptr, M, dB.ptr, N, 0.0, out.ptr, M, 0.25)
        else:
            ctx.dsyrk_ozaki(skg.COL_MAJOR, N, K, 1.0, dA.ptr, K, 0.0, out.ptr, N)
        singles.append(out)
    ctx.synchronize()
    for use_graph in (1, 1, 0):
        ctx.batched(probs, precision=skg.OZAKI_INT8, use_graph=use_graph)
        ctx.synchronize()
        for (kind, M, N, K), dev, s in zip(specs, keep, singles):
            got, want = dev[2].download(), s.download()
            if kind == 1:
                il = np.tril_indices(N)
                got, want = got.reshape(N, N)[il], want.reshape(N, N)[il]
            assert np.array_equal(got, want), (kind, M, N, K, use_graph)


def test_batch_argument_errors(skg, ctx, rng):
    probs, _, keep = make_batch(skg, ctx, rng, [(0, 10, 10, 10)])
    bad = skg.Problem(kind=3, M=1, N=1, K=1)
    with pytest.raises(skg.SkgemmError) as e:
        ctx.batched([probs[0], bad])
    assert e.value.code == skg.ERR_ARG
    syrk_bad = skg.Problem(kind=1, layout_a=0, M=4, N=5, K=3, alpha=1.0,
                           A=keep[0][0].ptr, lda=5, D=keep[0][2].ptr, ldd=5)
    with pytest.raises(skg.SkgemmError) as e:
        ctx.batched([syrk_bad])
    assert e.value.code == skg.ERR_ARG
    host = np.zeros(100)
    p = skg.Problem(kind=0, layout_a=0, layout_b=0, layout_d=0, M=10, N=10,
                    K=10, alpha=1.0, A=host.ctypes.data, lda=10,
                    B=keep[0][1].ptr, ldb=10, D=keep[0][2].ptr, ldd=10)
    with pytest.raises(skg.SkgemmError) as e:
        ctx.batched([p])
    assert e.value.code == skg.ERR_ARG
