"""GPU tests of the continuous-batching request loop (skgemm_server_*):
each result is the same alone and in the mix (plain FP64: within 1e-13
relative Frobenius, since the Stream-K partition depends on the batch;
split form: bit-identical, its integer steps do not depend on it), results
match the oracle, admission by HBM footprint, statistics and draining."""
import threading

import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu


def make_systems(skg, ctx, shapes, seed):
    systems = []
    for i, (Nd, Ns) in enumerate(shapes):
        psi = skg.DeviceArray(ctx, Nd * Ns)
        ctx.fill_normal(psi.ptr, Nd * Ns, seed + i, 1.0 / np.sqrt(Nd))
        q = skg.DeviceArray(ctx, Ns * Ns)
        ctx.fill_uniform(q.ptr, Ns * Ns, seed + 100 + i, -1.0, 1.0)
        systems.append((psi, q, Nd, Ns))
    ctx.synchronize()
    return systems


SHAPES = [(4096, 128), (8192, 200), (4096, 264), (13824, 136), (8192, 96)]


def run_requests(skg, ctx, srv, ids, systems, reqs, threads):
    """reqs: list of (system index, kind).  Submitted from `threads` Python
    threads at once; returns the output arrays and latencies."""
    outs = []
    for si, kind in reqs:
        psi, q, Nd, Ns = systems[si]
        outs.append(skg.DeviceArray(ctx, Ns * Ns if kind == 0 else Nd * Ns,
                                    zero=True))
    tickets = [None] * len(reqs)
    lat = [None] * len(reqs)

    def worker(k):
        mine = list(range(k, len(reqs), threads))
        for j in mine:
            si, kind = reqs[j]
            psi, q, Nd, Ns = systems[si]
            tickets[j] = srv.submit(ids[si], kind, outs[j].ptr,
                                    q.ptr if kind == 1 else None)
        for j in mine:
            lat[j] = srv.wait(tickets[j])

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return outs, lat


def host_result(out, kind, Nd, Ns):
    if kind == 0:
        return out.download().reshape(Ns, Ns)[np.tril_indices(Ns)]
    return out.download()


@pytest.mark.parametrize("precision", [0, 1])
def test_same_alone_and_in_mix(skg, ctx, orc, precision):
    systems = make_systems(skg, ctx, SHAPES, 60 + precision)
    srv = skg.Server(0, hbm_budget=8 << 30, max_batch=6, max_inflight=2,
                     precision=precision)
    try:
        ids = [srv.add_system(s[0].ptr, s[2], s[3]) for s in systems]
        alone = []
        for si in range(len(systems)):
            for kind in (0, 1):
                outs, _ = run_requests(skg, ctx, srv, ids, systems,
                                       [(si, kind)], 1)
                alone.append(host_result(outs[0], kind, systems[si][2],
                                         systems[si][3]))
        rng = np.random.default_rng(3)
        mix = [(int(rng.integers(len(systems))), int(rng.integers(2)))
               for _ in range(40)]
        outs, lat = run_requests(skg, ctx, srv, ids, systems, mix, 4)
        st = srv.stats()
    finally:
        srv.close()
    for (si, kind), out in zip(mix, outs):
        got = host_result(out, kind, systems[si][2], systems[si][3])
        want = alone[2 * si + kind]
        if precision == 1:
            assert np.array_equal(got, want), (si, kind)
        else:
            assert rel_fro(got, want) < 1e-13, (si, kind)
    assert all(x > 0 for x in lat)
    assert st["completed"] == 10 + 40
    assert 1 <= st["peak_inflight"] <= 2
    assert st["launches"] <= st["completed"]


def test_results_match_oracle(skg, ctx, orc):
    systems = make_systems(skg, ctx, [(4096, 72), (2048, 40)], 80)
    srv = skg.Server(0, hbm_budget=1 << 30, max_batch=4, max_inflight=2)
    try:
        ids = [srv.add_system(s[0].ptr, s[2], s[3]) for s in systems]
        reqs = [(0, 0), (0, 1), (1, 0), (1, 1)]
        outs, _ = run_requests(skg, ctx, srv, ids, systems, reqs, 2)
    finally:
        srv.close()
    for (si, kind), out in zip(reqs, outs):
        psi, q, Nd, Ns = systems[si]
        P = psi.download().reshape(Ns, Nd)           # row b = column b of Psi
        if kind == 0:
            ref = np.zeros(Ns * Ns)
            orc.dgemm(0, 1, 0, Ns, Ns, Nd, 1.0, P.ravel().copy(), Nd,
                      P.ravel().copy(), Nd, 0.0, ref, Ns, 0.0, dd=True)
            il = np.tril_indices(Ns)
            got = out.download().reshape(Ns, Ns)[il]
            assert rel_fro(got, ref.reshape(Ns, Ns)[il]) < 1e-13
        else:
            Q = q.download()
            ref = np.zeros(Nd * Ns)
            orc.dgemm(1, 0, 1, Nd, Ns, Ns, 1.0, P.ravel().copy(), Nd, Q, Ns,
                      0.0, ref, Nd, 0.0, dd=True)
            assert rel_fro(out.download(), ref) < 1e-13


def test_admission_by_footprint(skg, ctx):
    systems = make_systems(skg, ctx, [(4096, 128)], 90)
    psi, q, Nd, Ns = systems[0]
    srv = skg.Server(0, hbm_budget=Nd * Ns * 8 + 1024, max_batch=4,
                     max_inflight=1, precision=1)
    try:
        sid = srv.add_system(psi.ptr, Nd, Ns)
        with pytest.raises(skg.SkgemmError) as e:    # pool over budget
            srv.add_system(psi.ptr, Nd, Ns)
        assert e.value.code == skg.ERR_NOMEM
        out = skg.DeviceArray(ctx, Ns * Ns)
        with pytest.raises(skg.SkgemmError) as e:    # split workspace
            srv.submit(sid, 0, out.ptr)
        assert e.value.code == skg.ERR_NOMEM
        with pytest.raises(skg.SkgemmError) as e:
            srv.submit(sid + 5, 0, out.ptr)
        assert e.value.code == skg.ERR_ARG
        with pytest.raises(skg.SkgemmError) as e:    # host output buffer
            srv.submit(sid, 0, np.zeros(Ns * Ns).ctypes.data)
        assert e.value.code == skg.ERR_ARG
    finally:
        srv.close()


def test_close_drains_queue(skg, ctx):
    systems = make_systems(skg, ctx, [(8192, 256)], 95)
    psi, q, Nd, Ns = systems[0]
    srv = skg.Server(0, hbm_budget=4 << 30, max_batch=2, max_inflight=1)
    sid = srv.add_system(psi.ptr, Nd, Ns)
    outs = [skg.DeviceArray(ctx, Nd * Ns, zero=True) for _ in range(6)]
    for o in outs:
        srv.submit(sid, 1, o.ptr, q.ptr)
    srv.close()                       # returns after the queue ran
    first = outs[0].download()
    assert np.linalg.norm(first) > 0
    for o in outs[1:]:
        assert rel_fro(o.download(), first) < 1e-13
