"""The c3, c4 and c5 workloads at their full sizes through the C ABI, on
exactly the inputs bench.py times (stream-k-gemm_amd/workloads.py).

Bars: plain FP64 within 1e-13 relative Frobenius of the double-double
oracle on sampled rows; the split form (14 moduli) bit for bit equal to the
exact-integer restatement (oracle/ozaki.py through 128-bit C arithmetic) on
the same sampled rows, within 1e-13 of the double-double oracle and within
the bound the library returns; batched and served results equal to
stand-alone calls (bit for bit for the split form and for replays of one
plan, within 1e-13 for plain FP64 under another Stream-K partition)."""
import importlib.util
import os
import threading

import numpy as np
import pytest

from conftest import PKG_DIR, rel_fro
from oracle import ozaki

pytestmark = pytest.mark.gpu
TOL = 1e-13

_spec = importlib.util.spec_from_file_location(
    "skgemm_workloads", os.path.join(PKG_DIR, "workloads.py"))
wl = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(wl)


def near_pow2(x, L=14):
    """Squared norms whose scale code changes under a relative change of
    1e-12: the host's and the device's summation orders may give different
    codes there (the code is a step function of the norm)."""
    _, h = ozaki.tables(L)
    x = np.asarray(x, np.float64)
    c = ozaki.scale_code(x, h)
    return ((ozaki.scale_code(x * (1 + 1e-12), h) != c) |
            (ozaki.scale_code(x * (1 - 1e-12), h) != c))


def syrk_rows_dd(orc, P, rows):
    """Rows `rows` of Psi^T Psi in double-double; P is Ns x Nd (row b is
    column b of Psi)."""
    Ns, Nd = P.shape
    sub = np.ascontiguousarray(P[rows])
    ref = np.zeros(len(rows) * Ns)
    orc.dgemm(0, 1, 0, len(rows), Ns, Nd, 1.0, sub.ravel(), Nd, P.ravel(), Nd,
              0.0, ref, Ns, 0.0, dd=True)
    return ref.reshape(len(rows), Ns)


def lower_rows(C, rows):
    return np.concatenate([C[i, :r + 1] for i, r in enumerate(rows)])


def c3_rows(P, scales_seed=11, n=24):
    """Sampled columns of Psi: the 8 with the largest norms, the 8 with
    the smallest, 8 others."""
    Ns = P.shape[0]
    norms = np.einsum("ij,ij->i", P, P)
    order = np.argsort(norms)
    pick = set(order[:8]) | set(order[-8:])
    rest = [i for i in np.random.default_rng(scales_seed).permutation(Ns)
            if i not in pick]
    pick |= set(rest[:n - len(pick)])
    return np.array(sorted(pick)), norms


def test_c3_syrk_full_size(skg, ctx, orc):
    """c3 overlap Ms = Psi^T Psi, Psi 65536 x 1024 with 10^U(0,6) column
    scales: plain FP64 and split form (14 moduli) on sampled rows against
    the double-double oracle and the exact-integer restatement; diagonal
    against the column norms; the two forms against each other on the
    whole lower triangle."""
    Nd, Ns = wl.C3_ND, wl.C3_NS
    print("c3 syrk", Nd, Ns, flush=True)
    psi, q = wl.c3_operands(skg, ctx)
    plain = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    split = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    ctx.set_tuning(-1, 0)
    ctx.dsyrk(skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0, plain.ptr, Ns)
    bound = ctx.dsyrk_ozaki(skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0,
                            split.ptr, Ns, slices=14)
    ctx.synchronize()
    P = psi.download().reshape(Ns, Nd)
    Cp = plain.download().reshape(Ns, Ns)
    Cs = split.download().reshape(Ns, Ns)
    for x in (psi, q, plain, split):
        x.free()
    rows, norms = c3_rows(P)
    ref = syrk_rows_dd(orc, P, rows)
    want = lower_rows(ref, rows)
    e_plain = rel_fro(lower_rows(Cp[rows], rows), want)
    e_split = rel_fro(lower_rows(Cs[rows], rows), want)
    print("c3 syrk err plain %.3e split %.3e bound %.3e" %
          (e_plain, e_split, bound), flush=True)
    assert e_plain < TOL
    assert e_split < TOL and e_split <= bound
    # bit for bit against the exact integer product of the same rows
    ok = ~near_pow2(norms)
    exact, _, _ = ozaki.product_rows(np.ascontiguousarray(P[rows]), P, 14,
                                     norm2_a=norms[rows], norm2_b=norms)
    for i, r in enumerate(rows):
        cols = np.arange(r + 1)
        cols = cols[ok[cols]] if ok[r] else cols[:0]
        assert np.array_equal(Cs[r, cols], exact[i, cols]), r
    assert rel_fro(np.diag(Cp), norms) < 1e-14
    assert rel_fro(np.diag(Cs), norms) < 1e-14
    il = np.tril_indices(Ns)
    assert rel_fro(Cs[il], Cp[il]) < TOL


def test_c3_rotation_full_size(skg, ctx, orc):
    """c3 rotation Psi Q (65536 x 1024 x 1024), plain FP64 and split form,
    on 256 sampled grid rows against the double-double oracle; the split
    rows bit for bit against the exact-integer restatement."""
    Nd, Ns = wl.C3_ND, wl.C3_NS
    print("c3 rotation", Nd, Ns, flush=True)
    psi, q = wl.c3_operands(skg, ctx)
    rot = skg.DeviceArray(ctx, Nd * Ns, zero=True)
    rot2 = skg.DeviceArray(ctx, Nd * Ns, zero=True)
    ctx.dgemm(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, Nd, Ns, Ns, 1.0,
              psi.ptr, Nd, q.ptr, Ns, 0.0, rot.ptr, Nd, 0.0)
    bound = ctx.dgemm_ozaki(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, Nd,
                            Ns, Ns, 1.0, psi.ptr, Nd, q.ptr, Ns, 0.0,
                            rot2.ptr, Nd, 0.0, slices=14)
    ctx.synchronize()
    P = psi.download().reshape(Ns, Nd)
    Q = q.download().reshape(Ns, Ns)
    Rp = rot.download().reshape(Ns, Nd)       # column j of the output = row j
    Rs = rot2.download().reshape(Ns, Nd)
    for x in (psi, q, rot, rot2):
        x.free()
    g = np.sort(np.random.default_rng(12).choice(Nd, 256, replace=False))
    A = np.ascontiguousarray(P[:, g].T)       # Psi rows g, 256 x Ns
    ref = np.zeros(256 * Ns)
    orc.dgemm(0, 0, 0, 256, Ns, Ns, 1.0, A.ravel(), Ns, Q.ravel().copy(), Ns,
              0.0, ref, Ns, 0.0, dd=True)
    ref = ref.reshape(256, Ns)
    e_plain = rel_fro(Rp[:, g].T, ref)
    e_split = rel_fro(Rs[:, g].T, ref)
    print("c3 rotation err plain %.3e split %.3e bound %.3e" %
          (e_plain, e_split, bound), flush=True)
    assert e_plain < TOL
    assert e_split < TOL and e_split <= bound
    na = np.einsum("ij,ij->i", A, A)
    Qt = np.ascontiguousarray(Q.T)            # rows = columns of Q
    nb = np.einsum("ij,ij->i", Qt, Qt)
    exact, _, _ = ozaki.product_rows(A, Qt, 14, norm2_a=na, norm2_b=nb)
    okr, okc = ~near_pow2(na), ~near_pow2(nb)
    got = Rs[:, g].T
    assert np.array_equal(got[okr][:, okc], exact[okr][:, okc])


def test_c4_batch_full_size(skg, ctx, orc):
    """The c4 batch (64 systems, 128 problems, one persistent launch): the
    plain FP64 batch replayed from a captured graph equals the eager launch
    bit for bit and every problem's sampled rows are within 1e-13 of the
    double-double oracle; the split-form batch equals each problem's
    stand-alone split call bit for bit and is within 1e-13 of the oracle."""
    print("c4 batch", flush=True)
    probs, keep, flops = wl.c4_batch(skg, ctx)
    outs = [keep[4 * i + 2] if k == 0 else keep[4 * i + 3]
            for i in range(len(probs) // 2) for k in (0, 1)]
    ctx.set_tuning(-1, 0)

    def run(prec, graph):
        for o in outs:
            ctx.fill_uniform(o.ptr, o.n, 99, -1.0, 1.0)   # stale contents
        ctx.batched(probs, precision=prec, use_graph=graph)
        ctx.synchronize()
        return [o.download() for o in outs]

    eager = run(skg.FP64, 0)
    graph = run(skg.FP64, 1)
    graph2 = run(skg.FP64, 1)                 # a second replay
    split = run(skg.OZAKI_INT8, 1)
    rng = np.random.default_rng(13)
    for i in range(len(probs) // 2):
        psi, q = keep[4 * i], keep[4 * i + 1]
        Ns, Nd = probs[2 * i].N, probs[2 * i].K
        print("c4 system", i, Nd, Ns, flush=True)
        # SYRK (lower triangle, row-major)
        for a, b in ((eager, graph), (graph, graph2)):
            il = np.tril_indices(Ns)
            assert np.array_equal(a[2 * i].reshape(Ns, Ns)[il],
                                  b[2 * i].reshape(Ns, Ns)[il]), i
            assert np.array_equal(a[2 * i + 1], b[2 * i + 1]), i
        P = psi.download().reshape(Ns, Nd)
        rows = np.sort(rng.choice(Ns, 3, replace=False))
        want = lower_rows(syrk_rows_dd(orc, P, rows), rows)
        for C in (graph[2 * i], split[2 * i]):
            assert rel_fro(lower_rows(C.reshape(Ns, Ns)[rows], rows),
                           want) < TOL, i
        # rotation (column-major Nd x Ns)
        Q = q.download()
        g = np.sort(rng.choice(Nd, 4, replace=False))
        A = np.ascontiguousarray(P[:, g].T)
        ref = np.zeros(4 * Ns)
        orc.dgemm(0, 0, 0, 4, Ns, Ns, 1.0, A.ravel(), Ns, Q, Ns, 0.0, ref, Ns,
                  0.0, dd=True)
        for R in (graph[2 * i + 1], split[2 * i + 1]):
            assert rel_fro(R.reshape(Ns, Nd)[:, g].T.ravel(), ref) < TOL, i
        # the split batch equals the stand-alone split calls
        ms = skg.DeviceArray(ctx, Ns * Ns, zero=True)
        ctx.dsyrk_ozaki(skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0, ms.ptr,
                        Ns, slices=0)
        il = np.tril_indices(Ns)
        assert np.array_equal(ms.download().reshape(Ns, Ns)[il],
                              split[2 * i].reshape(Ns, Ns)[il]), i
        ms.free()
        rt = skg.DeviceArray(ctx, Nd * Ns, zero=True)
        ctx.dgemm_ozaki(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, Nd, Ns,
                        Ns, 1.0, psi.ptr, Nd, q.ptr, Ns, 0.0, rt.ptr, Nd, 0.0,
                        slices=0)
        assert np.array_equal(rt.download(), split[2 * i + 1]), i
        rt.free()
    for a in keep:
        a.free()


# c5 systems for the serving test: the five largest Psi, the three largest
# Ns, the three smallest Psi of the 200, and five others (indices into the
# c5 list, so each keeps its c5 seed).
C5_PICK = [13, 113, 117, 132, 53, 191, 64, 189, 127, 26, 99, 0, 1, 2, 3, 4]


@pytest.mark.parametrize("precision", [0, 1])
def test_c5_serving_full_size(skg, ctx, orc, precision):
    """A server over c5 systems (Ns up to 2044, Nd up to 65306, 8.6 GB of
    resident Psi): 48 overlap / rotation requests submitted from four
    threads at once (several launches, two in flight) give the same result
    as each request served alone (split form: bit for bit), and every
    result is within 1e-13 of the double-double oracle on sampled rows."""
    print("c5 serving precision", precision, flush=True)
    systems = wl.c5_systems(skg, ctx, pick=C5_PICK)
    pool = sum(s[2] * s[3] * 8 for s in systems)
    srv = skg.Server(0, hbm_budget=pool + (24 << 30), max_batch=16,
                     max_inflight=2, precision=precision)
    rng = np.random.default_rng(14 + precision)
    reqs = [(int(rng.integers(len(systems))), int(rng.integers(2)))
            for _ in range(48)]
    try:
        ids = [srv.add_system(s[0].ptr, s[2], s[3]) for s in systems]
        outs = [skg.DeviceArray(ctx, systems[si][3] ** 2 if kind == 0 else
                                systems[si][2] * systems[si][3], zero=True)
                for si, kind in reqs]
        tickets = [None] * len(reqs)

        def worker(k):
            mine = range(k, len(reqs), 4)
            for j in mine:
                si, kind = reqs[j]
                tickets[j] = srv.submit(ids[si], kind, outs[j].ptr,
                                        systems[si][1].ptr if kind else None)
            for j in mine:
                srv.wait(tickets[j])

        ts = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["launches"] >= 3
        mixed = [o.download() for o in outs]
        for o in outs:
            o.free()
        # each distinct request alone, then against the oracle
        for (si, kind) in sorted(set(reqs)):
            psi, q, Nd, Ns = systems[si]
            print("c5 system", C5_PICK[si], Nd, Ns, kind, flush=True)
            o = skg.DeviceArray(ctx, Ns * Ns if kind == 0 else Nd * Ns,
                                zero=True)
            srv.wait(srv.submit(ids[si], kind, o.ptr, q.ptr if kind else None))
            alone = o.download()
            o.free()
            il = np.tril_indices(Ns)
            for j, r in enumerate(reqs):
                if r != (si, kind):
                    continue
                if kind == 0:
                    a, b = mixed[j].reshape(Ns, Ns)[il], alone.reshape(Ns, Ns)[il]
                else:
                    a, b = mixed[j], alone
                if precision == 1:
                    assert np.array_equal(a, b), (si, kind)
                else:
                    assert rel_fro(a, b) < TOL, (si, kind)
            P = psi.download().reshape(Ns, Nd)
            if kind == 0:
                rows = np.sort(rng.choice(Ns, 2, replace=False))
                want = lower_rows(syrk_rows_dd(orc, P, rows), rows)
                got = lower_rows(alone.reshape(Ns, Ns)[rows], rows)
            else:
                g = np.sort(rng.choice(Nd, 4, replace=False))
                A = np.ascontiguousarray(P[:, g].T)
                want = np.zeros(4 * Ns)
                orc.dgemm(0, 0, 0, 4, Ns, Ns, 1.0, A.ravel(), Ns,
                          q.download(), Ns, 0.0, want, Ns, 0.0, dd=True)
                got = alone.reshape(Ns, Nd)[:, g].T.ravel()
            assert rel_fro(got, want) < TOL, (si, kind)
    finally:
        srv.close()
        for s in systems:
            s[0].free()
            s[1].free()
