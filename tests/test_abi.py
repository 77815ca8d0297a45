"""C ABI checks that need no GPU: the library loads, exports exactly what
include/skgemm.h declares, headers compile, argument errors are reported
without touching a device, and the Stream-K work split (restated from the
kernel) covers every MAC iteration once with deadlock-free fix-up order."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, PKG_DIR

HEADER = os.path.join(ROOT, "include", "skgemm.h")
FACADE = os.path.join(ROOT, "include", "skgemm.hpp")
LIB = os.path.join(PKG_DIR, "libskgemm.so")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(skgemm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_binding_exports(skg):
    assert header_functions() == sorted(skg.EXPORTS)


def test_library_exports_every_declared_symbol(skg):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB],
                         capture_output=True, text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [f for f in header_functions() if f not in syms]
    assert not missing, missing
    lib = skg.lib()  # ctypes resolves every one of them
    for f in header_functions():
        assert getattr(lib, f) is not None


def test_headers_compile_as_c_and_cpp(tmp_path):
    subprocess.run(["gcc", "-fsyntax-only", "-Wall", "-Werror", "-x", "c",
                    HEADER], check=True)
    src = tmp_path / "use.cpp"
    src.write_text('#include "skgemm.hpp"\n'
                   'int main() { gemm<double>::pack_A(16, 1, nullptr, nullptr);'
                   ' return 0; }\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I",
                    os.path.join(ROOT, "include"), str(src)], check=True)


def test_errors_without_device(skg):
    lib = skg.lib()
    assert lib.skgemm_version().startswith(b"skgemm")
    rc = lib.skgemm_dgemm(None, 0, 0, 0, 4, 4, 4, 1.0, None, 4, None, 4, 0.0,
                          None, 4, 0.0)
    assert rc == skg.ERR_STATE
    assert b"context" in lib.skgemm_last_error()
    assert lib.skgemm_tile_info(-1, None, None, None, None) == skg.ERR_ARG
    assert lib.skgemm_tile_info(skg.num_tiles(), None, None, None,
                                None) == skg.ERR_ARG
    n = skg.num_tiles()
    assert n >= 4
    for t in range(n):
        info = skg.tile_info(t)
        assert info["bm"] % 64 == 0 and info["bn"] % 64 == 0
        assert info["bk"] % 4 == 0 and info["waves"] in (4, 8)


def test_batch_and_server_argument_checks_without_device(skg):
    """Argument errors are reported before any device work."""
    import ctypes
    lib = skg.lib()
    assert lib.skgemm_batched(None, None, 1, 0, 0, 0) == skg.ERR_STATE
    out = ctypes.c_void_p()
    for args in ((0, 1 << 30, 0, 1, 0, 0), (0, 1 << 30, 4, 0, 0, 0),
                 (0, 1 << 30, 4, 9, 0, 0), (0, 1 << 30, 4, 2, 5, 0),
                 (0, 1 << 30, 4, 2, 1, 3)):
        assert lib.skgemm_server_create(*args, ctypes.byref(out)) == skg.ERR_ARG
        assert out.value is None
    assert lib.skgemm_server_wait(None, 1, None) == skg.ERR_ARG
    assert lib.skgemm_server_stats(None, None, None, None, None,
                                   None) == skg.ERR_ARG


# ---- Stream-K work split (restates sk_dgemm_kernel's scheduling) ---------

def wg_start(w, total, G):
    base, extra = divmod(total, G)
    return w * base + min(w, extra)


def logical_id(g, G):
    per, rem = divmod(G, 8)
    x = g % 8
    return x * per + min(x, rem) + g // 8


def simulate(num_tiles, ipt, G):
    """Returns per-tile (owner, [peers]) and per-(tile) covered iteration
    sets, following the kernel's loop."""
    total = num_tiles * ipt
    owners, peers, cover = {}, {}, np.zeros(total, np.int64)
    for w in range(G):
        it, end = wg_start(w, total, G), wg_start(w + 1, total, G)
        first = True
        while it < end:
            tile = it // ipt
            t0 = tile * ipt
            kb, ke = it - t0, min(end - t0, ipt)
            cover[t0 + kb:t0 + ke] += 1
            if kb != 0:
                assert first, "a workgroup spills only its first tile"
                peers.setdefault(tile, []).append(w)
            elif ke < ipt:
                owners[tile] = w
            first = False
            it = t0 + ke
    return owners, peers, cover, total


@pytest.mark.parametrize("num_tiles,ipt,G", [
    (64, 64, 256), (36, 4096, 512), (15, 2048, 512), (1280, 33, 512),
    (1, 4, 3), (45, 70, 37), (20, 32, 640), (7, 1, 5), (3, 1000, 1000)])
def test_streamk_split_covers_once_and_owners_wait_forward(num_tiles, ipt, G):
    G = min(G, num_tiles * ipt)
    owners, peers, cover, total = simulate(num_tiles, ipt, G)
    assert np.all(cover == 1)
    for tile, ps in peers.items():
        w = owners[tile]
        # the owner's peer loop in the kernel: w2 = w+1 .. while start < end
        expect = [w2 for w2 in range(w + 1, G)
                  if wg_start(w2, total, G) < (tile + 1) * ipt]
        assert ps == expect
        assert all(p > w for p in ps)  # waits only on later workgroups
    assert set(owners) == set(peers)


@pytest.mark.parametrize("tiles,kchunks,kci", [(36, 14, 146), (36, 7, 293),
                                                (45, 11, 187), (10, 25, 82)])
def test_data_parallel_grid_is_unit_aligned(tiles, kchunks, kci):
    """A grid of exactly one (K chunk, tile) unit per workgroup (kept as is
    by the planner, not rounded to whole XCD rounds, and run on the even
    split): every workgroup takes exactly one whole unit, so no tile needs
    a fix-up."""
    G = tiles * kchunks
    owners, peers, cover, total = simulate(G, kci, G)
    assert np.all(cover == 1)
    assert not owners and not peers
    assert all(wg_start(w, total, G) == w * kci for w in range(G + 1))


def test_committed_traffic_figures_follow_their_counters():
    """profiles/traffic.json: bytes per launch = (FETCH_SIZE x fetch factor
    + WRITE_SIZE x write factor) KiB, each factor the one that makes the
    copy kernel read and write exactly its 1 GiB."""
    import json
    d = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))
    cal = d["calibration"]
    ff = cal["copy_kernel_bytes_each_way"] / 1024 / cal["fetch_kib_raw"]
    wf = cal["copy_kernel_bytes_each_way"] / 1024 / cal["write_kib_raw"]
    assert abs(ff - cal["fetch_factor"]) < 1e-3
    assert abs(wf - cal["write_factor"]) < 1e-3
    for key in ("rotation", "syrk"):
        e = d[key]
        want = (e["fetch_kib_raw"] * ff + e["write_kib_raw"] * wf) * 1024
        assert abs(e["bytes_per_launch"] - want) <= 1e-4 * want, key
        assert e["bytes_per_launch"] > e["algorithmic_bytes"]
    # the split-form entries carry the factors of their own run
    c3 = d["calibration_c3"]
    ff3 = c3["copy_kernel_bytes_each_way"] / 1024 / c3["fetch_kib_raw"]
    wf3 = c3["copy_kernel_bytes_each_way"] / 1024 / c3["write_kib_raw"]
    for key in ("int8_products", "residue_pass"):
        e = d[key]
        assert abs(e["fetch_factor"] - ff3) < 1e-3 and abs(e["write_factor"] - wf3) < 1e-3
        want = (e["fetch_kib_raw"] * ff3 + e["write_kib_raw"] * wf3) * 1024
        assert abs(e["bytes_per_launch"] - want) <= 1e-4 * want, key
        assert e["bytes_per_launch"] > e["algorithmic_bytes"]


@pytest.mark.parametrize("G", [1, 7, 8, 9, 37, 256, 511, 512, 1024])
def test_xcd_remap_is_a_bijection(G):
    ids = sorted(logical_id(g, G) for g in range(G))
    assert ids == list(range(G))
    # consecutive logical ids share an XCD (hardware id g -> XCD g % 8)
    xcd_of = {logical_id(g, G): g % 8 for g in range(G)}
    changes = sum(xcd_of[i] != xcd_of[i + 1] for i in range(G - 1))
    assert changes <= 7


def test_triangle_tile_map():
    """SYRK tile index t -> (ti, tj), row-major lower triangle."""
    import math
    seen = []
    for t in range(5050):
        i = int((math.sqrt(8.0 * t + 1.0) - 1.0) * 0.5)
        while i * (i + 1) // 2 > t:
            i -= 1
        while (i + 1) * (i + 2) // 2 <= t:
            i += 1
        seen.append((i, t - i * (i + 1) // 2))
    assert seen == [(i, j) for i in range(100) for j in range(i + 1)]


def test_nloc_argument_checks_without_device(skg):
    lib = skg.lib()
    h = ctypes.c_void_p()
    assert lib.skgemm_nloc_create(None, 100, 0, None, None, None, None, None,
                                  None, ctypes.byref(h)) == skg.ERR_STATE
    assert lib.skgemm_nloc_apply(None, 16, 1.0, None, None, 32) == skg.ERR_ARG
    assert lib.skgemm_nloc_destroy(None) == skg.OK


def test_stencil_argument_checks_without_device(skg):
    lib = skg.lib()
    c = (ctypes.c_double * 7)()
    assert lib.skgemm_halo_extend(None, 4, 4, 4, 16, 2, None, None,
                                  1) == skg.ERR_STATE
    assert lib.skgemm_stencil_apply(None, 4, 4, 4, 16, 2, None, None, c, c, c,
                                    0.0, None, 32) == skg.ERR_STATE


# ---- tile orders (tile_coords, shared by the kernel and the host split) ----

@pytest.fixture(scope="module")
def tile_order_exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("toc") / "tile_order_check")
    subprocess.run([os.environ.get("HIPCC", "/opt/rocm/bin/hipcc"),
                    "-std=c++17", "-O1", "--offload-arch=gfx950",
                    "-D__HIP_PLATFORM_AMD__",
                    os.path.join(ROOT, "tests", "cpp", "tile_order_check.cpp"),
                    "-o", out], check=True, capture_output=True, text=True)
    return out


def tile_coords_of(exe, syrk, order, tm, tn):
    out = subprocess.run([exe, str(syrk), str(order), str(tm), str(tn)],
                         check=True, capture_output=True, text=True).stdout
    return [tuple(int(v) for v in line.split()) for line in out.splitlines()]


@pytest.mark.parametrize("order,tm,tn", [
    (0, 1, 1), (0, 5, 3), (1, 5, 3), (1, 1, 40), (2, 2, 17), (3, 2, 17),
    (3, 1, 16), (4, 3, 16), (8, 2, 64), (5, 4, 23), (7, 6, 20), (16, 2, 16),
    (3, 7, 2)])
def test_gemm_tile_orders_are_bijections(tile_order_exe, order, tm, tn):
    """Each GEMM tile order visits every (ti, tj) of the tm x tn grid once:
    ti fastest (0), tj fastest (1), and bands of `order` tile columns with
    tj fastest inside a band (>= 2), the last band narrower when order does
    not divide tn, with tm > 1."""
    seq = tile_coords_of(tile_order_exe, 0, order, tm, tn)
    assert sorted(seq) == [(i, j) for i in range(tm) for j in range(tn)]
    if order == 0:
        assert seq == [(i, j) for j in range(tn) for i in range(tm)]
    elif order == 1:
        assert seq == [(i, j) for i in range(tm) for j in range(tn)]
    else:
        want = []
        for b0 in range(0, tn, order):
            cols = range(b0, min(b0 + order, tn))
            want += [(i, j) for i in range(tm) for j in cols]
        assert seq == want


@pytest.mark.parametrize("tm", [1, 2, 4, 9, 33])
def test_syrk_tile_order_is_the_lower_triangle(tile_order_exe, tm):
    seq = tile_coords_of(tile_order_exe, 1, 0, tm, tm)
    assert seq == [(i, j) for i in range(tm) for j in range(i + 1)]


# ---- Stream-K fix-up flag protocol under arbitrary interleavings ----------

class FixupMemory:
    def __init__(self, G):
        self.slot = [None] * G      # partial-tile slot of each workgroup
        self.flag = [0] * G         # 1: that workgroup's partial is ready
        self.status = 0             # bit 0: a wait hit its spin limit
        self.out = {}               # tile -> value written


def workgroup(w, G, ntiles, ipt, mem, spin_limit, starts):
    """sk_dgemm_kernel's per-workgroup loop (skgemm_kernels.hip, fix-up
    section): a generator that yields at every shared-memory access, so a
    scheduler can interleave workgroups arbitrarily.  The value of a
    partial is the sum of its iteration indices, so a tile's exact result is
    the sum over the tile's iterations."""
    total = ntiles * ipt
    it, end = starts(w), starts(w + 1)
    while it < end:
        tile = it // ipt
        t0 = tile * ipt
        kb, ke = it - t0, min(end - t0, ipt)
        acc = sum(range(t0 + kb, t0 + ke))
        yield
        if kb != 0:
            # spill: stores complete (s_waitcnt 0, barrier), then the flag
            assert mem.slot[w] is None, "a slot is written once per launch"
            mem.slot[w] = acc
            yield
            mem.flag[w] = 1
            yield
            it = t0 + ke
            continue
        if ke < ipt:
            w2 = w + 1
            while w2 < G and starts(w2) < t0 + ipt:
                if starts(w2 + 1) == starts(w2):
                    w2 += 1
                    continue
                n = 0
                while mem.flag[w2] == 0:
                    yield
                    n += 1
                    if n > spin_limit:
                        mem.status |= 1
                        break
                mem.flag[w2] = 0          # consumed (cleared by its consumer)
                yield
                acc += mem.slot[w2] if mem.slot[w2] is not None else 0
                yield
                w2 += 1
        mem.out[tile] = acc
        yield
        it = t0 + ke


def run_launch(G, ntiles, ipt, mem, rng, spin_limit=1 << 40, hold=None,
               hold_steps=0):
    """Random interleaving of all G workgroups (all resident at once).  A
    workgroup in `hold` is not scheduled during the first hold_steps steps
    (a late peer).  Returns the number of scheduler steps."""
    total = ntiles * ipt
    starts = lambda x: wg_start(x, total, G)
    mem.slot = [None] * G
    mem.out = {}
    live = {w: workgroup(w, G, ntiles, ipt, mem, spin_limit, starts)
            for w in range(G)}
    steps = 0
    while live:
        ready = [w for w in live
                 if hold is None or w not in hold or steps >= hold_steps]
        w = ready[int(rng.integers(len(ready)))] if ready else None
        if w is None:
            steps += 1
            continue
        try:
            next(live[w])
        except StopIteration:
            del live[w]
        steps += 1
        assert steps < 10 ** 6, "no progress: a wait that never ends"
    return steps


@pytest.mark.parametrize("ntiles,ipt,G", [(1, 12, 5), (3, 7, 4), (10, 3, 7),
                                          (2, 40, 9), (5, 5, 25), (7, 9, 3)])
def test_fixup_protocol_random_interleavings(ntiles, ipt, G):
    """Every tile's result is exact, every flag is consumed and cleared by
    the end of the launch (so a replayed graph needs no memset), and no
    schedule deadlocks: an owner only waits on later workgroups, whose
    partial is the first unit of their range."""
    rng = np.random.default_rng(ntiles * 1000 + ipt * 10 + G)
    G = min(G, ntiles * ipt)
    mem = FixupMemory(G)
    for trial in range(200):
        run_launch(G, ntiles, ipt, mem, rng)
        assert mem.status == 0
        assert mem.flag == [0] * G
        for t in range(ntiles):
            assert mem.out[t] == sum(range(t * ipt, (t + 1) * ipt)), (trial, t)


def workgroup_waitfree(w, G, ntiles, ipt, mem, starts):
    """The wait-free (last-arriver) fix-up of sk_dgemm_kernel
    (skgemm_kernels.hip, L.waitfree): every contributor of a split tile
    stores its partial (the tile's first contributor wo in slot G + wo,
    the others in their own slot), then adds one to flags[wo]; the one that
    saw ncontrib - 1 resets the counter and adds the partials in contributor
    order.  A partial is (sum of its iterations, (contributor,)), so the
    order of the additions is checked along with the value."""
    it, end = starts(w), starts(w + 1)
    while it < end:
        tile = it // ipt
        t0 = tile * ipt
        kb, ke = it - t0, min(end - t0, ipt)
        acc = (sum(range(t0 + kb, t0 + ke)), (w,))
        yield
        if kb != 0 or ke < ipt:
            wo = w
            if kb != 0:
                wo = w - 1
                while starts(wo) > t0:
                    wo -= 1
            contrib = [x for x in range(wo, G)
                       if starts(x) < t0 + ipt and starts(x + 1) != starts(x)]
            mine = G + w if w == wo else w
            assert mem.slot[mine] is None, "a slot is written once per launch"
            mem.slot[mine] = acc
            yield                                   # stores complete
            before = mem.flag[wo]
            mem.flag[wo] = before + 1               # one atomic add
            last = before == len(contrib) - 1
            if last:
                mem.flag[wo] = 0
            yield
            if not last:
                it = t0 + ke
                continue
            total, order = 0, ()
            for x in contrib:
                v = mem.slot[G + x if x == wo else x]
                assert v is not None, "a contributor's partial is missing"
                total, order = total + v[0], order + v[1]
                yield
            acc = (total, order)
        mem.out[tile] = acc
        yield
        it = t0 + ke


def run_launch_waitfree(G, ntiles, ipt, mem, rng, resident):
    """At most `resident` workgroups run at once, admitted in a random
    order (a launch sharing the device with another one); the rest start
    only as running ones finish."""
    total = ntiles * ipt
    starts = lambda x: wg_start(x, total, G)
    mem.slot = [None] * (2 * G)
    mem.out = {}
    waiting = list(rng.permutation(G))
    live = {}
    steps = 0
    while live or waiting:
        while waiting and len(live) < resident:
            x = int(waiting.pop())
            live[x] = workgroup_waitfree(x, G, ntiles, ipt, mem, starts)
        w = list(live)[int(rng.integers(len(live)))]
        try:
            next(live[w])
        except StopIteration:
            del live[w]
        steps += 1
        assert steps < 10 ** 6, "no progress"
    return steps


@pytest.mark.parametrize("ntiles,ipt,G", [(1, 12, 5), (3, 7, 4), (10, 3, 7),
                                          (2, 40, 9), (5, 5, 25), (7, 9, 3)])
@pytest.mark.parametrize("resident", [1, 2, 1000])
def test_waitfree_fixup_any_residency(ntiles, ipt, G, resident):
    """The wait-free fix-up finishes under any interleaving and with any
    number of workgroups resident (one at a time included, where the
    waiting fix-up's owner would wait forever for a peer that cannot
    start), every tile is exact, the partials are added in contributor
    order (the waiting fix-up's order, so the two modes give the same
    bits), and every counter is back at zero."""
    rng = np.random.default_rng(ntiles * 1000 + ipt * 10 + G + resident)
    G = min(G, ntiles * ipt)
    total = ntiles * ipt
    mem = FixupMemory(G)
    for trial in range(100):
        run_launch_waitfree(G, ntiles, ipt, mem, rng, resident)
        assert mem.flag == [0] * G
        for t in range(ntiles):
            got, order = mem.out[t]
            assert got == sum(range(t * ipt, (t + 1) * ipt)), (trial, t)
            want = tuple(x for x in range(G)
                         if wg_start(x, total, G) < (t + 1) * ipt
                         and wg_start(x + 1, total, G) > t * ipt)
            assert order == want, (trial, t)


def test_waiting_fixup_needs_coresidency():
    """The contrast: with one workgroup resident at a time and the owner
    admitted first, the waiting fix-up's owner spins until its limit (the
    reason the waiting mode caps a launch at what the device holds)."""
    rng = np.random.default_rng(3)
    ntiles, ipt, G = 1, 20, 2
    total = ntiles * ipt
    starts = lambda x: wg_start(x, total, G)
    mem = FixupMemory(G)
    owner = workgroup(0, G, ntiles, ipt, mem, 100, starts)
    for _ in owner:       # runs alone: the peer never starts
        pass
    assert mem.status == 1


def test_fixup_timeout_then_reset():
    """A peer that is late beyond the owner's spin limit: the owner gives up
    (status bit 0, which the host reports as SKGEMM_ERR_TIMEOUT), every
    workgroup still finishes, and the late peer's flag is left set.  Without
    the host's reset (skgemm_ctx::check_status zeroes the flags on a
    timeout) the next launch's owner would take that stale flag and add a
    slot before its new partial is there; with the reset the next launch is
    exact."""
    rng = np.random.default_rng(7)
    ntiles, ipt, G = 1, 20, 2      # owner w0, peer w1 on the one tile
    mem = FixupMemory(G)
    run_launch(G, ntiles, ipt, mem, rng, spin_limit=2, hold={1},
               hold_steps=50)
    assert mem.status == 1
    assert mem.flag[1] == 1                           # stale flag
    exact = sum(range(ipt))
    assert mem.out[0] != exact
    # Next launch without the reset: the owner may consume the stale flag
    # before the peer rewrote its slot (peer held back so it is certain).
    stale = FixupMemory(G)
    stale.flag = list(mem.flag)
    run_launch(G, ntiles, ipt, stale, rng, hold={1}, hold_steps=50)
    assert stale.out[0] != exact
    # With the reset (status and flags zeroed), any interleaving is exact.
    mem.status, mem.flag = 0, [0] * G
    for _ in range(50):
        run_launch(G, ntiles, ipt, mem, rng)
        assert mem.status == 0 and mem.out[0] == exact and mem.flag == [0, 0]
