"""Benchmark of the FP64 GEMM/SYRK hot path on one MI355X.

One step = one DFT subspace step on one grid block: the overlap SYRK
Ms = Psi^T Psi (lower triangle) followed by the rotation GEMM Psi * Q, with
Psi 32768 x 520 (Nd = 32^3 grid points, Ns = 520 orbitals, normal(0,
1/sqrt(Nd)), seed 1, column-major, resident in HBM) and Q 520 x 520
uniform(-1, 1) (seed 2, row-major).  Plain FP64 on the FP64 MFMA path.

Prints ONE JSON line.  value = whole-job TFLOP/s (algorithmic FLOP:
SYRK N(N+1)K + rotation 2 Nd Ns^2) over the timed steps, inputs already in
HBM.  roofline: the dominant kernel's algorithmic FLOP per launch divided by
its mean launch time from HIP events recorded on the launch stream.
cpu_baseline: the oracle's C++/OpenMP restatement of the reference CPU path
(syrk tile pairs + trp_3d rotation), built -march=native here, timed on a
bounded sample of the same step.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 10] [--no-cpu]
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "stream-k-gemm_amd")
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 matrix (datasheet)
PEAK_HBM_GBS = 8000.0
METRIC = "FP64-accurate GEMM/SYRK TFLOP/s + p99 latency: plain FP64 vs Ozaki split"


def load_skgemm():
    name = "skgemm_amd"
    if name in sys.modules:
        return sys.modules[name].skgemm
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod.skgemm


def log(msg):
    sys.stderr.write("[bench] %s\n" % msg)
    sys.stderr.flush()


def traffic_from_profiles(kernel_key):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy-kernel calibration) and committed under
    profiles/; None when no measurement is committed."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(kernel_key, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(psi_host, q_host, Nd, Ns, gpu_ms, gpu_rot):
    """Time the oracle's restatement of the reference CPU path on the same
    step: oracle_syrk_tiles on P16 (syrk_crsp.cpp:306-348) and oracle_trp_3d
    on the grid-packed layout (include/sme_trp_3d.hpp:194-215); packing is
    outside the timed region as in the reference (gemm_crsp.cpp:395-405)."""
    sys.path.insert(0, ROOT)
    from oracle import oracle
    tmp = tempfile.mkdtemp(prefix="skgemm_oracle_")
    so = os.path.join(tmp, "liboracle_native.so")
    try:
        oracle.build(native=True, out=so)
    except subprocess.CalledProcessError:
        so = None
    L = oracle.load(so)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    P = psi_host.reshape(Ns, Nd)                      # row b = column b of Psi
    t = (Ns + 15) // 16
    p16 = np.zeros((t * 16, Nd))
    p16[:Ns] = P
    p16 = np.ascontiguousarray(p16.reshape(t, 16, Nd).transpose(0, 2, 1)).ravel()
    grid = np.ascontiguousarray(
        P.T.reshape(Nd // 16, 16, Ns).transpose(0, 2, 1)).ravel()
    C = np.zeros(Ns * Ns)
    D = np.zeros(t * Nd * 16)
    flops = Ns * (Ns + 1) * Nd + 2.0 * Nd * Ns * Ns

    def step():
        L.oracle_syrk_tiles(Nd, Ns, p16.ctypes.data, C.ctypes.data, 1.0, 0.0,
                            cores)
        L.oracle_trp_3d(Nd, Ns, grid.ctypes.data, q_host.ctypes.data,
                        D.ctypes.data, cores)

    step()  # warm-up (WITERS = 1)
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_start < 10.0
                             and len(times) < 50):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
    mean_s = float(np.mean(times))
    # the CPU result of the last step against the GPU rotation (parity check)
    rot = D.reshape(t, Nd, 16).transpose(1, 0, 2).reshape(Nd, t * 16)[:, :Ns]
    dev = float(np.linalg.norm(rot - gpu_rot) / np.linalg.norm(rot))
    return {
        "value": flops / mean_s / 1e12,
        "unit": "TFLOP/s",
        "cores": cores,
        "kind": "port",
        "sample": ("full step (SYRK 32768x520 + rotation 32768x520x520) x %d "
                   "after 1 warm-up, %.2f s mean per step, OpenMP threads=%d, "
                   "build %s" % (len(times), mean_s, cores,
                                 "-march=native" if so else "x86-64-v3")),
        "rotation_rel_diff_vs_gpu": dev,
    }


TUNING_CACHE = None  # set from --tuning-cache


def tuned(ctx, kind, M, N, K, la, lb=0):
    """The plan for this shape: from the tuning cache when it holds the
    shape (the cache was loaded into the context at start-up), else by the
    autotuner, which appends it to the cache."""
    if TUNING_CACHE and os.path.exists(TUNING_CACHE):
        with open(TUNING_CACHE) as f:
            for line in f:
                v = line.split()
                if (len(v) == 10 and int(v[0]) == kind and int(v[5]) == K and
                        sorted((int(v[3]), int(v[4]))) == sorted((M, N))):
                    p = ctx.query_plan(kind, M, N, K, la=la, lb=lb)
                    return {"tile": p["tile"], "grid": p["grid"],
                            "kchunks": p["kchunks"], "us": float(v[9])}
    return ctx.autotune(kind, M, N, K, la=la, lb=lb, cache_path=TUNING_CACHE)


def median_ms(ctx, fn, reps):
    """Median device time of `reps` calls from the context's timing ring
    (HIP events on the launch stream), after one warm-up call."""
    fn()
    ctx.synchronize()
    ctx.timing_enable(reps)
    for _ in range(reps):
        fn()
    ctx.synchronize()
    return float(np.median(ctx.timing_read(reps)))


def lower_rel(a, b, n):
    il = np.tril_indices(n)
    x = a.reshape(n, n)[il]
    y = b.reshape(n, n)[il]
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def bench_c1(skg, ctx, reps):
    """c1: A, B 1024 x 1024 uniform(-1, 1) (seeds 42, 43), row-major, D = A B,
    autotuned plan."""
    n = 1024
    a = skg.DeviceArray(ctx, n * n)
    b = skg.DeviceArray(ctx, n * n)
    d = skg.DeviceArray(ctx, n * n, zero=True)
    ctx.fill_uniform(a.ptr, n * n, 42, -1.0, 1.0)
    ctx.fill_uniform(b.ptr, n * n, 43, -1.0, 1.0)
    plan = tuned(ctx, 0, n, n, n, la=skg.ROW_MAJOR, lb=skg.ROW_MAJOR)
    t = median_ms(ctx, lambda: ctx.dgemm(
        skg.ROW_MAJOR, skg.ROW_MAJOR, skg.ROW_MAJOR, n, n, n, 1.0, a.ptr, n,
        b.ptr, n, 0.0, d.ptr, n, 0.0), reps)
    for x in (a, b, d):
        x.free()
    f = 2.0 * n ** 3
    return {"workload": "GEMM 1024^3 FP64 row-major, D = A B", "ms": t,
            "tflops": f / (t * 1e-3) / 1e12,
            "frac_fp64_peak": f / (t * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
            "plan": plan}


def bench_split(skg, ctx, reps):
    """Split-form SYRK on Psi 65536 x 1024 (columns scaled by 10^u, u ~
    U(0, 6), seed 3) against the autotuned plain FP64 SYRK on the same
    input, and the split-form rotation Psi Q against the plain FP64
    rotation; error of each split result against the plain FP64 result and
    the bound the library returns."""
    Nd, Ns = 65536, 1024
    psi = skg.DeviceArray(ctx, Nd * Ns)
    ctx.fill_normal(psi.ptr, Nd * Ns, 3, 1.0 / np.sqrt(Nd))
    ctx.scale_columns_pow10(psi.ptr, Nd, Ns, Nd, 4, 6.0)
    plain = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    split = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    plan = tuned(ctx, 1, Ns, Ns, Nd, la=skg.COL_MAJOR)
    t_plain = median_ms(ctx, lambda: ctx.dsyrk(
        skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0, plain.ptr, Ns), reps)
    out = {"workload": "SYRK Ms = Psi^T Psi (lower), Psi 65536x1024 FP64 "
                       "col-major, columns scaled by 10^U(0,6); split form "
                       "vs plain FP64; rotation Psi*Q, plain FP64 and "
                       "split form",
           "plain_fp64": {"ms": t_plain, "plan": plan}}
    f = float(Ns) * (Ns + 1) * Nd
    out["plain_fp64"]["tflops"] = f / (t_plain * 1e-3) / 1e12
    for L in (14, 13):
        bound = ctx.dsyrk_ozaki(skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0,
                                split.ptr, Ns, slices=L)
        t = median_ms(ctx, lambda: ctx.dsyrk_ozaki(
            skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0, split.ptr, Ns,
            slices=L), reps)
        err = lower_rel(split.download(), plain.download(), Ns)
        out["split_L%d" % L] = {
            "ms": t, "tflops_fp64_equivalent": f / (t * 1e-3) / 1e12,
            "speedup_vs_plain": t_plain / t, "rel_fro_vs_plain_fp64": err,
            "stated_bound": bound, "moduli": L,
            "int8_tops": L * f / (t * 1e-3) / 1e12}
    # plain FP64 rotation Psi * Q on the same Psi (Q uniform(-1, 1), seed 4)
    q = skg.DeviceArray(ctx, Ns * Ns)
    rot = skg.DeviceArray(ctx, Nd * Ns)
    ctx.fill_uniform(q.ptr, Ns * Ns, 4, -1.0, 1.0)
    rplan = tuned(ctx, 0, Nd, Ns, Ns, la=skg.COL_MAJOR, lb=skg.ROW_MAJOR)
    t_rot = median_ms(ctx, lambda: ctx.dgemm(
        skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, Nd, Ns, Ns, 1.0, psi.ptr,
        Nd, q.ptr, Ns, 0.0, rot.ptr, Nd, 0.0), reps)
    fr = 2.0 * Nd * Ns * Ns
    out["plain_fp64_rotation"] = {
        "ms": t_rot, "tflops": fr / (t_rot * 1e-3) / 1e12,
        "frac_fp64_peak": fr / (t_rot * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
        "plan": rplan}
    # Split-form rotation on the same operands (14 moduli), its error
    # against the plain FP64 rotation and the bound the library returns.
    rot2 = skg.DeviceArray(ctx, Nd * Ns)

    def split_rot():
        return ctx.dgemm_ozaki(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR,
                               Nd, Ns, Ns, 1.0, psi.ptr, Nd, q.ptr, Ns, 0.0,
                               rot2.ptr, Nd, 0.0, slices=14)
    rbound = split_rot()
    t_srot = median_ms(ctx, split_rot, reps)
    got, want = rot2.download(), rot.download()
    out["split_rotation_L14"] = {
        "ms": t_srot, "tflops_fp64_equivalent": fr / (t_srot * 1e-3) / 1e12,
        "speedup_vs_plain": t_rot / t_srot,
        "rel_fro_vs_plain_fp64": float(np.linalg.norm(got - want) /
                                       np.linalg.norm(want)),
        "stated_bound": rbound, "moduli": 14}
    del got, want
    out["plain_fp64"]["frac_fp64_peak"] = (out["plain_fp64"]["tflops"] /
                                           PEAK_FP64_TFLOPS)
    for x in (psi, q, rot, rot2):
        x.free()
    return out


def bench_projection(skg, ctx, reps):
    """Projection GEMM hp = Psi^T (H Psi), Ns = 616, Nd = 175616 (the
    reference's sweep shape, sub_gemm.sh:26-30), Psi band-packed 16 (PA)
    and H Psi band-packed 32 (PB), D row-major: the gemm_crsp packed entry
    point's layouts.  Default plan, then the autotuned one."""
    Ns, Nd = 616, 175616
    tA, tB = (Ns + 15) // 16, (Ns + 31) // 32
    psi = skg.DeviceArray(ctx, tA * Nd * 16)
    hpsi = skg.DeviceArray(ctx, tB * Nd * 32)
    hp = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    ctx.fill_normal(psi.ptr, tA * Nd * 16, 11, 1.0 / np.sqrt(Nd))
    ctx.fill_normal(hpsi.ptr, tB * Nd * 32, 12, 1.0 / np.sqrt(Nd))

    def run():
        ctx.dgemm(skg.PACK16, skg.PACK32, skg.ROW_MAJOR, Ns, Ns, Nd, 1.0,
                  psi.ptr, Nd, hpsi.ptr, Nd, 0.0, hp.ptr, Ns, 0.0)
    f = 2.0 * Ns * Ns * Nd
    t0 = median_ms(ctx, run, reps)
    plan = tuned(ctx, 0, Ns, Ns, Nd, la=skg.PACK16, lb=skg.PACK32)
    t1 = median_ms(ctx, run, reps)
    for x in (psi, hpsi, hp):
        x.free()
    return {"workload": "projection GEMM hp = Psi^T (H Psi), Ns 616, Nd "
                        "175616, FP64, PA / PB packed",
            "default_plan": {"ms": t0, "tflops": f / (t0 * 1e-3) / 1e12},
            "tuned": {"ms": t1, "tflops": f / (t1 * 1e-3) / 1e12,
                      "frac_fp64_peak": f / (t1 * 1e-3) / 1e12 /
                      PEAK_FP64_TFLOPS, "plan": plan}}


def bench_nloc(skg, ctx, reps):
    """Non-local projector operator result += V_nloc psi on the c2 grid
    (Nd 32768, 520 bands, psi 16-band-packed, result 32-band-packed), 64
    projectors of 512..1024 grid points (compact spheres) and 9, 13 or 16
    columns, seeded.  HBM-bound: the algorithmic traffic is psi read at the
    touched points plus result read and written there (3 x 8 x points x
    nb bytes); the implementation also writes and reads the gathered rows
    and the backward products (4 x 8 x rows x nb more)."""
    import time
    K, nb, count = 32768, 520, 64
    rng = np.random.default_rng(6)
    nrows = [int(rng.integers(512, 1025)) for _ in range(count)]
    ncols = [int(rng.choice([9, 13, 16])) for _ in range(count)]
    idx = []
    for n in nrows:
        c = int(rng.integers(0, K))
        idx.append(np.unique((c + np.arange(-(n // 2), n - n // 2) * 3) % K)[:n])
    nrows = [len(x) for x in idx]
    chi = rng.uniform(-1, 1, sum(r * c for r, c in zip(nrows, ncols)))
    gam = rng.uniform(-1, 1, sum(ncols))
    idx_all = np.concatenate(idx)
    points = len(np.unique(idx_all))
    rows = int(sum(nrows))
    h = skg.Nloc(ctx, K, nrows, ncols, list(range(count)), idx_all, chi, gam)
    psi = skg.DeviceArray(ctx, ((nb + 15) // 16) * K * 16)
    res = skg.DeviceArray(ctx, ((nb + 31) // 32) * K * 32, zero=True)
    ctx.fill_normal(psi.ptr, ((nb + 15) // 16) * K * 16, 7, 1.0 / np.sqrt(K))
    h.apply(nb, 0.02, psi.ptr, res.ptr, 32)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        h.apply(nb, 0.02, psi.ptr, res.ptr, 32)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    h.close()
    for x in (psi, res):
        x.free()
    alg = 3.0 * 8 * points * nb
    moved = alg + 4.0 * 8 * rows * nb
    flop = 4.0 * nb * sum(r * c for r, c in zip(nrows, ncols))
    return {"workload": "nloc projectors: Nd 32768, 520 bands, 64 projectors "
                        "(512..1024 points, 9/13/16 columns), 16->32 packed",
            "ms": ms, "timing": "host wall clock over %d applications" % reps,
            "grid_points_touched": points, "projector_rows": rows,
            "algorithmic_gbps": alg / (ms * 1e-3) / 1e9,
            "moved_gbps": moved / (ms * 1e-3) / 1e9,
            "frac_hbm_peak_algorithmic": alg / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "gflops": flop / (ms * 1e-3) / 1e9}


def bench_stencil(skg, ctx, reps):
    """Real-space stencil on the c2 grid: 32^3 points, 520 bands, FDn 6
    (12th order), periodic halo.  HBM-bound; algorithmic bytes: halo fill
    reads the local grid and writes the extended one, the stencil reads the
    extended grid and the diagonal once and writes the 32-band output."""
    n, nb, F = 32, 520, 6
    K, K_ex = n ** 3, (n + 2 * F) ** 3
    t16, t32 = (nb + 15) // 16, (nb + 31) // 32
    loc = skg.DeviceArray(ctx, t16 * K * 16)
    ext = skg.DeviceArray(ctx, t16 * K_ex * 16, zero=True)
    A = skg.DeviceArray(ctx, K)
    out = skg.DeviceArray(ctx, t32 * K * 32, zero=True)
    ctx.fill_normal(loc.ptr, t16 * K * 16, 8, 1.0 / np.sqrt(K))
    ctx.fill_uniform(A.ptr, K, 9, -1.0, 1.0)
    c = [0, -1.0 / 12, 2.0 / 3, -2.5, 8.0 / 3, -1.0 / 4, 1.0 / 30]
    th = median_ms(ctx, lambda: ctx.halo_extend(n, n, n, nb, F, loc.ptr,
                                                ext.ptr, True), reps)
    ts = median_ms(ctx, lambda: ctx.stencil_apply(n, n, n, nb, F, ext.ptr,
                                                  A.ptr, c, c, c, 0.5,
                                                  out.ptr, 32), reps)
    for x in (loc, ext, A, out):
        x.free()
    bh = 8.0 * 16 * t16 * (K + K_ex)
    bs = 8.0 * (16 * t16 * K_ex + K + nb * K)
    return {"workload": "stencil H psi, 32^3 grid, 520 bands, FDn 6, "
                        "16-band-packed extended in, 32-band-packed out",
            "halo_extend": {"ms": th, "gbps": bh / (th * 1e-3) / 1e9,
                            "frac_hbm_peak": bh / (th * 1e-3) / 1e9 /
                            PEAK_HBM_GBS},
            "stencil": {"ms": ts, "gbps": bs / (ts * 1e-3) / 1e9,
                        "frac_hbm_peak": bs / (ts * 1e-3) / 1e9 / PEAK_HBM_GBS,
                        "gflops": (1 + 6 * F) * 2.0 * K * nb / (ts * 1e-3) /
                        1e9}}


def bench_layouts(skg, ctx, reps):
    """Layout conversions of the SYRK / rotation operand Psi, Nd x nb =
    65536 x 1024 FP64 (512 MiB): column-major -> grid-packed (pack_c16),
    grid-packed -> band-packed and band-packed -> grid-packed (the
    band/grid tile conversions).  Algorithmic traffic: 8 bytes read + 8
    written per element."""
    Nd, nb = 65536, 1024
    n = Nd * nb
    col = skg.DeviceArray(ctx, n)
    grid = skg.DeviceArray(ctx, n, zero=True)
    band = skg.DeviceArray(ctx, n, zero=True)
    ctx.fill_normal(col.ptr, n, 7, 1.0)
    runs = {
        "pack_c16": lambda: ctx.pack(skg.ROW_MAJOR, col.ptr, Nd, skg.PACK16,
                                     grid.ptr, nb, Nd, nb),
        "grid_to_band": lambda: ctx.transpose(skg.PACK16, grid.ptr, nb,
                                              skg.PACK16, band.ptr, Nd, Nd,
                                              nb),
        "band_to_grid": lambda: ctx.transpose(skg.PACK16, band.ptr, Nd,
                                              skg.PACK16, grid.ptr, nb, nb,
                                              Nd),
    }
    out = {"workload": "Psi 65536 x 1024 FP64 layout conversions",
           "bytes_per_launch": 16 * n}
    for name, fn in runs.items():
        t = median_ms(ctx, fn, reps)
        gbps = 16.0 * n / (t * 1e-3) / 1e9
        out[name] = {"ms": t, "gbps": gbps, "frac_hbm_peak": gbps / PEAK_HBM_GBS}
    for x in (col, grid, band):
        x.free()
    return out


def c4_batch(skg, ctx, seed=5, count=64):
    """64 systems: Ns ~ U{128, 136, ..., 1536}, Nd ~ U{4096, 13824, 32768,
    65536}, Psi ~ N(0, 1/sqrt(Nd)) (seeded per system), Q ~ U(-1, 1); one
    overlap SYRK and one rotation GEMM per system."""
    rng = np.random.default_rng(seed)
    probs, keep, flops = [], [], 0.0
    for i in range(count):
        Ns = int(rng.choice(np.arange(128, 1537, 8)))
        Nd = int(rng.choice([4096, 13824, 32768, 65536]))
        psi = skg.DeviceArray(ctx, Nd * Ns)
        ctx.fill_normal(psi.ptr, Nd * Ns, 1000 + i, 1.0 / np.sqrt(Nd))
        q = skg.DeviceArray(ctx, Ns * Ns)
        ctx.fill_uniform(q.ptr, Ns * Ns, 2000 + i, -1.0, 1.0)
        ms = skg.DeviceArray(ctx, Ns * Ns, zero=True)
        rot = skg.DeviceArray(ctx, Nd * Ns, zero=True)
        probs.append(skg.Problem(kind=1, layout_a=skg.COL_MAJOR, M=Ns, N=Ns,
                                 K=Nd, alpha=1.0, A=psi.ptr, lda=Nd,
                                 D=ms.ptr, ldd=Ns))
        probs.append(skg.Problem(kind=0, layout_a=skg.COL_MAJOR,
                                 layout_b=skg.ROW_MAJOR,
                                 layout_d=skg.COL_MAJOR, M=Nd, N=Ns, K=Ns,
                                 alpha=1.0, A=psi.ptr, lda=Nd, B=q.ptr,
                                 ldb=Ns, D=rot.ptr, ldd=Nd))
        flops += float(Ns) * (Ns + 1) * Nd + 2.0 * Nd * Ns * Ns
        keep += [psi, q, ms, rot]
    ctx.synchronize()
    return probs, keep, flops


def bench_batched(skg, ctx, reps):
    probs, keep, flops = c4_batch(skg, ctx)
    out = {"workload": "64 systems x (overlap SYRK + rotation GEMM), Ns ~ "
                       "U{128..1536 step 8}, Nd ~ U{4096,13824,32768,65536}, "
                       "seed 5; one persistent launch per batch",
           "problems": len(probs), "flop": flops}
    for name, prec, graph, r in (("plain_fp64_graph", skg.FP64, 1, reps),
                                 ("plain_fp64_eager", skg.FP64, 0, reps),
                                 ("split_graph", skg.OZAKI_INT8, 1,
                                  max(2, reps // 2))):
        t = median_ms(ctx, lambda: ctx.batched(probs, precision=prec,
                                               use_graph=graph), r)
        out[name] = {"ms": t, "tflops": flops / (t * 1e-3) / 1e12}
    for a in keep:
        a.free()
    return out


def bench_serving(skg, ctx, c4_tflops, seconds, precision):
    """200 resident systems (Ns ~ U[128, 2048], Nd ~ U[8192, 65536], seed 6,
    ~64 GB of Psi); overlap and rotation requests arrive as a Poisson
    process at ~70% of the plain FP64 batched throughput; latency = submit
    -> completion observed by the server.  `precision`: plain FP64 or the
    split form (14 moduli), same arrivals."""
    import threading
    rng = np.random.default_rng(6)
    systems = []
    for i in range(200):
        Ns = int(rng.integers(128, 2049))
        Nd = int(rng.integers(8192, 65537))
        psi = skg.DeviceArray(ctx, Nd * Ns)
        ctx.fill_normal(psi.ptr, Nd * Ns, 3000 + i, 1.0 / np.sqrt(Nd))
        q = skg.DeviceArray(ctx, Ns * Ns)
        ctx.fill_uniform(q.ptr, Ns * Ns, 4000 + i, -1.0, 1.0)
        systems.append((psi, q, Nd, Ns))
    ctx.synchronize()
    pool_gb = sum(s[2] * s[3] * 8 for s in systems) / 1e9
    mean_flop = np.mean([s[3] * (s[3] + 1) * s[2] + 2.0 * s[2] * s[3] ** 2
                         for s in systems]) / 2.0
    lam = 0.7 * c4_tflops * 1e12 / mean_flop           # requests per second
    n_req = max(50, int(lam * seconds))
    kinds = rng.integers(0, 2, n_req)
    picks = rng.integers(0, len(systems), n_req)
    gaps = rng.exponential(1.0 / lam, n_req)
    # Outputs: a ring of buffers sized for the largest result, longer than
    # the number of requests that can be in flight (2 launches x 16).
    ring = 48
    max_out = max(s[2] * s[3] for s in systems)
    outs = [skg.DeviceArray(ctx, max_out) for _ in range(ring)]
    srv = skg.Server(0, hbm_budget=int(pool_gb * 1e9) + (16 << 30),
                     max_batch=16, max_inflight=2, precision=precision)
    lat = []
    try:
        ids = [srv.add_system(s[0].ptr, s[2], s[3]) for s in systems]
        pending = []
        lock = threading.Lock()
        done = threading.Event()

        def waiter():
            while True:
                with lock:
                    item = pending.pop(0) if pending else None
                if item is None:
                    if done.is_set():
                        with lock:
                            if not pending:
                                return
                    time.sleep(0.0002)
                    continue
                lat.append(srv.wait(item))

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        t_next = t0
        for j in range(n_req):
            t_next += gaps[j]
            while time.perf_counter() < t_next:
                pass
            si, kind = int(picks[j]), int(kinds[j])
            psi, q, Nd, Ns = systems[si]
            tk = srv.submit(ids[si], kind, outs[j % ring].ptr,
                            q.ptr if kind == 1 else None)
            with lock:
                pending.append(tk)
        done.set()
        th.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
    finally:
        srv.close()
    lat = np.array(lat)
    out = {"workload": "200 resident systems (Ns ~ U[128,2048], Nd ~ "
                       "U[8192,65536], seed 6), Poisson arrivals, overlap or "
                       "rotation per request, " +
                       ("plain FP64" if precision == skg.FP64 else
                        "split form (14 moduli)"),
           "resident_psi_gb": pool_gb, "arrival_rate_per_s": lam,
           "requests": int(n_req), "wall_s": wall,
           "latency_ms_p50": float(np.percentile(lat, 50)) / 1e3,
           "latency_ms_p99": float(np.percentile(lat, 99)) / 1e3,
           "tflops": st["flops"] / wall / 1e12,
           "launches": st["launches"], "peak_inflight": st["peak_inflight"],
           "requests_per_launch": st["completed"] / max(1, st["launches"])}
    for s in systems:
        s[0].free()
        s[1].free()
    for o in outs:
        o.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the c1, split-form, batched and serving "
                         "sections")
    ap.add_argument("--extra", default="c1,split,projection,nloc,stencil,layouts,batched,serving",
                    help="comma list of extra sections to run")
    ap.add_argument("--tuning-cache", default=None,
                    help="autotuner cache file: read if it exists (no "
                         "tuning runs), else written by the tuning runs")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(a K-sharded SYRK would add one all-reduce of Ms; see the "
              "design notes).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    skg = load_skgemm()
    ctx = skg.Context(0)
    info = ctx.device_info()
    Nd, Ns = 32768, 520
    log("device %s, workload c2 Nd=%d Ns=%d" % (info, Nd, Ns))
    psi = skg.DeviceArray(ctx, Nd * Ns)
    q = skg.DeviceArray(ctx, Ns * Ns)
    ms = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    rot = skg.DeviceArray(ctx, Nd * Ns, zero=True)
    ctx.fill_normal(psi.ptr, Nd * Ns, 1, 1.0 / np.sqrt(Nd))
    ctx.fill_uniform(q.ptr, Ns * Ns, 2, -1.0, 1.0)
    global TUNING_CACHE
    TUNING_CACHE = args.tuning_cache
    plans = {}
    if args.tile >= 0 or args.grid > 0:
        ctx.set_tuning(args.tile, args.grid)
    elif args.tuning_cache and os.path.exists(args.tuning_cache):
        ctx.load_tuning(args.tuning_cache)
        for name, kind, M, N, K in (("syrk", 1, Ns, Ns, Nd),
                                    ("rotation", 0, Nd, Ns, Ns)):
            p = ctx.query_plan(kind, M, N, K, la=skg.COL_MAJOR,
                               lb=skg.ROW_MAJOR)
            plans[name] = {"tile": p["tile"], "grid": p["grid"],
                           "kchunks": p["kchunks"], "us": None}
        log("plans from %s: %s" % (args.tuning_cache, plans))
    elif not args.no_tune:
        # Autotuner: picks tile / grid / K chunks per shape with real
        # launches, before (and outside) the timed region.
        plans["syrk"] = ctx.autotune(1, Ns, Ns, Nd, la=skg.COL_MAJOR,
                                     cache_path=args.tuning_cache)
        plans["rotation"] = ctx.autotune(0, Nd, Ns, Ns, la=skg.COL_MAJOR,
                                         lb=skg.ROW_MAJOR,
                                         cache_path=args.tuning_cache)
        log("autotuned plans %s" % plans)

    def step():
        ctx.dsyrk(skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0, ms.ptr, Ns)
        ctx.dgemm(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, Nd, Ns, Ns, 1.0,
                  psi.ptr, Nd, q.ptr, Ns, 0.0, rot.ptr, Nd, 0.0)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    ctx.timing_enable(2 * args.steps)
    ctx.synchronize()                       # barrier before the timed region
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()                       # and after it
    wall = time.perf_counter() - t0
    ev = ctx.timing_read(2 * args.steps)
    syrk_ms, rot_ms = ev[0::2], ev[1::2]
    log("timed %d steps in %.3f ms" % (args.steps, wall * 1e3))

    f_syrk = float(Ns) * (Ns + 1) * Nd
    f_rot = 2.0 * Nd * Ns * Ns
    flops = f_syrk + f_rot
    ms_per_step = wall * 1e3 / args.steps
    value = flops / (ms_per_step * 1e-3) / 1e12
    step_dev = syrk_ms + rot_ms
    kernels = {
        "syrk": {"mean_ms": float(np.mean(syrk_ms)), "flop": f_syrk,
                 "bytes": 8.0 * (Ns * Nd + Ns * (Ns + 1) / 2)},
        "rotation": {"mean_ms": float(np.mean(rot_ms)), "flop": f_rot,
                     "bytes": 8.0 * (2.0 * Nd * Ns + Ns * Ns)},
    }
    for k in kernels.values():
        k["tflops"] = k["flop"] / (k["mean_ms"] * 1e-3) / 1e12
        k["frac_fp64_peak"] = k["tflops"] / PEAK_FP64_TFLOPS
    dom = max(kernels, key=lambda k: kernels[k]["mean_ms"])
    dk = kernels[dom]
    roofline = {
        "bound": "mfma",
        "kernel": dom,
        "achieved": dk["tflops"],
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": dk["tflops"] / PEAK_FP64_TFLOPS,
        "traffic": traffic_from_profiles(dom),
    }

    probes = {"fp64_mfma_tflops": ctx.peak_probe(0),
              "int8_mfma_tops": ctx.peak_probe(1),
              "hbm_copy_gbps_1GiB": ctx.copy_probe(1 << 30)}

    extra = {}
    sections = set() if args.no_extra else set(
        x for x in args.extra.split(",") if x)
    if "c1" in sections:
        log("c1 GEMM 1024^3 ...")
        extra["c1_gemm"] = bench_c1(skg, ctx, 20)
    if "split" in sections:
        log("split form (SYRK 65536x1024) ...")
        extra["split_syrk"] = bench_split(skg, ctx, 10)
    if "projection" in sections:
        log("projection GEMM (616x616x175616) ...")
        extra["projection"] = bench_projection(skg, ctx, 10)
    if "nloc" in sections:
        log("nloc projectors (c2 grid) ...")
        extra["nloc"] = bench_nloc(skg, ctx, 20)
    if "stencil" in sections:
        log("stencil (32^3, 520 bands) ...")
        extra["stencil"] = bench_stencil(skg, ctx, 10)
    if "layouts" in sections:
        log("layout conversions (65536x1024) ...")
        extra["layouts"] = bench_layouts(skg, ctx, 10)
    if "batched" in sections or "serving" in sections:
        log("batched (64 systems) ...")
        extra["batched"] = bench_batched(skg, ctx, 5)
    if "serving" in sections:
        log("serving (200 systems, Poisson) ...")
        rate_tflops = extra["batched"]["plain_fp64_graph"]["tflops"]
        extra["serving"] = bench_serving(skg, ctx, rate_tflops,
                                         args.serve_seconds, skg.FP64)
        log("serving, split form ...")
        extra["serving"]["split_form"] = bench_serving(
            skg, ctx, rate_tflops, args.serve_seconds, skg.OZAKI_INT8)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        psi_h = psi.download()
        q_h = q.download()
        rot_h = rot.download().reshape(Ns, Nd).T
        cpu = cpu_baseline(psi_h, q_h, Nd, Ns, ms_per_step, rot_h)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64: Psi normal(0,1/sqrt(Nd)), "
                "Q uniform(-1,1)), generated in HBM",
        "config": {"workload": "DFT subspace step: SYRK Ms=Psi^T Psi (lower) + "
                               "rotation Psi*Q, Psi 32768x520 FP64, plain FP64",
                   "Nd": Nd, "Ns": Ns, "precision": "plain FP64 MFMA"},
        "latency_ms": {"step_p50": float(np.percentile(step_dev, 50)),
                       "step_p99": float(np.percentile(step_dev, 99)),
                       "source": "HIP events per launch, device time"},
        "kernels": kernels,
        "plans": {k: {"tile": v["tile"], "tile_shape": skg.tile_info(v["tile"]),
                      "grid": v["grid"], "kchunks": v["kchunks"]}
                  for k, v in plans.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "probes": probes,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
