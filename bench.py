"""Benchmark of the FP64 GEMM/SYRK hot path on one MI355X.

Headline (the metric's configuration): the overlap SYRK Ms = Psi^T Psi
(lower triangle) on Psi 65536 x 1024 FP64 (column-major, normal(0,
1/sqrt(Nd)) with column j scaled by 10^u_j, u_j ~ U(0, 6); seeded fills,
resident in HBM).  One step = one such SYRK at FP64 accuracy: the split
form (Ozaki scheme II on the INT8 matrix cores) in its certifying mode,
which takes the fewest moduli whose rigorous error bound is at most 1e-13
(or plain FP64 when none is), with the plain FP64 SYRK timed beside it on
the same input.  value = FP64-equivalent TFLOP/s (N(N+1)K FLOP per SYRK)
over the timed steps' wall time; the c5 serving p50 / p99 of both
precisions ride in the same line.

roofline: the dominant stage of the split sequence (residue pass, INT8
products or CRT), its algorithmic bytes or INT8 ops per launch over its
mean duration from HIP events recorded on the launch stream.
cpu_baseline: the oracle's C++/OpenMP restatement of the reference CPU
path (the sme_syrk tile pairs on the P16 layout), built -march=native,
timed on the same SYRK on the host threads of the GPU's share, with its
result compared to both GPU results.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--no-cpu]
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
PEAK_INT8_TOPS = 5000.0   # MI355X dense INT8 matrix (datasheet)
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


def load_workloads():
    spec = importlib.util.spec_from_file_location(
        "skgemm_workloads", os.path.join(PKG_DIR, "workloads.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


wl = load_workloads()


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


def host_threads():
    """Host threads of this GPU's share: OMP_NUM_THREADS when the launcher
    sets it (16 per GPU on the MI355X boxes), else the CPUs this process
    may run on."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else avail
    return min(n, avail), avail


def cpu_baseline(psi_host, Nd, Ns, gpu_plain, gpu_split):
    """Time the oracle's restatement of the reference CPU path on the
    headline SYRK: oracle_syrk_tiles on the P16 layout (syrk_crsp.cpp:306-348,
    include/sme_syrk.hpp:111-149); packing is outside the timed region as in
    the reference (gemm_crsp.cpp:395-405), warm-up and mean as in
    syrk_crsp.cpp:443-462.  The CPU result is compared with both GPU
    results (lower triangle)."""
    sys.path.insert(0, ROOT)
    from oracle import oracle
    tmp = tempfile.mkdtemp(prefix="skgemm_oracle_")
    so = os.path.join(tmp, "liboracle_native.so")
    try:
        oracle.build(native=True, out=so)
    except subprocess.CalledProcessError:
        so = None
    L = oracle.load(so)
    threads, avail = host_threads()
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    P = psi_host.reshape(Ns, Nd)                      # row b = column b of Psi
    t = (Ns + 15) // 16
    p16 = np.zeros((t * 16, Nd))
    p16[:Ns] = P
    p16 = np.ascontiguousarray(p16.reshape(t, 16, Nd).transpose(0, 2, 1)).ravel()
    C = np.zeros(Ns * Ns)
    flops = float(Ns) * (Ns + 1) * Nd

    def step():
        L.oracle_syrk_tiles(Nd, Ns, p16.ctypes.data, C.ctypes.data, 1.0, 0.0,
                            threads)

    step()  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_start < 15.0
                             and len(times) < 50):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
    mean_s = float(np.mean(times))
    # C is column-major upper: C[m1 + m2 Ns] with m1 <= m2, i.e. the
    # row-major lower triangle of C.reshape(Ns, Ns), as on the GPU
    cpu = C.reshape(Ns, Ns)
    il = np.tril_indices(Ns)

    def rel(x):
        x = x.reshape(Ns, Ns)[il]
        return float(np.linalg.norm(x - cpu[il]) / np.linalg.norm(cpu[il]))
    return {
        "value": flops / mean_s / 1e12,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": ("the headline SYRK (Psi 65536x1024, P16 tile pairs) x %d "
                   "after 1 warm-up, %.3f s mean per SYRK, OpenMP threads=%d "
                   "(the GPU's host share; %d CPUs available to the "
                   "process, nproc %d), build %s" %
                   (len(times), mean_s, threads, avail, os.cpu_count() or 0,
                    "-march=native" if so else "x86-64-v3")),
        "rel_diff_gpu_plain_vs_cpu": rel(gpu_plain),
        "rel_diff_gpu_split_vs_cpu": rel(gpu_split),
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


def timed_calls(ctx, fn, steps, warmup):
    """Wall time per call of `steps` calls after `warmup`, bracketed by
    synchronisations, and the per-call device times from the context's
    timing ring (HIP events on the launch stream)."""
    for _ in range(warmup):
        fn()
    ctx.synchronize()
    ctx.timing_enable(steps)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ev = ctx.timing_read(steps)
    ctx.timing_enable(0)
    return wall * 1e3 / steps, ev


def bench_headline(skg, ctx, steps, warmup, moduli):
    """The c3 overlap SYRK: the split form (certifying mode by default) and
    the autotuned plain FP64 SYRK on the same resident Psi, their per-stage
    and per-call device times, and both results for the CPU comparison."""
    Nd, Ns = wl.C3_ND, wl.C3_NS
    psi, q = wl.c3_operands(skg, ctx)
    q.free()
    plain_out = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    split_out = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    f = float(Ns) * (Ns + 1) * Nd
    plan = tuned(ctx, 1, Ns, Ns, Nd, la=skg.COL_MAJOR)

    def plain():
        ctx.dsyrk(skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0, plain_out.ptr,
                  Ns)

    def split(L=moduli, bound=False):
        return ctx.dsyrk_ozaki(skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0,
                               split_out.ptr, Ns, slices=L, bound=bound)
    bound = split(bound=True)
    L = ctx.split_moduli()
    split_ms, split_ev = timed_calls(ctx, split, steps, warmup)
    plain_ms, plain_ev = timed_calls(ctx, plain, steps, warmup)
    # per-stage device times of the same calls
    ctx.timing_stages(True)
    for _ in range(2):
        split()
    ctx.synchronize()
    ctx.timing_enable(3 * steps)
    for _ in range(steps):
        split()
    ctx.synchronize()
    st = ctx.timing_read(3 * steps).reshape(-1, 3)
    ctx.timing_enable(0)
    ctx.timing_stages(False)
    # the fixed-modulus forms beside it
    fixed = {}
    for Lf in (14, 13):
        bf = split(Lf, True)
        t, _ = timed_calls(ctx, lambda: split(Lf), max(5, steps // 2), 2)
        fixed["split_L%d" % Lf] = {
            "ms": t, "tflops_fp64_equivalent": f / (t * 1e-3) / 1e12,
            "speedup_vs_plain": plain_ms / t, "stated_bound": bf,
            "moduli": Lf, "rel_fro_vs_plain_fp64": lower_rel(
                split_out.download(), plain_out.download(), Ns)}
    split(bound=True)                      # the headline result again
    ctx.synchronize()
    res_plain = plain_out.download()
    res_split = split_out.download()
    psi_host = psi.download()
    for x in (psi, plain_out, split_out):
        x.free()
    stage_ms = [float(np.mean(st[:, i])) for i in range(3)]
    # algorithmic work per stage launch
    prep_bytes = 8.0 * Nd * Ns + float(L) * Nd * Ns   # read Psi, write planes
    int8_ops = float(L) * f                           # L products, N(N+1)K ops
    stages = {
        "residue_pass": {"mean_ms": stage_ms[0], "algorithmic_bytes": prep_bytes,
                         "gbps": prep_bytes / (stage_ms[0] * 1e-3) / 1e9,
                         "frac_hbm_peak": prep_bytes / (stage_ms[0] * 1e-3)
                         / 1e9 / PEAK_HBM_GBS},
        "int8_products": {"mean_ms": stage_ms[1], "int8_ops": int8_ops,
                          "tops": int8_ops / (stage_ms[1] * 1e-3) / 1e12,
                          "frac_int8_peak": int8_ops / (stage_ms[1] * 1e-3)
                          / 1e12 / PEAK_INT8_TOPS},
        "crt": {"mean_ms": stage_ms[2]},
    }
    return {
        "split_ms": split_ms, "plain_ms": plain_ms, "flop": f,
        "moduli": L, "stated_bound": bound, "plan_plain": plan,
        "split_device_ms_p50": float(np.percentile(split_ev, 50)),
        "split_device_ms_p99": float(np.percentile(split_ev, 99)),
        "plain_device_ms_p50": float(np.percentile(plain_ev, 50)),
        "plain_device_ms_p99": float(np.percentile(plain_ev, 99)),
        "stages": stages, "fixed": fixed,
        "rel_fro_split_vs_plain": lower_rel(res_split, res_plain, Ns),
        "_plain": res_plain, "_split": res_split, "_psi": psi_host,
    }


def bench_split_rotation(skg, ctx, reps):
    """The c3 rotation Psi Q (65536 x 1024 x 1024): autotuned plain FP64
    against the split form (14 moduli and certifying)."""
    Nd, Ns = wl.C3_ND, wl.C3_NS
    psi, q = wl.c3_operands(skg, ctx)
    rot = skg.DeviceArray(ctx, Nd * Ns)
    rot2 = skg.DeviceArray(ctx, Nd * Ns)
    rplan = tuned(ctx, 0, Nd, Ns, Ns, la=skg.COL_MAJOR, lb=skg.ROW_MAJOR)
    t_rot = median_ms(ctx, lambda: ctx.dgemm(
        skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, Nd, Ns, Ns, 1.0, psi.ptr,
        Nd, q.ptr, Ns, 0.0, rot.ptr, Nd, 0.0), reps)
    fr = 2.0 * Nd * Ns * Ns
    out = {"workload": "rotation Psi*Q, Psi 65536x1024 FP64 col-major with "
                       "10^U(0,6) column scales, Q 1024x1024 uniform(-1,1)",
           "plain_fp64": {"ms": t_rot, "tflops": fr / (t_rot * 1e-3) / 1e12,
                          "frac_fp64_peak": fr / (t_rot * 1e-3) / 1e12 /
                          PEAK_FP64_TFLOPS, "plan": rplan}}
    want = rot.download()
    for name, L in (("split_L14", 14), ("split_certifying", skg.SLICES_AUTO)):
        def run(b=False):
            return ctx.dgemm_ozaki(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR,
                                   Nd, Ns, Ns, 1.0, psi.ptr, Nd, q.ptr, Ns,
                                   0.0, rot2.ptr, Nd, 0.0, slices=L, bound=b)
        b = run(True)
        used = ctx.split_moduli()
        t = median_ms(ctx, run, reps)
        got = rot2.download()
        out[name] = {"ms": t, "tflops_fp64_equivalent": fr / (t * 1e-3) / 1e12,
                     "speedup_vs_plain": t_rot / t, "moduli": used,
                     "stated_bound": b,
                     "rel_fro_vs_plain_fp64": float(
                         np.linalg.norm(got - want) / np.linalg.norm(want))}
        del got
    for x in (psi, q, rot, rot2):
        x.free()
    return out


def bench_c2_step(skg, ctx, steps, warmup, tuning_cache):
    """The c2 DFT subspace step on one grid block (plain FP64): SYRK
    Ms = Psi^T Psi then rotation Psi Q, Psi 32768 x 520 normal(0, 1/sqrt(Nd))
    (seed 1) column-major, Q 520 x 520 uniform(-1, 1) (seed 2)."""
    Nd, Ns = 32768, 520
    psi = skg.DeviceArray(ctx, Nd * Ns)
    q = skg.DeviceArray(ctx, Ns * Ns)
    ms = skg.DeviceArray(ctx, Ns * Ns, zero=True)
    rot = skg.DeviceArray(ctx, Nd * Ns, zero=True)
    ctx.fill_normal(psi.ptr, Nd * Ns, 1, 1.0 / np.sqrt(Nd))
    ctx.fill_uniform(q.ptr, Ns * Ns, 2, -1.0, 1.0)
    plans = {"syrk": tuned(ctx, 1, Ns, Ns, Nd, la=skg.COL_MAJOR),
             "rotation": tuned(ctx, 0, Nd, Ns, Ns, la=skg.COL_MAJOR,
                               lb=skg.ROW_MAJOR)}

    def step():
        ctx.dsyrk(skg.COL_MAJOR, Ns, Nd, 1.0, psi.ptr, Nd, 0.0, ms.ptr, Ns)
        ctx.dgemm(skg.COL_MAJOR, skg.ROW_MAJOR, skg.COL_MAJOR, Nd, Ns, Ns, 1.0,
                  psi.ptr, Nd, q.ptr, Ns, 0.0, rot.ptr, Nd, 0.0)
    for _ in range(warmup):
        step()
    ctx.synchronize()
    ctx.timing_enable(2 * steps)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ev = ctx.timing_read(2 * steps)
    ctx.timing_enable(0)
    for x in (psi, q, ms, rot):
        x.free()
    f_syrk = float(Ns) * (Ns + 1) * Nd
    f_rot = 2.0 * Nd * Ns * Ns
    ms_step = wall * 1e3 / steps
    kernels = {}
    for name, f, t in (("syrk", f_syrk, ev[0::2]), ("rotation", f_rot,
                                                    ev[1::2])):
        m = float(np.mean(t))
        kernels[name] = {"mean_ms": m, "flop": f,
                         "tflops": f / (m * 1e-3) / 1e12,
                         "frac_fp64_peak": f / (m * 1e-3) / 1e12 /
                         PEAK_FP64_TFLOPS, "plan": plans[name]}
    return {"workload": "DFT subspace step: SYRK Ms=Psi^T Psi (lower) + "
                        "rotation Psi*Q, Psi 32768x520 FP64, plain FP64",
            "ms_per_step": ms_step,
            "tflops": (f_syrk + f_rot) / (ms_step * 1e-3) / 1e12,
            "kernels": kernels}


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
    nb bytes); the implementation reads psi once per projector row (the
    repeats of overlapping projectors mostly from L2) and writes and reads
    the forward parts of the 256-row chunks and X (2 x 8 x (parts + columns)
    x nb)."""
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
    parts = sum(-(-r // 256) * c for r, c in zip(nrows, ncols))
    moved = 8.0 * nb * (rows + 2 * points + 2 * parts + 2 * sum(ncols))
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


def bench_batched(skg, ctx, reps):
    probs, keep, flops = wl.c4_batch(skg, ctx)
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
    """200 resident systems (workloads.c5_systems: Ns ~ U[128, 2048], Nd ~
    U[8192, 65536], seed 6, ~68 GB of Psi); overlap and rotation requests
    (arrival stream seed 60) arrive as a Poisson
    process at ~70% of the plain FP64 batched throughput; latency = submit
    -> completion observed by the server.  `precision`: plain FP64 or the
    split form (14 moduli), same arrivals."""
    import threading
    systems = wl.c5_systems(skg, ctx)
    rng = np.random.default_rng(60)         # arrivals, kinds and systems
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


def bench_int8_library(reps, timeout_s=180):
    """The vendor library's INT8 GEMM, measured in a child process: torch's
    own HIP runtime does not find the device in a process where this
    library's runtime already holds it, so the child imports torch alone.
    A failure of the child is recorded, never fatal to the bench."""
    cmd = [sys.executable, os.path.abspath(__file__), "--int8-library-child",
           "--int8-reps", str(reps)]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True,
                           timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"error": "timed out after %d s" % timeout_s}
    if r.returncode != 0:
        return {"error": "child exit %d: %s" % (r.returncode,
                                                 r.stderr.strip()[-300:])}
    try:
        return json.loads(r.stdout.strip().splitlines()[-1])
    except (ValueError, IndexError):
        return {"error": "unreadable child output: %s" % r.stdout[-300:]}


def int8_library_child(reps):
    """The vendor library's INT8 GEMM (torch._int_mm, hipBLASLt on ROCm) on
    shapes beside the split form's products: one 1024 x 1024 x 65536
    product, the MACs of 14 of them stacked (14336 x 1024 x 65536), and a
    large square (8192^3).  A scale for the INT8 roofline fraction, not a
    path of this library."""
    import torch
    out = {"source": "torch._int_mm, int8 x int8 -> int32, A row-major, "
                     "B column-major, uniform int8 operands"}
    g = torch.Generator(device="cuda").manual_seed(5)
    for M, N, K in ((1024, 1024, 65536), (14336, 1024, 65536),
                    (8192, 8192, 8192)):
        name = "%dx%dx%d" % (M, N, K)
        try:
            a = torch.randint(-128, 128, (M, K), dtype=torch.int8,
                              device="cuda", generator=g)
            b = torch.randint(-128, 128, (N, K), dtype=torch.int8,
                              device="cuda", generator=g).t()
            for _ in range(3):
                torch._int_mm(a, b)
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            t0.record()
            for _ in range(reps):
                torch._int_mm(a, b)
            t1.record()
            torch.cuda.synchronize()
            ms = t0.elapsed_time(t1) / reps
            out[name] = {"ms": ms, "tops": 2.0 * M * N * K / (ms * 1e-3) / 1e12}
            del a, b
        except Exception as e:  # the library may lack a kernel for a shape
            out[name] = {"error": str(e)[:200]}
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--moduli", default="auto",
                    help="split-form modulus count for the headline: auto "
                         "(certifying, bound <= 1e-13) or 8..16")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip every section but the headline")
    ap.add_argument("--extra", default="c2,c1,rotation,projection,nloc,"
                                       "stencil,layouts,batched,serving,"
                                       "int8_library",
                    help="comma list of extra sections to run")
    ap.add_argument("--tuning-cache", default=None,
                    help="autotuner cache file: read if it exists (no "
                         "tuning runs), else written by the tuning runs")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--int8-library-child", action="store_true",
                    help=argparse.SUPPRESS)
    ap.add_argument("--int8-reps", type=int, default=10,
                    help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.int8_library_child:
        print(json.dumps(int8_library_child(args.int8_reps)))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(a row-sharded SYRK would add one all-reduce of Ms; see the "
              "design notes).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    skg = load_skgemm()
    ctx = skg.Context(0)
    info = ctx.device_info()
    global TUNING_CACHE
    TUNING_CACHE = args.tuning_cache
    if args.tuning_cache and os.path.exists(args.tuning_cache):
        ctx.load_tuning(args.tuning_cache)
    moduli = skg.SLICES_AUTO if args.moduli == "auto" else int(args.moduli)
    Nd, Ns = wl.C3_ND, wl.C3_NS
    log("device %s, headline: c3 overlap SYRK Nd=%d Ns=%d, split moduli %s"
        % (info, Nd, Ns, args.moduli))
    h = bench_headline(skg, ctx, args.steps, args.warmup, moduli)
    log("split %.4f ms (%d moduli, bound %.3g), plain %.4f ms" %
        (h["split_ms"], h["moduli"], h["stated_bound"], h["plain_ms"]))
    f = h["flop"]
    value = f / (h["split_ms"] * 1e-3) / 1e12
    plain_tflops = f / (h["plain_ms"] * 1e-3) / 1e12
    # Roofline of the dominant stage of the split sequence.
    st = h["stages"]
    dom = max(st, key=lambda k: st[k]["mean_ms"])
    if dom == "int8_products":
        roofline = {"bound": "mfma", "kernel": "split-form INT8 products "
                    "(gemm_i8_kernel, %d moduli)" % h["moduli"],
                    "achieved": st[dom]["tops"], "peak": PEAK_INT8_TOPS,
                    "unit": "TOP/s", "frac": st[dom]["frac_int8_peak"]}
    elif dom == "residue_pass":
        roofline = {"bound": "hbm", "kernel": "split-form norms and residue "
                    "planes", "achieved": st[dom]["gbps"],
                    "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": st[dom]["frac_hbm_peak"]}
    else:
        roofline = {"bound": "hbm", "kernel": "split-form CRT",
                    "achieved": None, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": None}
    roofline["mean_ms"] = st[dom]["mean_ms"]
    roofline["traffic"] = traffic_from_profiles(dom)

    probes = {"fp64_mfma_tflops": ctx.peak_probe(0),
              "int8_mfma_tops": ctx.peak_probe(1),
              "hbm_copy_gbps_1GiB": ctx.copy_probe(1 << 30)}

    extra = {}
    sections = set() if args.no_extra else set(
        x for x in args.extra.split(",") if x)
    def section(key, label, fn):
        # A section that raises is recorded with its error; the headline
        # line is printed whatever an extra does.
        log(label + " ...")
        try:
            extra[key] = fn()
        except Exception as e:  # noqa: BLE001
            log("%s failed: %s" % (key, e))
            extra[key] = {"error": str(e)[:300]}

    if "c2" in sections:
        section("c2_step", "c2 step (SYRK + rotation, 32768 x 520)",
                lambda: bench_c2_step(skg, ctx, 20, 5, args.tuning_cache))
    if "c1" in sections:
        section("c1_gemm", "c1 GEMM 1024^3", lambda: bench_c1(skg, ctx, 20))
    if "rotation" in sections:
        section("c3_rotation", "c3 rotation, plain and split",
                lambda: bench_split_rotation(skg, ctx, 10))
    if "projection" in sections:
        section("projection", "projection GEMM (616x616x175616)",
                lambda: bench_projection(skg, ctx, 10))
    if "nloc" in sections:
        section("nloc", "nloc projectors (c2 grid)",
                lambda: bench_nloc(skg, ctx, 20))
    if "stencil" in sections:
        section("stencil", "stencil (32^3, 520 bands)",
                lambda: bench_stencil(skg, ctx, 10))
    if "layouts" in sections:
        section("layouts", "layout conversions (65536x1024)",
                lambda: bench_layouts(skg, ctx, 10))
    if "int8_library" in sections:
        section("int8_library", "library INT8 GEMM for scale (child process)",
                lambda: bench_int8_library(10))
    if "batched" in sections or "serving" in sections:
        section("batched", "batched (64 systems)",
                lambda: bench_batched(skg, ctx, 5))
    latency = {"c3_split_device_ms_p50": h["split_device_ms_p50"],
               "c3_split_device_ms_p99": h["split_device_ms_p99"],
               "c3_plain_device_ms_p50": h["plain_device_ms_p50"],
               "c3_plain_device_ms_p99": h["plain_device_ms_p99"],
               "source": "HIP events per call on the launch stream"}
    if "serving" in sections and "error" not in extra.get("batched", {}):
        log("serving (200 systems, Poisson), plain FP64 ...")
        rate_tflops = extra["batched"]["plain_fp64_graph"]["tflops"]
        sp = bench_serving(skg, ctx, rate_tflops, args.serve_seconds, skg.FP64)
        log("serving, split form ...")
        ss = bench_serving(skg, ctx, rate_tflops, args.serve_seconds,
                           skg.OZAKI_INT8)
        extra["serving"] = {"plain_fp64": sp, "split_form": ss}
        latency.update({
            "c5_serving_plain_fp64_p50_ms": sp["latency_ms_p50"],
            "c5_serving_plain_fp64_p99_ms": sp["latency_ms_p99"],
            "c5_serving_split_p50_ms": ss["latency_ms_p50"],
            "c5_serving_split_p99_ms": ss["latency_ms_p99"],
            "c5_source": "submit -> completion seen by the server, Poisson "
                         "arrivals at %.0f requests/s" %
                         sp["arrival_rate_per_s"]})

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (the headline SYRK) ...")
        cpu = cpu_baseline(h["_psi"], Nd, Ns, h["_plain"], h["_split"])

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["split_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 in HBM: Psi normal(0,1/sqrt(Nd)) "
                "with column j scaled by 10^u_j, u_j ~ U(0,6))",
        "config": {"workload": "overlap SYRK Ms = Psi^T Psi (lower), Psi "
                               "65536x1024 FP64 column-major with 10^U(0,6) "
                               "column scales, FP64-accurate: split form "
                               "(Ozaki II, INT8 MFMA) certified to 1e-13 vs "
                               "plain FP64 MFMA",
                   "Nd": Nd, "Ns": Ns,
                   "precision": "split form, %d moduli (certifying mode)"
                                % h["moduli"] if moduli == skg.SLICES_AUTO
                                else "split form, %d moduli" % h["moduli"]},
        "plain_vs_split": {
            "plain_fp64": {"ms": h["plain_ms"], "tflops": plain_tflops,
                           "frac_fp64_peak": plain_tflops / PEAK_FP64_TFLOPS,
                           "plan": h["plan_plain"]},
            "split": {"ms": h["split_ms"], "tflops_fp64_equivalent": value,
                      "moduli": h["moduli"], "stated_bound": h["stated_bound"],
                      "speedup_vs_plain": h["plain_ms"] / h["split_ms"],
                      "rel_fro_vs_plain_fp64": h["rel_fro_split_vs_plain"]},
            **h["fixed"],
            "flop_per_syrk": f,
            "flop_reference_convention": float(Ns) * Ns * Nd,
            "tflops_reference_convention_split":
                float(Ns) * Ns * Nd / (h["split_ms"] * 1e-3) / 1e12,
            "stages_ms": st,
        },
        "latency_ms": latency,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "probes": probes,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
