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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true")
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
    plans = {}
    if args.tile >= 0 or args.grid > 0:
        ctx.set_tuning(args.tile, args.grid)
    elif not args.no_tune:
        # Autotuner: picks tile / grid / K chunks per shape with real
        # launches, before (and outside) the timed region.
        plans["syrk"] = ctx.autotune(1, Ns, Ns, Nd, la=skg.COL_MAJOR)
        plans["rotation"] = ctx.autotune(0, Nd, Ns, Ns, la=skg.COL_MAJOR,
                                         lb=skg.ROW_MAJOR)
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
              "int8_mfma_tops": ctx.peak_probe(1)}

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
    print(json.dumps(out))


if __name__ == "__main__":
    main()
