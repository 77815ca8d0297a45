"""Seeded synthetic inputs of the benchmark configurations, generated in
HBM through the C ABI (counter-based SplitMix64 fills), shared by bench.py
and the GPU tests so both run exactly the same problems.

  c3  Psi 65536 x 1024 FP64 column-major, normal(0, 1/sqrt(Nd)) (seed 3),
      column j scaled by 10^u_j, u_j ~ U(0, 6) (seed 4); Q 1024 x 1024
      uniform(-1, 1) row-major (seed 4).
  c4  64 systems: Ns ~ U{128, 136, ..., 1536}, Nd ~ U{4096, 13824, 32768,
      65536} (numpy seed 5), Psi ~ N(0, 1/sqrt(Nd)) (fill seed 1000 + i),
      Q ~ U(-1, 1) (fill seed 2000 + i); one overlap SYRK and one rotation
      GEMM per system.
  c5  200 resident systems: Ns ~ U[128, 2048], Nd ~ U[8192, 65536] (numpy
      seed 6), Psi ~ N(0, 1/sqrt(Nd)) (fill seed 3000 + i), Q ~ U(-1, 1)
      (fill seed 4000 + i).
"""
import numpy as np

C3_ND, C3_NS = 65536, 1024


def c3_operands(skg, ctx):
    """(psi, q) device arrays of the c3 workload."""
    Nd, Ns = C3_ND, C3_NS
    psi = skg.DeviceArray(ctx, Nd * Ns)
    ctx.fill_normal(psi.ptr, Nd * Ns, 3, 1.0 / np.sqrt(Nd))
    ctx.scale_columns_pow10(psi.ptr, Nd, Ns, Nd, 4, 6.0)
    q = skg.DeviceArray(ctx, Ns * Ns)
    ctx.fill_uniform(q.ptr, Ns * Ns, 4, -1.0, 1.0)
    return psi, q


def c4_shapes(seed=5, count=64):
    """(Nd, Ns) of the c4 systems, in batch order."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        Ns = int(rng.choice(np.arange(128, 1537, 8)))
        Nd = int(rng.choice([4096, 13824, 32768, 65536]))
        out.append((Nd, Ns))
    return out


def c4_batch(skg, ctx, seed=5, count=64):
    """The c4 batch: problems (SYRK then rotation per system), the device
    arrays behind them [psi, q, ms, rot] per system, total FLOP."""
    probs, keep, flops = [], [], 0.0
    for i, (Nd, Ns) in enumerate(c4_shapes(seed, count)):
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


def c5_shapes(seed=6, count=200):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        Ns = int(rng.integers(128, 2049))
        Nd = int(rng.integers(8192, 65537))
        out.append((Nd, Ns))
    return out


def c5_systems(skg, ctx, seed=6, count=200, pick=None):
    """Resident systems of the c5 workload as (psi, q, Nd, Ns); `pick`
    (indices into the c5 list) selects a subset with the same seeds."""
    shapes = c5_shapes(seed, count)
    systems = []
    for i in (range(count) if pick is None else pick):
        Nd, Ns = shapes[i]
        psi = skg.DeviceArray(ctx, Nd * Ns)
        ctx.fill_normal(psi.ptr, Nd * Ns, 3000 + i, 1.0 / np.sqrt(Nd))
        q = skg.DeviceArray(ctx, Ns * Ns)
        ctx.fill_uniform(q.ptr, Ns * Ns, 4000 + i, -1.0, 1.0)
        systems.append((psi, q, Nd, Ns))
    ctx.synchronize()
    return systems
