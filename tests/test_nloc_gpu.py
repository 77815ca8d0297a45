"""GPU parity of the non-local projector operator (skgemm_nloc_*) against
the oracle's restatement of include/nloc_packed_opt2.hpp and the reference
test's unpacked formulation: the reference test's five configurations for
both output packings, shared grid points, repeated atoms, empty projectors,
points listed twice by one projector, crowded points, band counts that
leave partial band groups (bands past nb untouched), a c2-sized case (Nd
32768, 520 bands), determinism and argument errors.
Bar: the added term (result after - before) within 1e-13 relative Frobenius
of the oracle's."""
import numpy as np
import pytest

from conftest import rel_fro
from test_oracle import (NLOC_CONFIGS, band_pack_cm, band_unpack_cm,
                         nloc_projectors, nloc_unpacked)

pytestmark = pytest.mark.gpu
TOL = 1e-13


def run_gpu(skg, ctx, K, nb, dv, proj, psi16, res0, tile):
    nrows, ncols, atoms, idx, chi, gam = proj
    h = skg.Nloc(ctx, K, nrows, ncols, atoms, idx, chi, gam)
    try:
        d_psi = skg.DeviceArray.from_host(ctx, psi16)
        d_res = skg.DeviceArray.from_host(ctx, res0)
        h.apply(nb, dv, d_psi.ptr, d_res.ptr, tile)
        ctx.synchronize()
        return d_res.download()
    finally:
        h.close()


def oracle(orc, K, nb, dv, proj, psi16, res0, tile):
    nrows, ncols, atoms, idx, chi, gam = proj
    out = res0.copy()
    orc.nloc_apply(K, nb, dv, nrows, ncols, atoms, idx, chi, gam, psi16, out,
                   tile)
    return out


@pytest.mark.parametrize("cfg", NLOC_CONFIGS)
@pytest.mark.parametrize("tile", [32, 16])
def test_reference_configurations(skg, ctx, orc, cfg, tile):
    """nloc_packed_opt2_test.cpp:289-296 configurations (K, nb, nrow, ncol,
    projectors), dv = 0.01, random psi and a non-zero starting result."""
    K, nb, nrow, ncol, nproj = cfg
    print("cfg", cfg, "tile", tile, flush=True)
    rng = np.random.default_rng(K + nb + tile)
    proj = nloc_projectors(rng, K, nrow, ncol, nproj)
    psi = rng.uniform(-1, 1, K * nb)
    res0 = band_pack_cm(rng.uniform(-1, 1, K * nb), K, nb, tile)
    psi16 = band_pack_cm(psi, K, nb, 16)
    got = run_gpu(skg, ctx, K, nb, 0.01, proj, psi16, res0, tile)
    want = oracle(orc, K, nb, 0.01, proj, psi16, res0, tile)
    assert rel_fro(got - res0, want - res0) < TOL
    unpacked = nloc_unpacked(K, nb, 0.01, *proj, psi,
                             band_unpack_cm(res0, K, nb, tile))
    assert np.max(np.abs(band_unpack_cm(got, K, nb, tile) - unpacked)) < 1e-12


def test_shared_points_repeated_atoms_empty_projectors(skg, ctx, orc):
    K, nb = 700, 40
    rng = np.random.default_rng(21)
    nrows = [120, 0, 80, 120, 50]
    ncols = [7, 5, 0, 9, 3]
    atoms = [3, 3, 4, 4, 9]
    idx = np.concatenate([rng.permutation(K)[:n] for n in nrows])
    idx[200:220] = idx[:20]                     # points shared by 0 and 3
    chi = rng.uniform(-1, 1, sum(r * c for r, c in zip(nrows, ncols)))
    gam = rng.uniform(-1, 1, sum(ncols))
    proj = (nrows, ncols, atoms, idx, chi, gam)
    psi16 = band_pack_cm(rng.uniform(-1, 1, K * nb), K, nb, 16)
    res0 = band_pack_cm(rng.uniform(-1, 1, K * nb), K, nb, 32)
    got = run_gpu(skg, ctx, K, nb, 0.25, proj, psi16, res0, 32)
    want = oracle(orc, K, nb, 0.25, proj, psi16, res0, 32)
    assert rel_fro(got - res0, want - res0) < TOL


def test_repeated_rows_and_crowded_points(skg, ctx, orc):
    """A projector that lists one grid point several times (each row adds in
    row order), 40 projectors of 16-20 columns sharing one point (more rows
    than a block holds points), and runs of consecutive grid points longer
    than a block."""
    K, nb = 300, 70
    rng = np.random.default_rng(5)
    nrows, ncols, atoms, idx = [], [], [], []
    rep = rng.permutation(K)[:30]
    rep[[3, 9, 17]] = rep[0]                 # point rep[0] four times
    nrows.append(30); ncols.append(7); atoms.append(1); idx.append(rep)
    for p in range(40):                       # all of them touch point 7
        pts = rng.permutation(K)[:25]
        pts[pts == 7] = (7 + 1 + p) % K
        pts[p % 25] = 7
        nrows.append(25); ncols.append(16 + p % 5); atoms.append(2 + p // 3)
        idx.append(pts)
    nrows.append(K); ncols.append(3); atoms.append(99); idx.append(np.arange(K))
    idx = np.concatenate(idx)
    chi = rng.uniform(-1, 1, sum(r * c for r, c in zip(nrows, ncols)))
    gam = rng.uniform(-1, 1, sum(ncols))
    proj = (nrows, ncols, atoms, idx, chi, gam)
    for tile in (32, 16):
        psi16 = band_pack_cm(rng.uniform(-1, 1, K * nb), K, nb, 16)
        res0 = band_pack_cm(rng.uniform(-1, 1, K * nb), K, nb, tile)
        got = run_gpu(skg, ctx, K, nb, 0.5, proj, psi16, res0, tile)
        want = oracle(orc, K, nb, 0.5, proj, psi16, res0, tile)
        assert rel_fro(got - res0, want - res0) < TOL


@pytest.mark.parametrize("nb", [1, 3, 17, 65, 100, 130])
def test_odd_band_counts_random_projectors(skg, ctx, orc, nb):
    """Band counts that leave partial 16- and 64-band groups, random
    projector sets with 1..24 columns (past 16: the extra column pass),
    repeated atoms and grid points listed twice by one projector."""
    K = 600
    rng = np.random.default_rng(1000 + nb)
    count = 9
    nrows = [int(rng.integers(1, 200)) for _ in range(count)]
    ncols = [int(rng.integers(1, 25)) for _ in range(count)]
    atoms = sorted(int(a) for a in rng.integers(1, 5, count))
    idx = []
    for n in nrows:
        pts = rng.integers(0, K, n)          # repeats within a projector too
        idx.append(pts)
    idx = np.concatenate(idx)
    chi = rng.uniform(-1, 1, sum(r * c for r, c in zip(nrows, ncols)))
    gam = rng.uniform(-1, 1, sum(ncols))
    proj = (nrows, ncols, atoms, idx, chi, gam)
    for tile in (32, 16):
        print("nb", nb, "tile", tile, flush=True)
        psi16 = band_pack_cm(rng.uniform(-1, 1, K * nb), K, nb, 16)
        res0 = band_pack_cm(rng.uniform(-1, 1, K * nb), K, nb, tile)
        got = run_gpu(skg, ctx, K, nb, 0.3, proj, psi16, res0, tile)
        want = oracle(orc, K, nb, 0.3, proj, psi16, res0, tile)
        assert rel_fro(got - res0, want - res0) < TOL
        # bands past nb in the last packed tile are not written
        pad = np.ones(((nb + tile - 1) // tile, K, tile), bool)
        for b in range(nb):
            pad[b // tile, :, b % tile] = False
        assert np.array_equal(got[pad.ravel()], res0[pad.ravel()])


def c2_projectors(rng, K, count):
    """Compact projector spheres: nrow 512..1024 consecutive-ish grid
    indices around random centres, ncol in {9, 13, 16}."""
    nrows = [int(rng.integers(512, 1025)) for _ in range(count)]
    ncols = [int(rng.choice([9, 13, 16])) for _ in range(count)]
    idx = []
    for n in nrows:
        c = int(rng.integers(0, K))
        pts = (c + np.arange(-(n // 2), n - n // 2) * 3) % K
        idx.append(np.unique(pts)[:n])
    nrows = [len(x) for x in idx]
    chi = rng.uniform(-1, 1, sum(r * c for r, c in zip(nrows, ncols)))
    gam = rng.uniform(-1, 1, sum(ncols))
    atoms = list(range(count))
    return nrows, ncols, atoms, np.concatenate(idx), chi, gam


def test_c2_sized_case_and_determinism(skg, ctx, orc):
    """Nd 32768, 520 bands, 24 projectors; two applications give the same
    bits."""
    K, nb = 32768, 520
    rng = np.random.default_rng(33)
    proj = c2_projectors(rng, K, 24)
    psi16 = band_pack_cm(rng.standard_normal(K * nb) / np.sqrt(K), K, nb, 16)
    res0 = np.zeros(((nb + 31) // 32) * K * 32)
    h = skg.Nloc(ctx, K, *proj)
    try:
        d_psi = skg.DeviceArray.from_host(ctx, psi16)
        r1 = skg.DeviceArray.from_host(ctx, res0)
        r2 = skg.DeviceArray.from_host(ctx, res0)
        h.apply(nb, 0.02, d_psi.ptr, r1.ptr, 32)
        h.apply(nb, 0.02, d_psi.ptr, r2.ptr, 32)
        ctx.synchronize()
        a, b = r1.download(), r2.download()
    finally:
        h.close()
    assert np.array_equal(a, b)
    want = oracle(orc, K, nb, 0.02, proj, psi16, res0, 32)
    assert rel_fro(a, want) < TOL


def test_argument_errors(skg, ctx):
    with pytest.raises(skg.SkgemmError) as e:       # index out of range
        skg.Nloc(ctx, 10, [2], [1], [1], [3, 10], [1.0, 1.0], [1.0])
    assert e.value.code == skg.ERR_ARG
    h = skg.Nloc(ctx, 10, [2], [1], [1], [3, 4], [1.0, 1.0], [1.0])
    try:
        psi = skg.DeviceArray(ctx, 10 * 16, zero=True)
        res = skg.DeviceArray(ctx, 10 * 32, zero=True)
        with pytest.raises(skg.SkgemmError) as e:   # bad packing
            h.apply(16, 1.0, psi.ptr, res.ptr, 8)
        assert e.value.code == skg.ERR_ARG
        with pytest.raises(skg.SkgemmError) as e:   # host memory
            h.apply(16, 1.0, np.zeros(160).ctypes.data, res.ptr, 32)
        assert e.value.code == skg.ERR_ARG
    finally:
        h.close()
