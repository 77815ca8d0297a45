"""GPU parity of the real-space stencil and the single-domain halo fill
(skgemm_stencil_apply, skgemm_halo_extend) against the oracle's restatement
of include/stencil_extended_packed.hpp and include/halo_packed.hpp: the
reference test's grid and 12th-order coefficients, every FDn 1..8, both
output packings, zero and periodic halos, the c2 grid (32^3, 520 bands).
Bar: halo bit-exact; stencil within 1e-13 relative Frobenius of the oracle
(the device fuses each term's multiply-add)."""
import numpy as np
import pytest

from conftest import rel_fro
from test_oracle import STENCIL_COEF, grid_pack16

pytestmark = pytest.mark.gpu
TOL = 1e-13


def run(skg, ctx, orc, ni, nj, nk, nb, F, periodic, tile, cx, cy, cz, c0,
        seed):
    rng = np.random.default_rng(seed)
    K = ni * nj * nk
    K_ex = (ni + 2 * F) * (nj + 2 * F) * (nk + 2 * F)
    t16 = (nb + 15) // 16
    local = grid_pack16(rng.uniform(-1, 1, (nb, nk, nj, ni)), nb)
    A = rng.uniform(-1, 1, K)
    d_loc = skg.DeviceArray.from_host(ctx, local)
    d_ext = skg.DeviceArray(ctx, t16 * K_ex * 16, zero=True)
    d_A = skg.DeviceArray.from_host(ctx, A)
    n_out = ((nb + tile - 1) // tile) * K * tile
    sentinel = np.full(n_out, -3.5)
    d_out = skg.DeviceArray.from_host(ctx, sentinel)
    ctx.halo_extend(ni, nj, nk, nb, F, d_loc.ptr, d_ext.ptr, periodic)
    ctx.stencil_apply(ni, nj, nk, nb, F, d_ext.ptr, d_A.ptr, cx, cy, cz, c0,
                      d_out.ptr, tile)
    ctx.synchronize()
    ext_want = orc.halo_extend(ni, nj, nk, nb, F, local, periodic)
    assert np.array_equal(d_ext.download(), ext_want)
    want = orc.stencil_apply(ni, nj, nk, nb, F, ext_want, A, cx, cy, cz, c0,
                             tile, out=sentinel.copy())
    got = d_out.download()
    assert np.array_equal(got == -3.5, want == -3.5)   # padding untouched
    assert rel_fro(got, want) < TOL
    return got


@pytest.mark.parametrize("tile", [32, 16])
def test_reference_grid_and_coefficients(skg, ctx, orc, tile):
    """stencil_extended_test.cpp:136-153: 12 x 10 x 8 grid, 48 bands,
    FDn 6, coef 0 -1/12 2/3 -5/2 8/3 -1/4 1/30 on all three axes."""
    c = STENCIL_COEF
    run(skg, ctx, orc, 12, 10, 8, 48, 6, True, tile, c, c, c, 0.75, 11)


@pytest.mark.parametrize("F", range(1, 9))
def test_every_fdn_zero_halo(skg, ctx, orc, F):
    rng = np.random.default_rng(F)
    cx, cy, cz = (rng.uniform(-1, 1, F + 1) for _ in range(3))
    print("FDn", F, flush=True)
    run(skg, ctx, orc, 7, 5, 9, 40, F, False, 32, cx, cy, cz, -0.5, 20 + F)


def test_c2_grid(skg, ctx, orc):
    """32^3 grid (Nd = 32768), 520 bands, FDn 6, periodic."""
    c = STENCIL_COEF
    run(skg, ctx, orc, 32, 32, 32, 520, 6, True, 32, c, c, c, 1.25, 5)


def test_argument_errors(skg, ctx):
    a = skg.DeviceArray(ctx, 4096, zero=True)
    b = skg.DeviceArray(ctx, 1 << 16, zero=True)
    c = np.zeros(10)
    with pytest.raises(skg.SkgemmError) as e:       # FDn out of range
        ctx.stencil_apply(4, 4, 4, 16, 9, b.ptr, a.ptr, c, c, c, 0.0, a.ptr)
    assert e.value.code == skg.ERR_ARG
    with pytest.raises(skg.SkgemmError) as e:       # bad packing
        ctx.stencil_apply(4, 4, 4, 16, 1, b.ptr, a.ptr, c, c, c, 0.0, a.ptr, 8)
    assert e.value.code == skg.ERR_ARG
    with pytest.raises(skg.SkgemmError) as e:       # host memory
        ctx.halo_extend(4, 4, 4, 16, 1, np.zeros(1024).ctypes.data, b.ptr)
    assert e.value.code == skg.ERR_ARG


def test_interior_copy_leaves_ghosts(skg, ctx, orc, rng):
    """Boundary 2 (copy_interior_to_extended_packed, halo_packed.hpp:86-123):
    the interior is copied, every ghost point keeps its previous value."""
    ni, nj, nk, nb, F = 6, 5, 4, 20, 2
    local = grid_pack16(rng.uniform(-1, 1, (nb, nk, nj, ni)), nb)
    K_ex = (ni + 2 * F) * (nj + 2 * F) * (nk + 2 * F)
    before = np.full(((nb + 15) // 16) * K_ex * 16, 7.0)
    d_loc = skg.DeviceArray.from_host(ctx, local)
    d_ext = skg.DeviceArray.from_host(ctx, before)
    ctx.halo_extend(ni, nj, nk, nb, F, d_loc.ptr, d_ext.ptr, boundary=2)
    ctx.synchronize()
    got = d_ext.download()
    zero = orc.halo_extend(ni, nj, nk, nb, F, local, False)
    e = np.zeros((nk + 2 * F, nj + 2 * F, ni + 2 * F), bool)
    e[F:F + nk, F:F + nj, F:F + ni] = True
    inside = np.repeat(np.tile(e.ravel(), (nb + 15) // 16), 16)
    assert np.array_equal(got[inside], zero[inside])
    assert np.all(got[~inside] == 7.0)


@pytest.mark.parametrize("shape", [(1, 1, 1, 3, 1), (3, 2, 5, 17, 3),
                                   (50, 3, 2, 16, 4), (7, 5, 9, 40, 8)])
@pytest.mark.parametrize("offset", [0, 1])
def test_halo_shapes_modes_alignment(skg, ctx, orc, shape, offset):
    """Every boundary mode on planes smaller than one workgroup's share,
    ghost widths beyond the grid (periodic wraps more than once), and grids
    one double off 16-byte alignment (offset 1: the one-double-per-lane
    kernel; offset 0: the 16-byte kernel)."""
    ni, nj, nk, nb, F = shape
    print("halo", shape, "offset", offset, flush=True)
    rng = np.random.default_rng(sum(shape) + offset)
    t = (nb + 15) // 16
    local = rng.uniform(-1, 1, t * ni * nj * nk * 16)
    n_ext = t * (ni + 2 * F) * (nj + 2 * F) * (nk + 2 * F) * 16
    e = np.zeros((nk + 2 * F, nj + 2 * F, ni + 2 * F), bool)
    e[F:F + nk, F:F + nj, F:F + ni] = True
    inside = np.repeat(np.tile(e.ravel(), t), 16)
    d_loc = skg.DeviceArray.from_host(ctx, np.concatenate([[0.0] * offset,
                                                           local]))
    for mode in (0, 1, 2):
        d_ext = skg.DeviceArray.from_host(ctx, np.full(n_ext + offset, 7.0))
        ctx.halo_extend(ni, nj, nk, nb, F, d_loc.ptr + 8 * offset,
                        d_ext.ptr + 8 * offset, boundary=mode)
        ctx.synchronize()
        got = d_ext.download()[offset:]
        want = orc.halo_extend(ni, nj, nk, nb, F, local, mode == 1)
        if mode == 2:
            assert np.array_equal(got[inside], want[inside])
            assert np.all(got[~inside] == 7.0)
        else:
            assert np.array_equal(got, want), mode


@pytest.mark.parametrize("shape", [(7, 5, 9, 33, 6), (9, 17, 3, 17, 4),
                                   (16, 8, 17, 31, 2)])
@pytest.mark.parametrize("tile", [32, 16])
def test_odd_bands_short_chunks(skg, ctx, orc, shape, tile):
    """Odd band counts (a last band pair with one band), grids that leave
    partial 8 x 8 blocks and a last k chunk shorter than 8 planes."""
    ni, nj, nk, nb, F = shape
    print("stencil", shape, "tile", tile, flush=True)
    rng = np.random.default_rng(sum(shape) + tile)
    cx, cy, cz = (rng.uniform(-1, 1, F + 1) for _ in range(3))
    run(skg, ctx, orc, ni, nj, nk, nb, F, True, tile, cx, cy, cz, 0.3,
        sum(shape))


@pytest.mark.parametrize("F", [2, 6])
def test_aligned_and_unaligned_kernels_agree(skg, ctx, orc, F):
    """16-byte aligned grids and outputs take the band-pair kernel, grids
    one double off alignment the one-double kernel: the same terms in the
    same order, so the two give identical bits (and the oracle's values)."""
    ni, nj, nk, nb, tile = 10, 9, 11, 37, 32
    rng = np.random.default_rng(70 + F)
    cx, cy, cz = (rng.uniform(-1, 1, F + 1) for _ in range(3))
    local = grid_pack16(rng.uniform(-1, 1, (nb, nk, nj, ni)), nb)
    K = ni * nj * nk
    A = rng.uniform(-1, 1, K)
    t16 = (nb + 15) // 16
    n_ext = t16 * (ni + 2 * F) * (nj + 2 * F) * (nk + 2 * F) * 16
    n_out = ((nb + tile - 1) // tile) * K * tile
    d_A = skg.DeviceArray.from_host(ctx, A)
    outs = []
    for off in (0, 1):
        print("FDn", F, "offset", off, flush=True)
        d_loc = skg.DeviceArray.from_host(ctx, np.concatenate([[0.0] * off,
                                                               local]))
        d_ext = skg.DeviceArray(ctx, n_ext + off, zero=True)
        d_out = skg.DeviceArray.from_host(ctx, np.full(n_out + off, -3.5))
        ctx.halo_extend(ni, nj, nk, nb, F, d_loc.ptr + 8 * off,
                        d_ext.ptr + 8 * off, True)
        ctx.stencil_apply(ni, nj, nk, nb, F, d_ext.ptr + 8 * off, d_A.ptr,
                          cx, cy, cz, 0.4, d_out.ptr + 8 * off, tile)
        ctx.synchronize()
        outs.append(d_out.download()[off:])
    assert np.array_equal(outs[0], outs[1])
    ext_want = orc.halo_extend(ni, nj, nk, nb, F, local, True)
    want = orc.stencil_apply(ni, nj, nk, nb, F, ext_want, A, cx, cy, cz, 0.4,
                             tile, out=np.full(n_out, -3.5))
    assert np.array_equal(outs[0] == -3.5, want == -3.5)
    assert rel_fro(outs[0], want) < TOL
