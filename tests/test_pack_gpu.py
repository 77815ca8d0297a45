"""GPU parity of the layout conversions (skgemm_pack, skgemm_transpose)
against the oracle's restatement of include/sme_pack.hpp and numpy index
maps: bit-exact, elements outside the matrix left untouched, ragged extents,
every layout pair, the band <-> grid conversions of the rotation's operands
at the full 65536 x 1024 shape."""
import numpy as np
import pytest

from test_oracle import band_pack, decode_d, encode_b, encode_d, grid_pack

pytestmark = pytest.mark.gpu
SENTINEL = -7.25


def expected_buffer(X, layout, pad):
    """Buffer `layout` holding X (D convention) with SENTINEL everywhere the
    matrix does not reach, and its ld."""
    R, O = X.shape
    idx = np.arange(1, R * O + 1, dtype=np.float64).reshape(R, O)
    enc, ld = encode_d(idx, layout, pad)
    out = np.full(enc.shape, SENTINEL)
    hit = enc > 0
    out[hit] = X.ravel()[enc[hit].astype(np.int64) - 1]
    return out, ld


def run_convert(skg, ctx, swap, X, ls, ldst, pad_src=3, pad_dst=2, off=0):
    """X is R x O (element (r, o)); src is encoded in the B convention.
    off: both buffers start `off` doubles into their allocations."""
    R, O = X.shape
    src, ld_s = encode_b(X, ls, pad_src)
    want, ld_d = expected_buffer(X.T if swap else X, ldst, pad_dst)
    d_src = skg.DeviceArray.from_host(ctx, np.concatenate([[0.0] * off, src]))
    d_dst = skg.DeviceArray.from_host(ctx, np.full(want.size + off,
                                                   SENTINEL))
    f = ctx.transpose if swap else ctx.pack
    f(ls, d_src.ptr + 8 * off, ld_s, ldst, d_dst.ptr + 8 * off, ld_d, O, R)
    ctx.synchronize()
    got = d_dst.download()
    assert np.all(got[:off] == SENTINEL)
    return got[off:], want


@pytest.mark.parametrize("swap", [False, True])
@pytest.mark.parametrize("shape", [(45, 70), (130, 300), (64, 64), (1, 17)])
@pytest.mark.parametrize("pads", [(3, 2), (2, 4)])
def test_every_layout_pair(skg, ctx, rng, swap, shape, pads):
    """Odd and even leading dimensions (pads): 16-byte pairs on neither,
    one or both sides of the conversion; odd extents leave single
    elements at the edges."""
    X = rng.uniform(-1, 1, shape)
    for ls in range(4):
        for ldst in range(4):
            print("swap", swap, "shape", shape, "pads", pads, "layouts", ls,
                  ldst, flush=True)
            got, want = run_convert(skg, ctx, swap, X, ls, ldst, *pads)
            assert np.array_equal(got, want), (ls, ldst)


@pytest.mark.parametrize("swap", [False, True])
def test_unaligned_buffers(skg, ctx, rng, swap):
    """Buffers one double off 16-byte alignment take the one-element path."""
    X = rng.uniform(-1, 1, (70, 45))
    for ls in range(4):
        for ldst in range(4):
            print("unaligned swap", swap, "layouts", ls, ldst, flush=True)
            got, want = run_convert(skg, ctx, swap, X, ls, ldst, 2, 4, off=1)
            assert np.array_equal(got, want), (ls, ldst)


def test_band_grid_known_answer(skg, ctx, orc):
    """pack.cpp:190-258: value[g, b] = g + 1000 b, band -> grid -> band."""
    for Nd, nb in ((64, 48), (4096, 256)):
        print("band/grid", Nd, nb, flush=True)
        g, b = np.meshgrid(np.arange(Nd), np.arange(nb), indexing="ij")
        X = (g + 1000.0 * b)
        band = band_pack(X)
        d_band = skg.DeviceArray.from_host(ctx, band)
        d_grid = skg.DeviceArray(ctx, band.size, zero=True)
        d_back = skg.DeviceArray(ctx, band.size, zero=True)
        ctx.transpose(skg.PACK16, d_band.ptr, Nd, skg.PACK16, d_grid.ptr, nb,
                      nb, Nd)
        ctx.transpose(skg.PACK16, d_grid.ptr, nb, skg.PACK16, d_back.ptr, Nd,
                      Nd, nb)
        ctx.synchronize()
        grid = d_grid.download()
        assert np.array_equal(grid, orc.band_to_grid(Nd, nb, band))
        assert np.array_equal(grid, grid_pack(X))
        assert np.array_equal(d_back.download(), band)


def test_band_grid_ragged(skg, ctx, rng):
    """Extents that are not multiples of 16 (the reference requires them):
    padded panels are neither read nor written."""
    Nd, nb = 100, 40
    X = rng.uniform(-1, 1, (Nd, nb))
    band, ld_b = encode_b(X, 2, 0)                 # (r = g, o = b), ld = Nd
    d_band = skg.DeviceArray.from_host(ctx, band)
    # grid-packed is the D-convention PACK16 of X^T: element (m = b, n = g)
    want, ld_g = expected_buffer(X.T, 2, 0)
    d_grid = skg.DeviceArray.from_host(ctx, np.full(want.shape, SENTINEL))
    ctx.transpose(skg.PACK16, d_band.ptr, ld_b, skg.PACK16, d_grid.ptr, ld_g,
                  nb, Nd)
    ctx.synchronize()
    got = d_grid.download()
    assert np.array_equal(got, want)
    assert np.array_equal(decode_d(got, 2, nb, Nd, ld_g), X.T)


def test_many_row_blocks(skg, ctx, rng):
    """More than 65535 blocks of 64 along o (the grid's y loop)."""
    O, R = 65535 * 64 + 100, 3
    X = rng.uniform(-1, 1, (R, O))
    src = X.ravel().copy()                          # ROW_MAJOR: r*O + o
    d_src = skg.DeviceArray.from_host(ctx, src)
    d_dst = skg.DeviceArray(ctx, O * R, zero=True)
    ctx.pack(skg.ROW_MAJOR, d_src.ptr, O, skg.COL_MAJOR, d_dst.ptr, R, O, R)
    ctx.synchronize()
    assert np.array_equal(d_dst.download().reshape(O, R), X.T)


def test_rotation_operands_full_size(skg, ctx, rng):
    """Nd x nb = 65536 x 1024 (the SYRK / rotation shape): column-major ->
    grid-packed -> band-packed -> column-major, each stage checked against
    numpy's reshapes of the same matrix."""
    Nd, nb = 65536, 1024
    X = rng.standard_normal((Nd, nb))
    col = X.ravel(order="F").copy()
    d_col = skg.DeviceArray.from_host(ctx, col)
    d_grid = skg.DeviceArray(ctx, Nd * nb, zero=True)
    d_band = skg.DeviceArray(ctx, Nd * nb, zero=True)
    d_back = skg.DeviceArray(ctx, Nd * nb, zero=True)
    # column-major [Nd, nb]: element (r = b, o = g) at r*Nd + o
    ctx.pack(skg.ROW_MAJOR, d_col.ptr, Nd, skg.PACK16, d_grid.ptr, nb, Nd, nb)
    ctx.transpose(skg.PACK16, d_grid.ptr, nb, skg.PACK16, d_band.ptr, Nd, Nd,
                  nb)
    ctx.transpose(skg.PACK16, d_band.ptr, Nd, skg.ROW_MAJOR, d_back.ptr, Nd,
                  nb, Nd)
    ctx.synchronize()
    assert np.array_equal(d_grid.download(), grid_pack(X))
    assert np.array_equal(d_band.download(), band_pack(X))
    assert np.array_equal(d_back.download(), col)


def test_argument_errors(skg, ctx):
    a = skg.DeviceArray(ctx, 64 * 64, zero=True)
    b = skg.DeviceArray(ctx, 64 * 64, zero=True)
    with pytest.raises(skg.SkgemmError) as e:      # ld_dst < o_extent of src
        ctx.transpose(skg.ROW_MAJOR, a.ptr, 64, skg.COL_MAJOR, b.ptr, 32, 64, 32)
    assert e.value.code == skg.ERR_ARG
    with pytest.raises(skg.SkgemmError) as e:      # host pointer
        ctx.transpose(skg.ROW_MAJOR, np.zeros(64 * 64).ctypes.data, 64,
                      skg.COL_MAJOR, b.ptr, 64, 64, 64)
    assert e.value.code == skg.ERR_ARG
    with pytest.raises(skg.SkgemmError) as e:
        ctx.pack(7, a.ptr, 64, skg.COL_MAJOR, b.ptr, 64, 64, 64)
    assert e.value.code == skg.ERR_ARG
