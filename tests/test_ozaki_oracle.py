"""Pins the split-form (Ozaki scheme II) restatement in oracle/ozaki.py.
CPU only."""
import numpy as np
import pytest

from conftest import rel_fro
from oracle import ozaki


def c3_like(rng, K, N):
    """c3 inputs at reduced size: normal(0, 1/sqrt(K)) columns scaled by
    10^u, u ~ U(0, 6)."""
    P = rng.standard_normal((K, N)) / np.sqrt(K)
    return P * 10.0 ** rng.uniform(0, 6, N)[None, :]


def test_moduli_pairwise_coprime_and_headroom():
    import math
    m = ozaki.MODULI
    for i in range(len(m)):
        for j in range(i):
            assert math.gcd(m[i], m[j]) == 1
    _, h = ozaki.tables(14)
    assert h == 55  # log2(P) = 110.16 for the first 14 moduli
    # every modulus count keeps 2^(2h-2) below P / 4 with room for the
    # rounding of the scaled rows
    for L in range(8, 17):
        mL, hL = ozaki.tables(L)
        bits = sum(math.log2(x) for x in mL)
        assert 2 * hL - 2 <= bits - 2 - 0.1 and 2 * hL + 2 > bits - 2


def test_crt_reconstructs_exact_product(rng):
    """Residues of the exact product, reconstructed by CRT, give the exact
    product back (the scaling keeps |C'| < P/2)."""
    m, h = ozaki.tables(14)
    X = c3_like(rng, 300, 12)
    s = ozaki.sigma(np.einsum("ij,ij->j", X, X), h)
    x = ozaki.scaled_integers(X.T, s)
    cint = ozaki.exact_product(x, x)
    res = [np.vectorize(lambda v, ml=ml: v % ml, otypes=[object])(cint)
           for ml in m]
    assert np.array_equal(ozaki.crt(res, m), cint)


def test_scaling_bounds_the_integer_product(rng):
    m, h = ozaki.tables(14)
    P = 1
    for x in m:
        P *= x
    X = c3_like(rng, 2000, 20)
    s = ozaki.sigma(np.einsum("ij,ij->j", X, X), h)
    x = ozaki.scaled_integers(X.T, s)
    cint = ozaki.exact_product(x, x)
    assert max(abs(int(v)) for v in cint.ravel()) <= P // 4
    # and the integer rows use the headroom to within one eighth of an
    # octave: 2^(h-1) 8/9 < ||x'_o|| <= 2^(h-1) + sqrt(K)/2
    norms = np.sqrt(np.einsum("ij,ij->i", x, x))
    assert np.all(norms > 2.0 ** (h - 1) * 8 / 9)
    assert np.all(norms <= 2.0 ** (h - 1) + np.sqrt(2000) / 2)


def test_scale_code_is_the_largest_admissible_step():
    """scale_code against its definition in exact rational arithmetic: s
    = 2^(sigma-3) m (m in 8..15) is the largest value of that grid with
    norm2 s^2 <= 2^(2h-2) (1 - 2^-19), up to the one rounding of fl(g m^2)
    the code makes (it can only err towards the smaller step)."""
    from fractions import Fraction as F
    rng = np.random.default_rng(9)
    h = 55
    lim = F(2) ** (2 * h - 2) * (1 - F(1, 2 ** 19))
    vals = np.concatenate([10.0 ** rng.uniform(-250, 250, 3000),
                           np.ldexp(1.0, rng.integers(-500, 500, 200)),
                           np.nextafter(np.ldexp(1.0, rng.integers(-50, 50, 200)), 0)])
    codes = ozaki.scale_code(vals, h)
    for v, c in zip(vals, codes):
        e, m = int(c) >> 3, 8 + (int(c) & 7)
        s = F(m, 8) * F(2) ** e
        n2 = F(float(v))
        assert n2 * s * s <= lim * (1 + F(1, 2 ** 50))
        # the next grid value up is not admissible
        nxt = F(m + 1, 8) * F(2) ** e if m < 15 else F(2) ** (e + 1)
        assert n2 * nxt * nxt > lim * (1 - F(1, 2 ** 50))
    assert ozaki.scale_code([0.0], h)[0] == 0


def test_small_integer_inputs_are_exact(rng):
    """Inputs with few significant bits are represented exactly after
    scaling, so the split product equals the exact product."""
    A = rng.integers(-50, 50, (7, 40)).astype(np.float64)
    B = rng.integers(-50, 50, (9, 40)).astype(np.float64)
    C, _, _, _ = ozaki.product(A, B, 14)
    assert np.array_equal(C, A @ B.T)


@pytest.mark.parametrize("L,tol", [(14, 1e-13), (16, 1e-15)])
def test_error_against_double_double(orc, rng, L, tol):
    K, N = 4096, 48
    X = c3_like(rng, K, N)
    C, _, _, _ = ozaki.product(X.T.copy(), X.T.copy(), L)
    ref = np.zeros(N * N)
    orc.dsyrk(0, N, K, 1.0, np.ascontiguousarray(X.ravel()), N, 0.0, ref, N,
              dd=True)
    ref = ref.reshape(N, N)
    il = np.tril_indices(N)
    assert rel_fro(C[il], ref[il]) < tol
    # fewer moduli leave less headroom: the error grows as predicted
    C12, _, _, _ = ozaki.product(X.T.copy(), X.T.copy(), 12)
    assert rel_fro(C12[il], ref[il]) > rel_fro(C[il], ref[il])


def test_exact_to_double_matches_python_integers(rng):
    """The 128-bit C product and conversion equal the Python-integer
    restatement bit for bit (c3-like rows at the full scaling headroom,
    mixed signs)."""
    for L in (13, 14, 16):
        X = c3_like(rng, 700, 24)
        C, sa, sb, _ = ozaki.product(X.T.copy(), X.T.copy(), L)
        m, h = ozaki.tables(L)
        xa = ozaki.scaled_integers(X.T, sa)
        got = ozaki.exact_to_double(xa, xa, sa, sb)
        assert np.array_equal(got, C)
        rows = [0, 5, 23]
        got2, _, _ = ozaki.product_rows(X.T[rows].copy(), X.T.copy(), L)
        assert np.array_equal(got2, C[rows])
    A = rng.standard_normal((9, 333)) * 10.0 ** rng.uniform(0, 5, 9)[:, None]
    B = rng.standard_normal((11, 333))
    C, sa, sb, _ = ozaki.product(A, B, 14)
    got, _, _ = ozaki.product_rows(A, B, 14)
    assert np.array_equal(got, C)


# ---- the device's residue and CRT arithmetic, emulated in exact FP64 ------
# The GPU's INT8-product epilogue and CRT are FP64 sequences whose
# exactness rests on magnitude arguments; these tests run the same
# sequences with every FP64 operation rounded exactly as the device rounds
# it (fma through exact rationals) and compare with integer arithmetic.

def _fma(x, y, z):
    from fractions import Fraction
    return float(Fraction(x) * Fraction(y) + Fraction(z))


def _tables(L):
    """The CRT tables the library builds (make_tables): limbs of P and of
    W_l = (P / m_l) ((P / m_l)^-1 mod m_l), 37-bit limbs up to 14 moduli,
    32-bit above; wfrac_l = inv_l / m_l rounded once."""
    m, _ = ozaki.tables(L)
    P = 1
    for x in m:
        P *= x
    nl, lb = (3, 37) if L <= 14 else (4, 32)
    mask = (1 << lb) - 1
    limbs = lambda v: [float((v >> (lb * j)) & mask) for j in range(nl)]
    W, wf = [], []
    for ml in m:
        Pl = P // ml
        inv = pow(Pl % ml, -1, ml)
        W.append(limbs(Pl * inv))
        wf.append(inv / ml)
    return m, P, limbs(P), W, wf, nl, lb


def _device_crt(r, Pj, W, wf, nl, lb):
    """crt4 for one element: limb sums, quotient estimate, carries, then
    x = hi 2^64 + lo as the kernel assembles it."""
    a = [0.0] * nl
    f = 0.5
    for l, rl in enumerate(r):
        rl = float(rl)
        for j in range(nl):
            a[j] = _fma(rl, W[l][j], a[j])
        f = _fma(rl, wf[l], f)
    q = float(np.floor(f))
    two, inv = float(1 << lb), 1.0 / float(1 << lb)
    b = [_fma(-q, Pj[j], a[j]) for j in range(nl)]
    for j in range(nl - 1):
        c = float(np.floor(b[j] * inv))
        b[j] = _fma(-c, two, b[j])
        b[j + 1] = b[j + 1] + c
    if nl == 4:
        lo = (int(b[0]) & 0xFFFFFFFF) | ((int(b[1]) & 0xFFFFFFFF) << 32)
        hi = (int(b[3]) << 32) | (int(b[2]) & 0xFFFFFFFF)
    else:
        l0, l1 = int(b[0]), int(b[1])
        lo = (l0 | (l1 << 37)) & ((1 << 64) - 1)
        hi = int(b[2]) * 1024 + (l1 >> 27)
    return hi * (1 << 64) + lo


@pytest.mark.parametrize("L", [14, 16])
def test_device_crt_exact_on_chunk_sums(rng, L):
    """The CRT recovers x (|x| <= P / 4) from any representatives of its
    residue classes with |r_l| <= 128 x 32: symmetric residue bytes, the
    plain sums of up to 32 K chunks' residue bytes, and the extreme
    representatives of that range."""
    m, P, Pj, W, wf, nl, lb = _tables(L)
    lim = 128 * 32
    xs = [int(v) for v in rng.integers(-2 ** 62, 2 ** 62, 6)]
    xs += [P // 4, -(P // 4), 0, 1, -1, P // 4 - 12345]
    xs += [int(rng.integers(0, 2 ** 60)) * (P // 2 ** 62) // 4 for _ in range(4)]
    for x in xs:
        sym = [((x + ml // 2) % ml) - ml // 2 for ml in m]
        cases = [sym]
        # chunk sums: other members of each class inside [-lim, lim]
        cases.append([s + ml * int(rng.integers(-((lim - ml) // ml),
                                                (lim - ml) // ml + 1))
                      for s, ml in zip(sym, m)])
        # extremes: the largest-magnitude members, random signs
        ext = []
        for s, ml in zip(sym, m):
            up = s + ml * ((lim - s) // ml)
            dn = s - ml * ((lim + s) // ml)
            ext.append(up if rng.random() < 0.5 else dn)
        cases.append(ext)
        for r in cases:
            assert all(abs(v) <= lim for v in r)
            assert all((v - x) % ml == 0 for v, ml in zip(r, m))
            assert _device_crt(r, Pj, W, wf, nl, lb) == x


def test_device_residue_epilogue_symmetric(rng):
    """The INT8 products' epilogue: r = fma(-m, rint(a fl(1/m)), a) on an
    exact INT32 sum a is the symmetric residue of a for odd m (no
    correction step), and for m = 256 a member of [-128, 128] whose byte
    is a's class."""
    m, _ = ozaki.tables(16)
    edges = [0, 1, -1, 2 ** 31 - 1, -2 ** 31, 2 ** 30, -2 ** 30 + 7]
    for ml in m:
        inv = 1.0 / ml
        half = [k * ml + d for k in (-(2 ** 31) // ml + 1, -3, 0, 5,
                                     (2 ** 31 - 1) // ml - 1)
                for d in (-(ml // 2), ml // 2, (ml - 1) // 2, -((ml - 1) // 2))]
        vals = edges + half + [int(v) for v in rng.integers(-2 ** 31, 2 ** 31, 2000)]
        for a in vals:
            if not -2 ** 31 <= a < 2 ** 31:
                continue
            q = round(float(a) * inv)      # rint: ties to even
            r = _fma(-float(ml), float(q), float(a))
            assert r == a - ml * q
            r = int(r)
            if ml % 2:
                assert abs(r) <= (ml - 1) // 2 and (r - a) % ml == 0
            else:
                assert -128 <= r <= 128 and ((r & 255) - a) % 256 == 0
