"""CPU restatement of the split-precision (Ozaki scheme II) product used by
the library's skgemm_dgemm_ozaki / skgemm_dsyrk_ozaki.  TEST INFRASTRUCTURE
ONLY (see oracle.py); no counterpart in the reference.

The restatement follows the same steps with numpy / Python integers:
  h = floor(log2 P / 2 - 0.05); for ||x_o||^2 = g 2^e (g in [0.5, 1)) the
  scale s_o = 2^(sigma - 3) m, m in 8..15 the largest with
  fl(g m^2) <= 64 (1 - 2^-19) (e even, sigma = h - 1 - e/2) or
  <= 128 (1 - 2^-19) (e odd, sigma = h - (e + 3)/2), coded 8 sigma + m - 8
  x'_o = rint(fl(x_o s_o))            (exact FP64 integers)
  C'   = A' B'                        (exact, Python integers -- no CRT)
  C    = ldexp(fl(d / (m_a m_b)), -(sigma_a + sigma_b - 6)),
         d = fl(fl(C' >> 64) * 2^64 + fl(C' mod 2^64))
which is the value the GPU's residue / INT8-product / CRT pipeline must
reproduce bit for bit whenever |C'| < P / 2 (guaranteed by the scaling).

Pinning: tests/test_ozaki_oracle.py checks the exact product against an
independent residue + CRT reconstruction, the scaling bound |C'| <= P/4,
exactness on small-integer inputs, and the error against the double-double
oracle.
"""
import math

import numpy as np

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211,
          199, 197, 193, 191, 181, 179, 173]


def tables(L):
    m = MODULI[:L]
    bits = sum(math.log2(x) for x in m)
    h = int(math.floor(bits / 2.0 - 0.05))
    return m, h


def scale_code(norm2, h):
    """Scale codes 8 sigma + (m - 8) from squared norms (the GPU's
    scale_code: exact steps and one rounded product fl(g m^2))."""
    norm2 = np.asarray(norm2, np.float64)
    out = np.zeros(norm2.shape, np.int64)
    lim_even = 64.0 * (1.0 - 2.0 ** -19)
    lim_odd = 128.0 * (1.0 - 2.0 ** -19)
    for i, v in np.ndenumerate(norm2):
        if not (v > 0.0):
            continue
        g, e = math.frexp(float(v))
        odd = (e & 1) != 0
        sig = h - ((e + 3) >> 1) if odd else h - 1 - (e >> 1)
        lim = lim_odd if odd else lim_even
        if not odd and not (g * 64.0 <= lim):
            sig -= 1
            lim = 256.0 * (1.0 - 2.0 ** -19)
        m = 15
        while m > 8 and not (g * float(m * m) <= lim):
            m -= 1
        out[i] = 8 * sig + (m - 8)
    return out


def sigma(norm2, h):
    """Scale codes (kept under the old name for callers)."""
    return scale_code(norm2, h)


def code_exp(code):
    return np.asarray(code, np.int64) >> 3


def code_mant(code):
    return 8 + (np.asarray(code, np.int64) & 7)


def scale_of(code):
    """s = 2^(sigma - 3) m, exact."""
    return np.ldexp(code_mant(code).astype(np.float64), code_exp(code) - 3)


def scaled_integers(X, code):
    """X: O x K float64 (free index first).  Returns the exact
    integer-valued float64 array rint(fl(X_o * s_o))."""
    s = scale_of(code)
    return np.rint(X * s[:, None])


def residues(xint, m):
    """Symmetric residues of integer-valued float64 xint modulo m."""
    xi = xint.astype(np.int64)
    r = np.mod(xi, m)
    hi = (m - 1) // 2
    r = np.where(r > hi, r - m, r)
    return r.astype(np.int64)


def exact_product(xa, xb):
    """Exact C' = xa @ xb.T in Python integers (object arrays)."""
    A = np.array([[int(v) for v in row] for row in xa], dtype=object)
    B = np.array([[int(v) for v in row] for row in xb], dtype=object)
    return A.dot(B.T)


def crt(res_list, m):
    """Reconstruct integers from residues (symmetric range), independent of
    the GPU's Garner form: sum r_l * w_l mod P with w_l = (P/m_l) *
    ((P/m_l)^{-1} mod m_l)."""
    P = 1
    for x in m:
        P *= x
    out = np.zeros(res_list[0].shape, dtype=object)
    for r, ml in zip(res_list, m):
        Pl = P // ml
        w = Pl * pow(Pl, -1, ml)
        out = out + np.vectorize(int, otypes=[object])(r) * w
    out = np.vectorize(lambda v: v % P, otypes=[object])(out)
    return np.vectorize(lambda v: v - P if v > P // 2 else v,
                        otypes=[object])(out)


def to_double(cint, sa, sb):
    """The GPU's conversion of the exact integer product to FP64 and its
    unscaling by s_a s_b = 2^(sigma_a + sigma_b - 6) m_a m_b."""
    O1, O2 = cint.shape
    out = np.empty((O1, O2), np.float64)
    for i in range(O1):
        ea, ma = int(code_exp(sa[i])), int(code_mant(sa[i]))
        for j in range(O2):
            eb, mb = int(code_exp(sb[j])), int(code_mant(sb[j]))
            v = int(cint[i, j])
            hi = v >> 64
            lo = v & ((1 << 64) - 1)
            d = float(hi) * 18446744073709551616.0 + float(lo)
            out[i, j] = math.ldexp(d / float(ma * mb), -(ea + eb - 6))
    return out


def product(Xa, Xb, L=14):
    """Split-form product of Xa (M x K) and Xb^T (Xb: N x K, free index
    first), returning (C, sa, sb, exact integer product)."""
    m, h = tables(L)
    sa = sigma(np.einsum("ij,ij->i", Xa, Xa), h)
    sb = sigma(np.einsum("ij,ij->i", Xb, Xb), h)
    xa = scaled_integers(Xa, sa)
    xb = scaled_integers(Xb, sb)
    cint = exact_product(xa, xb)
    return to_double(cint, sa, sb), sa, sb, cint


def exact_to_double(xa, xb, sa, sb):
    """exact_product + to_double in 128-bit C arithmetic (oracle.cpp
    oracle_ozaki_exact), for sizes where Python integers are too slow.
    xa: M x K, xb: N x K integer-valued float64 (|x| < 2^55)."""
    from oracle import oracle
    a = np.ascontiguousarray(xa).astype(np.int64)
    b = np.ascontiguousarray(xb).astype(np.int64)
    assert np.array_equal(a, xa) and np.array_equal(b, xb)
    M, K = a.shape
    N = b.shape[0]
    sa = np.ascontiguousarray(sa, np.int32)
    sb = np.ascontiguousarray(sb, np.int32)
    out = np.empty(M * N, np.float64)
    oracle.load().oracle_ozaki_exact(M, N, K, a.ctypes.data, b.ctypes.data,
                                     sa.ctypes.data, sb.ctypes.data,
                                     out.ctypes.data)
    return out.reshape(M, N)


def product_rows(Xa, Xb, L=14, norm2_a=None, norm2_b=None):
    """The split-form product of the rows Xa (M x K) with every row of Xb
    (N x K), computed as product() does, through exact_to_double.  norm2_a /
    norm2_b: squared row norms when they are known for the full operand
    (the scale of a row depends on that row alone)."""
    m, h = tables(L)
    na = np.einsum("ij,ij->i", Xa, Xa) if norm2_a is None else norm2_a
    nb = np.einsum("ij,ij->i", Xb, Xb) if norm2_b is None else norm2_b
    sa, sb = sigma(na, h), sigma(nb, h)
    xa = scaled_integers(Xa, sa)
    xb = scaled_integers(Xb, sb)
    return exact_to_double(xa, xb, sa, sb), sa, sb
