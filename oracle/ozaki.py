"""CPU restatement of the split-precision (Ozaki scheme II) product used by
the library's skgemm_dgemm_ozaki / skgemm_dsyrk_ozaki.  TEST INFRASTRUCTURE
ONLY (see oracle.py); no counterpart in the reference.

The restatement follows the same steps with numpy / Python integers:
  sigma_o = h - 1 - ceil(e_o / 2),  ||x_o||^2 = f 2^e_o, h = floor(log2 P / 2 - 0.05)
  x'_o = rint(x_o 2^sigma_o)          (exact FP64 integers)
  C'   = A' B'                        (exact, Python integers -- no CRT)
  C    = (double(C' >> 64) * 2^64 + double(C' mod 2^64)) * 2^-(sa_i + sb_j)
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


def sigma(norm2, h):
    """Scale exponents from squared norms (same formula as the GPU)."""
    norm2 = np.asarray(norm2, np.float64)
    e = np.frexp(norm2)[1].astype(np.int64)
    s = h - 1 - ((e + 1) >> 1)
    s[~(norm2 > 0)] = 0
    return s


def scaled_integers(X, sig):
    """X: O x K float64 (free index first).  Returns exact integer-valued
    float64 array rint(X * 2^sigma_o)."""
    return np.rint(np.ldexp(X, np.asarray(sig)[:, None]))


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
    """The GPU's conversion of the exact integer product to FP64."""
    O1, O2 = cint.shape
    out = np.empty((O1, O2), np.float64)
    for i in range(O1):
        for j in range(O2):
            v = int(cint[i, j])
            hi = v >> 64
            lo = v & ((1 << 64) - 1)
            d = float(hi) * 18446744073709551616.0 + float(lo)
            out[i, j] = math.ldexp(d, -(int(sa[i]) + int(sb[j])))
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
