"""Write the small golden fixtures under tests/golden/.

gemm_kat.npz   - the reference's GEMM known-answer cases (gemm_crsp.cpp:555-565,
                 inputs gemm_crsp.cpp:451-460: A[i,k] = 0.01(i+1) + 0.001(k+1)
                 column-major, B[k,j] = 0.001(k+1) + 0.01(j+1) row-major).
                 Expected D from the closed form
                   D_ij = alpha*(1e-5 (I+J) S1 + 1e-4 I J K + 1e-6 S2) + gamma,
                   I = i+1, J = j+1, S1 = K(K+1)/2, S2 = K(K+1)(2K+1)/6,
                 evaluated in exact rational arithmetic and rounded once.
syrk_glibc.npz - the first SYRK case of syrk_crsp.cpp:474 (K=64, M=32,
                 seed 42): glibc srand/rand()%100 integers in the P16 layout
                 (syrk_crsp.cpp:379-395) and the exact A^T A in integer
                 arithmetic (entries are integers / 100).

Run: python tests/golden/gen_golden.py  (needs the oracle library for
glibc's rand stream).
"""
import os
import sys
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

GEMM_CASES = [(32, 64, 128, 1.0, 0.0), (32, 64, 4096, 1.0, 0.0),
              (32, 64, 4096, 2.0, 0.5), (64, 128, 2048, 1.0, 0.0)]


def gemm_closed_form(M, N, K, alpha, gamma):
    S1 = Fraction(K * (K + 1), 2)
    S2 = Fraction(K * (K + 1) * (2 * K + 1), 6)
    a, g = Fraction(alpha), Fraction(gamma)
    out = np.empty((M, N), np.float64)
    for i in range(M):
        I = i + 1
        for j in range(N):
            J = j + 1
            v = (Fraction(1, 10**5) * (I + J) * S1 + Fraction(1, 10**4) * I * J * K
                 + Fraction(1, 10**6) * S2)
            out[i, j] = float(a * v + g)
    return out


def main():
    arrays = {}
    for t, (M, N, K, alpha, gamma) in enumerate(GEMM_CASES):
        arrays["case%d_shape" % t] = np.array([M, N, K], np.int64)
        arrays["case%d_coef" % t] = np.array([alpha, gamma], np.float64)
        arrays["case%d_D" % t] = gemm_closed_form(M, N, K, alpha, gamma)
    np.savez_compressed(os.path.join(HERE, "gemm_kat.npz"), **arrays)

    from oracle import oracle
    K, M, seed = 64, 32, 42
    ap = oracle.init_A_packed(K, M, seed)              # values r/100
    ints = np.rint(ap * 100).astype(np.int64)          # r = rand() % 100
    A = ints.reshape(M // 16, K, 16).transpose(1, 0, 2).reshape(K, M)
    exact = A.T @ A                                     # integer, exact
    np.savez_compressed(os.path.join(HERE, "syrk_glibc.npz"),
                        shape=np.array([K, M, seed], np.int64),
                        a_packed_int=ints.astype(np.uint8),
                        ata_int=exact)
    print("wrote gemm_kat.npz, syrk_glibc.npz")


if __name__ == "__main__":
    main()
