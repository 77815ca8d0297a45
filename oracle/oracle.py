"""ctypes binding of the oracle (oracle.cpp): TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker (or as the timed CPU port of the
reference path for the CPU baseline).  The library under test never uses it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build(native=False, out=None):
    """Compile oracle.cpp.  native=True builds with -march=native into
    `out` (used for the CPU baseline on the machine that runs it)."""
    if out is None:
        out = LIB_PATH
    march = "-march=native" if native else "-march=x86-64-v3"
    cmd = ["g++", "-O3", march, "-fopenmp", "-fPIC", "-std=c++17",
           "-ffp-contract=off", "-shared", "-o", out,
           os.path.join(HERE, "oracle.cpp")]
    subprocess.run(cmd, check=True)
    return out


def load(path=None):
    global _lib
    if path is None and _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        build(out=p)
    L = ctypes.CDLL(p)
    D = ctypes.c_double
    I = ctypes.c_int
    P = ctypes.c_void_p
    L64 = ctypes.c_long
    sig = {
        "oracle_fill_uniform": [P, L64, ctypes.c_uint64, D, D],
        "oracle_gemm_test_inputs": [I, I, I, P, P],
        "oracle_init_A_packed": [I, I, P, I],
        "oracle_trp_test_inputs": [I, I, P, P],
        "oracle_pack_A": [I, I, P, P],
        "oracle_pack_B": [I, I, P, P],
        "oracle_pack_c16": [I, I, P, P],
        "oracle_unpack_c16": [I, I, P, P],
        "oracle_band_to_grid": [I, I, P, P],
        "oracle_nloc_apply": [I, I, D, I, P, P, P, P, P, P, P, P, I],
        "oracle_halo_extend": [I, I, I, I, I, P, P, I],
        "oracle_stencil_apply": [I, I, I, I, I, P, P, P, P, P, D, P, I],
        "oracle_grid_to_band": [I, I, P, P],
        "oracle_gemm_crsp_packed": [I, I, I, P, P, P, D, D, I],
        "oracle_gemm_opt_packed": [I, I, I, P, P, P, D, D, I],
        "oracle_gemm_naive": [I, I, I, P, P, P, D, D],
        "oracle_syrk_reference": [I, I, P, P, D, D],
        "oracle_syrk_tiles": [I, I, P, P, D, D, I],
        "oracle_trp_reference": [I, I, P, P, P],
        "oracle_trp_3d": [I, I, P, P, P, I],
        "oracle_dgemm": [I, I, I, I, I, I, D, P, L64, P, L64, D, P, L64, D, I],
        "oracle_dsyrk": [I, I, I, D, P, L64, D, P, L64, I],
        "oracle_ozaki_exact": [I, I, I, P, P, P, P, P],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.restype = None
        f.argtypes = args
    if path is None:
        _lib = L
    return L


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data


def fill_uniform(n, seed, lo=-1.0, hi=1.0):
    x = np.empty(n, np.float64)
    load().oracle_fill_uniform(_p(x), n, seed, lo, hi)
    return x


def gemm_test_inputs(M, N, K):
    """gemm_crsp.cpp:451-460 inputs: A column-major (flat), B row-major."""
    A = np.empty(M * K, np.float64)
    B = np.empty(K * N, np.float64)
    load().oracle_gemm_test_inputs(M, N, K, _p(A), _p(B))
    return A, B


def init_A_packed(K, M, seed):
    """syrk_crsp.cpp:379-395 (glibc rand) -> P16 [ceil(M/16)][K][16]."""
    t = (M + 15) // 16
    A = np.empty(t * K * 16, np.float64)
    load().oracle_init_A_packed(K, M, _p(A), seed)
    return A


def trp_test_inputs(Nd, nb):
    A = np.empty((Nd // 16) * nb * 16, np.float64)
    B = np.empty(nb * nb, np.float64)
    load().oracle_trp_test_inputs(Nd, nb, _p(A), _p(B))
    return A, B


def pack_A(M, K, A_colmajor):
    out = np.zeros((M // 16) * K * 16, np.float64)
    load().oracle_pack_A(M, K, _p(A_colmajor), _p(out))
    return out


def pack_B(N, K, B_rowmajor):
    out = np.zeros((N // 32) * K * 32, np.float64)
    load().oracle_pack_B(N, K, _p(B_rowmajor), _p(out))
    return out


def pack_c16(Nd, nb, src):
    out = np.zeros((Nd // 16) * nb * 16, np.float64)
    load().oracle_pack_c16(Nd, nb, _p(src), _p(out))
    return out


def unpack_c16(Nd, nb, src):
    out = np.zeros(Nd * nb, np.float64)
    load().oracle_unpack_c16(Nd, nb, _p(src), _p(out))
    return out


def nloc_apply(K, nb, dv, nrow, ncol, atom_index, index_all, chi_all,
               gamma_all, psi16, result, out_tile):
    """In place on result (out_tile = 32 or 16 band packing)."""
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    nrow, ncol, atom_index, index_all = (i32(nrow), i32(ncol), i32(atom_index),
                                         i32(index_all))
    chi_all, gamma_all, psi16 = f64(chi_all), f64(gamma_all), f64(psi16)
    pi = lambda a: a.ctypes.data
    load().oracle_nloc_apply(K, nb, dv, len(nrow), pi(nrow), pi(ncol),
                             pi(atom_index), pi(index_all), _p(chi_all),
                             _p(gamma_all), _p(psi16), _p(result), out_tile)
    return result


def halo_extend(ni, nj, nk, nb, F, local16, periodic):
    t = (nb + 15) // 16
    out = np.zeros(t * (ni + 2 * F) * (nj + 2 * F) * (nk + 2 * F) * 16)
    load().oracle_halo_extend(ni, nj, nk, nb, F, _p(local16), _p(out),
                              1 if periodic else 0)
    return out


def stencil_apply(ni, nj, nk, nb, F, ext16, A_diag, cx, cy, cz, c0, out_tile,
                  out=None):
    t = (nb + out_tile - 1) // out_tile
    if out is None:
        out = np.zeros(t * ni * nj * nk * out_tile)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    load().oracle_stencil_apply(ni, nj, nk, nb, F, _p(f64(ext16)),
                                _p(f64(A_diag)), _p(f64(cx)), _p(f64(cy)),
                                _p(f64(cz)), c0, _p(out), out_tile)
    return out


def band_to_grid(Nd, nb, src):
    out = np.zeros((Nd // 16) * nb * 16, np.float64)
    load().oracle_band_to_grid(Nd, nb, _p(src), _p(out))
    return out


def grid_to_band(Nd, nb, src):
    out = np.zeros((nb // 16) * Nd * 16, np.float64)
    load().oracle_grid_to_band(Nd, nb, _p(src), _p(out))
    return out


def gemm_crsp(M, N, K, A_packed, B_packed, alpha, gamma, nthreads=0, D=None):
    if D is None:
        D = np.zeros(M * N, np.float64)
    load().oracle_gemm_crsp_packed(M, N, K, _p(A_packed), _p(B_packed), _p(D),
                                   alpha, gamma, nthreads)
    return D


def gemm_opt(M, N, K, A_packed, B_packed, alpha, gamma, nthreads=0, D=None):
    if D is None:
        D = np.zeros(M * N, np.float64)
    load().oracle_gemm_opt_packed(M, N, K, _p(A_packed), _p(B_packed), _p(D),
                                  alpha, gamma, nthreads)
    return D


def gemm_naive(M, N, K, A_colmajor, B_rowmajor, alpha, gamma):
    D = np.zeros(M * N, np.float64)
    load().oracle_gemm_naive(M, N, K, _p(A_colmajor), _p(B_rowmajor), _p(D),
                             alpha, gamma)
    return D


def syrk_reference(K, M, A_packed, alpha, beta, C=None):
    if C is None:
        C = np.zeros(M * M, np.float64)
    load().oracle_syrk_reference(K, M, _p(A_packed), _p(C), alpha, beta)
    return C


def syrk_tiles(K, M, A_packed, alpha, beta, C=None, nthreads=0):
    if C is None:
        C = np.zeros(M * M, np.float64)
    load().oracle_syrk_tiles(K, M, _p(A_packed), _p(C), alpha, beta, nthreads)
    return C


def trp_reference(Nd, nb, A_grid, B):
    D = np.zeros(((nb + 15) // 16) * Nd * 16, np.float64)
    load().oracle_trp_reference(Nd, nb, _p(A_grid), _p(B), _p(D))
    return D


def trp_3d(Nd, nb, A_grid, B, nthreads=0):
    D = np.zeros(((nb + 15) // 16) * Nd * 16, np.float64)
    load().oracle_trp_3d(Nd, nb, _p(A_grid), _p(B), _p(D), nthreads)
    return D


def dgemm(la, lb, ldl, M, N, K, alpha, A, lda, B, ldb, beta, D, ldd, gamma=0.0,
          dd=True):
    """General-layout checker (in place on D); dd = double-double sums."""
    load().oracle_dgemm(la, lb, ldl, M, N, K, alpha, _p(A), lda, _p(B), ldb,
                        beta, _p(D), ldd, gamma, 1 if dd else 0)
    return D


def dsyrk(la, N, K, alpha, A, lda, beta, C, ldc, dd=True):
    """Lower-triangle SYRK checker (in place on C, row-major N x N)."""
    load().oracle_dsyrk(la, N, K, alpha, _p(A), lda, beta, _p(C), ldc,
                        1 if dd else 0)
    return C


def rel_fro(x, ref):
    """||x - ref||_F / ||ref||_F."""
    x = np.asarray(x, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (den if den > 0 else 1.0))
