"""ctypes binding of libskgemm.so (the C ABI in include/skgemm.h).

Plumbing for tests and the benchmark: it loads the in-tree HIP library and
raises if it is missing.  There is no fallback path: every compute call goes
through the library's HIP kernels.

Device buffers are passed as integer addresses (e.g. torch tensor
``data_ptr()``), host buffers as numpy arrays.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libskgemm.so")

ROW_MAJOR, COL_MAJOR, PACK16, PACK32 = 0, 1, 2, 3
FP64, OZAKI_INT8 = 0, 1
SYRK_LOWER, SYRK_DIAG_BLOCKS = 0, 1
SLICES_AUTO = -1
OK = 0
ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_TIMEOUT, ERR_NOTSUP, ERR_STATE = -1, -2, -3, -4, -5, -6

# Every entry point include/skgemm.h declares (checked by the CPU tests).
EXPORTS = [
    "skgemm_last_error", "skgemm_version", "skgemm_create", "skgemm_destroy",
    "skgemm_set_stream", "skgemm_synchronize", "skgemm_set_spin_limit",
    "skgemm_set_waitfree", "skgemm_set_fixup_hold",
    "skgemm_set_syrk_store",
    "skgemm_device_info", "skgemm_dev_alloc", "skgemm_dev_free",
    "skgemm_memcpy_h2d", "skgemm_memcpy_d2h", "skgemm_dgemm",
    "skgemm_dgemm_host", "skgemm_dsyrk", "skgemm_dsyrk_host",
    "skgemm_dsyrk_noalloc",
    "skgemm_dgemm_ozaki", "skgemm_dsyrk_ozaki", "skgemm_dgemm_ozaki_host",
    "skgemm_dsyrk_ozaki_host", "skgemm_pack", "skgemm_transpose",
    "skgemm_nloc_create", "skgemm_nloc_apply", "skgemm_nloc_destroy",
    "skgemm_halo_extend", "skgemm_stencil_apply",
    "skgemm_batched", "skgemm_set_tuning", "skgemm_set_kchunks",
    "skgemm_num_tiles",
    "skgemm_tile_info", "skgemm_autotune", "skgemm_load_tuning",
    "skgemm_query_plan", "skgemm_timing_enable", "skgemm_timing_read",
    "skgemm_timing_stages", "skgemm_split_info",
    "skgemm_last_kernel_ms", "skgemm_peak_probe",
    "skgemm_copy_probe", "skgemm_fill_uniform", "skgemm_fill_normal",
    "skgemm_scale_columns_pow10", "skgemm_server_create",
    "skgemm_server_destroy", "skgemm_server_add_system",
    "skgemm_server_submit", "skgemm_server_wait", "skgemm_server_stats",
]


class SkgemmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("skgemm error %d: %s" % (code, msg))
        self.code = code


_lib = None

c_int, c_double, c_void_p, c_int64, c_size_t, c_uint64 = (
    ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int64,
    ctypes.c_size_t, ctypes.c_uint64)


class Problem(ctypes.Structure):
    """skgemm_problem."""
    _fields_ = [("kind", c_int), ("layout_a", c_int), ("layout_b", c_int),
                ("layout_d", c_int), ("M", c_int), ("N", c_int), ("K", c_int),
                ("alpha", c_double), ("beta", c_double), ("gamma", c_double),
                ("A", c_void_p), ("lda", c_int64), ("B", c_void_p),
                ("ldb", c_int64), ("D", c_void_p), ("ldd", c_int64)]


def lib():
    """Load libskgemm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libskgemm.so not built: run "
                           "python stream-k-gemm_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    P = c_void_p
    sig = {
        "skgemm_last_error": (ctypes.c_char_p, []),
        "skgemm_version": (ctypes.c_char_p, []),
        "skgemm_create": (c_int, [c_int, ctypes.POINTER(P)]),
        "skgemm_destroy": (c_int, [P]),
        "skgemm_set_stream": (c_int, [P, P]),
        "skgemm_synchronize": (c_int, [P]),
        "skgemm_set_spin_limit": (c_int, [P, c_int64]),
        "skgemm_set_waitfree": (c_int, [P, c_int]),
        "skgemm_set_fixup_hold": (c_int, [P, c_int]),
        "skgemm_set_syrk_store": (c_int, [P, c_int]),
        "skgemm_device_info": (c_int, [P, P, P, P, P]),
        "skgemm_dev_alloc": (c_int, [P, c_size_t, c_int, ctypes.POINTER(P)]),
        "skgemm_dev_free": (c_int, [P, P]),
        "skgemm_memcpy_h2d": (c_int, [P, P, P, c_size_t]),
        "skgemm_memcpy_d2h": (c_int, [P, P, P, c_size_t]),
        "skgemm_dgemm": (c_int, [P, c_int, c_int, c_int, c_int, c_int, c_int,
                                 c_double, P, c_int64, P, c_int64, c_double,
                                 P, c_int64, c_double]),
        "skgemm_dgemm_host": (c_int, [P, c_int, c_int, c_int, c_int, c_int,
                                      c_int, c_double, P, c_int64, P, c_int64,
                                      c_double, P, c_int64, c_double]),
        "skgemm_dsyrk": (c_int, [P, c_int, c_int, c_int, c_double, P, c_int64,
                                 c_double, P, c_int64]),
        "skgemm_dsyrk_host": (c_int, [P, c_int, c_int, c_int, c_double, P,
                                      c_int64, c_double, P, c_int64]),
        "skgemm_dsyrk_noalloc": (c_int, [P, c_int, c_int, c_int, c_double, P,
                                         c_int64, c_double, P, c_int64]),
        "skgemm_dgemm_ozaki": (c_int, [P, c_int, c_int, c_int, c_int, c_int,
                                       c_int, c_double, P, c_int64, P, c_int64,
                                       c_double, P, c_int64, c_double, c_int,
                                       P]),
        "skgemm_dsyrk_ozaki": (c_int, [P, c_int, c_int, c_int, c_double, P,
                                       c_int64, c_double, P, c_int64, c_int,
                                       P]),
        "skgemm_dgemm_ozaki_host": (c_int, [P, c_int, c_int, c_int, c_int,
                                            c_int, c_int, c_double, P, c_int64,
                                            P, c_int64, c_double, P, c_int64,
                                            c_double, c_int, P]),
        "skgemm_dsyrk_ozaki_host": (c_int, [P, c_int, c_int, c_int, c_double,
                                            P, c_int64, c_double, P, c_int64,
                                            c_int, P]),
        "skgemm_pack": (c_int, [P, c_int, P, c_int64, c_int, P, c_int64, c_int,
                                c_int]),
        "skgemm_transpose": (c_int, [P, c_int, P, c_int64, c_int, P, c_int64,
                                     c_int, c_int]),
        "skgemm_nloc_create": (c_int, [P, c_int, c_int, P, P, P, P, P, P,
                                       ctypes.POINTER(c_void_p)]),
        "skgemm_nloc_apply": (c_int, [P, c_int, c_double, P, P, c_int]),
        "skgemm_nloc_destroy": (c_int, [P]),
        "skgemm_halo_extend": (c_int, [P, c_int, c_int, c_int, c_int, c_int,
                                       P, P, c_int]),
        "skgemm_stencil_apply": (c_int, [P, c_int, c_int, c_int, c_int,
                                         c_int, P, P, P, P, P, c_double, P,
                                         c_int]),
        "skgemm_batched": (c_int, [P, P, c_int, c_int, c_int, c_int]),
        "skgemm_set_tuning": (c_int, [P, c_int, c_int]),
        "skgemm_set_kchunks": (c_int, [P, c_int]),
        "skgemm_num_tiles": (c_int, []),
        "skgemm_tile_info": (c_int, [c_int, P, P, P, P]),
        "skgemm_autotune": (c_int, [P, c_int, c_int, c_int, c_int, c_int,
                                    c_int, ctypes.c_char_p, P, P, P, P]),
        "skgemm_load_tuning": (c_int, [P, ctypes.c_char_p]),
        "skgemm_query_plan": (c_int, [P, c_int, c_int, c_int, c_int, c_int,
                                      c_int, P, P, P, P, P]),
        "skgemm_timing_enable": (c_int, [P, c_int]),
        "skgemm_timing_read": (c_int, [P, P, c_int, P]),
        "skgemm_timing_stages": (c_int, [P, c_int]),
        "skgemm_split_info": (c_int, [P, P]),
        "skgemm_last_kernel_ms": (c_int, [P, P]),
        "skgemm_peak_probe": (c_int, [P, c_int, P]),
        "skgemm_copy_probe": (c_int, [P, c_size_t, P]),
        "skgemm_fill_uniform": (c_int, [P, P, c_int64, c_uint64, c_double,
                                        c_double]),
        "skgemm_fill_normal": (c_int, [P, P, c_int64, c_uint64, c_double]),
        "skgemm_scale_columns_pow10": (c_int, [P, P, c_int, c_int, c_int64,
                                               c_uint64, c_double]),
        "skgemm_server_create": (c_int, [c_int, c_size_t, c_int, c_int, c_int,
                                         c_int, ctypes.POINTER(P)]),
        "skgemm_server_destroy": (c_int, [P]),
        "skgemm_server_add_system": (c_int, [P, P, c_int, c_int, P]),
        "skgemm_server_submit": (c_int, [P, c_int, c_int, P, P, P]),
        "skgemm_server_wait": (c_int, [P, c_int64, P]),
        "skgemm_server_stats": (c_int, [P, P, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise SkgemmError(rc, lib().skgemm_last_error().decode())
    return rc


def _ptr(x):
    """Address of a numpy array, a torch tensor, an int or None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    raise TypeError("cannot take the address of %r" % type(x))


def _out_double():
    return ctypes.c_double(0.0)


class Context:
    """One skgemm_ctx (HIP stream + Stream-K workspace + tuning table)."""

    def __init__(self, device=0):
        h = c_void_p()
        check(lib().skgemm_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            check(lib().skgemm_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- control ---------------------------------------------------------
    def set_stream(self, stream_ptr):
        check(lib().skgemm_set_stream(self.h, stream_ptr))

    def synchronize(self):
        check(lib().skgemm_synchronize(self.h))

    def set_spin_limit(self, n):
        check(lib().skgemm_set_spin_limit(self.h, n))

    def set_waitfree(self, on):
        """Last-arriver Stream-K fix-up (no waits between workgroups)."""
        check(lib().skgemm_set_waitfree(self.h, 1 if on else 0))

    def set_fixup_hold(self, on):
        """Test hook: peers never raise their fix-up flags, so waiting owners
        time out (spin limit clamped to 64 rounds)."""
        check(lib().skgemm_set_fixup_hold(self.h, 1 if on else 0))

    def set_syrk_store(self, mode):
        """SYRK_LOWER or SYRK_DIAG_BLOCKS (whole 16 x 16 diagonal blocks)."""
        check(lib().skgemm_set_syrk_store(self.h, mode))

    def device_info(self):
        v = [c_int() for _ in range(4)]
        check(lib().skgemm_device_info(self.h, *[ctypes.byref(x) for x in v]))
        return dict(num_cus=v[0].value, num_xcds=v[1].value,
                    lds_per_cu=v[2].value, clock_mhz=v[3].value)

    def set_tuning(self, tile=-1, grid=0):
        check(lib().skgemm_set_tuning(self.h, tile, grid))

    def set_kchunks(self, n=0):
        check(lib().skgemm_set_kchunks(self.h, n))

    def query_plan(self, kind, M, N, K, la=COL_MAJOR, lb=ROW_MAJOR):
        t, g, kc = c_int(), c_int(), c_int()
        nt, ni = c_int64(), c_int64()
        check(lib().skgemm_query_plan(self.h, kind, la, lb, M, N, K,
                                      ctypes.byref(t), ctypes.byref(g),
                                      ctypes.byref(kc), ctypes.byref(nt),
                                      ctypes.byref(ni)))
        return dict(tile=t.value, grid=g.value, kchunks=kc.value,
                    tiles=nt.value, iters=ni.value)

    def timing_enable(self, capacity):
        check(lib().skgemm_timing_enable(self.h, capacity))

    def timing_read(self, max_n=100000):
        buf = np.zeros(max_n, np.float64)
        n = c_int()
        check(lib().skgemm_timing_read(self.h, buf.ctypes.data, max_n,
                                       ctypes.byref(n)))
        return buf[:n.value].copy()

    def timing_stages(self, on=True):
        check(lib().skgemm_timing_stages(self.h, 1 if on else 0))

    def split_moduli(self):
        """Modulus count of the last split-form call (0: plain FP64)."""
        v = c_int()
        check(lib().skgemm_split_info(self.h, ctypes.byref(v)))
        return v.value

    def last_kernel_ms(self):
        d = _out_double()
        check(lib().skgemm_last_kernel_ms(self.h, ctypes.byref(d)))
        return d.value

    # -- compute (device pointers) ------------------------------------------
    def dgemm(self, la, lb, ld_, M, N, K, alpha, A, lda, B, ldb, beta, D, ldd,
              gamma=0.0):
        check(lib().skgemm_dgemm(self.h, la, lb, ld_, M, N, K, alpha, _ptr(A),
                                 lda, _ptr(B), ldb, beta, _ptr(D), ldd, gamma))

    def dsyrk(self, la, N, K, alpha, A, lda, beta, C, ldc):
        check(lib().skgemm_dsyrk(self.h, la, N, K, alpha, _ptr(A), lda, beta,
                                 _ptr(C), ldc))

    def dsyrk_noalloc(self, la, N, K, alpha, A, lda, beta, C, ldc):
        check(lib().skgemm_dsyrk_noalloc(self.h, la, N, K, alpha, _ptr(A),
                                         lda, beta, _ptr(C), ldc))

    def dgemm_ozaki(self, la, lb, ld_, M, N, K, alpha, A, lda, B, ldb, beta, D,
                    ldd, gamma=0.0, slices=0, bound=True):
        e = _out_double()
        check(lib().skgemm_dgemm_ozaki(self.h, la, lb, ld_, M, N, K, alpha,
                                       _ptr(A), lda, _ptr(B), ldb, beta,
                                       _ptr(D), ldd, gamma, slices,
                                       ctypes.byref(e) if bound else None))
        return e.value if bound else None

    def dsyrk_ozaki(self, la, N, K, alpha, A, lda, beta, C, ldc, slices=0,
                    bound=True):
        """Returns the error bound (bound=True: synchronous), else None."""
        e = _out_double()
        check(lib().skgemm_dsyrk_ozaki(self.h, la, N, K, alpha, _ptr(A), lda,
                                       beta, _ptr(C), ldc, slices,
                                       ctypes.byref(e) if bound else None))
        return e.value if bound else None

    def dgemm_ozaki_host(self, la, lb, ld_, M, N, K, alpha, A, lda, B, ldb,
                         beta, D, ldd, gamma=0.0, slices=0):
        for x in (A, B, D):
            assert x.dtype == np.float64 and x.flags.c_contiguous
        e = _out_double()
        check(lib().skgemm_dgemm_ozaki_host(self.h, la, lb, ld_, M, N, K,
                                            alpha, _ptr(A), lda, _ptr(B), ldb,
                                            beta, _ptr(D), ldd, gamma, slices,
                                            ctypes.byref(e)))
        return e.value

    def dsyrk_ozaki_host(self, la, N, K, alpha, A, lda, beta, C, ldc,
                         slices=0):
        for x in (A, C):
            assert x.dtype == np.float64 and x.flags.c_contiguous
        e = _out_double()
        check(lib().skgemm_dsyrk_ozaki_host(self.h, la, N, K, alpha, _ptr(A),
                                            lda, beta, _ptr(C), ldc, slices,
                                            ctypes.byref(e)))
        return e.value

    def pack(self, src_layout, src, ld_src, dst_layout, dst, ld_dst, o_ext,
             r_ext):
        check(lib().skgemm_pack(self.h, src_layout, _ptr(src), ld_src,
                                dst_layout, _ptr(dst), ld_dst, o_ext, r_ext))

    def halo_extend(self, ni, nj, nk, nb, F, local16, ext16, periodic=True,
                    boundary=None):
        """boundary: 0 zero, 1 periodic, 2 interior only (default from
        `periodic`)."""
        if boundary is None:
            boundary = 1 if periodic else 0
        check(lib().skgemm_halo_extend(self.h, ni, nj, nk, nb, F,
                                       _ptr(local16), _ptr(ext16), boundary))

    def stencil_apply(self, ni, nj, nk, nb, F, ext16, A_diag, cx, cy, cz, c0,
                      out, out_tile=32):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        cx, cy, cz = f(cx), f(cy), f(cz)
        check(lib().skgemm_stencil_apply(self.h, ni, nj, nk, nb, F,
                                         _ptr(ext16), _ptr(A_diag),
                                         cx.ctypes.data, cy.ctypes.data,
                                         cz.ctypes.data, c0, _ptr(out),
                                         out_tile))

    def transpose(self, src_layout, src, ld_src, dst_layout, dst, ld_dst,
                  o_ext, r_ext):
        check(lib().skgemm_transpose(self.h, src_layout, _ptr(src), ld_src,
                                     dst_layout, _ptr(dst), ld_dst, o_ext,
                                     r_ext))

    def batched(self, problems, precision=FP64, slices=0, use_graph=0):
        arr = (Problem * len(problems))(*problems)
        check(lib().skgemm_batched(self.h, ctypes.cast(arr, c_void_p),
                                   len(problems), precision, slices,
                                   use_graph))

    # -- compute (host numpy arrays) ----------------------------------------
    def dgemm_host(self, la, lb, ld_, M, N, K, alpha, A, lda, B, ldb, beta, D,
                   ldd, gamma=0.0):
        for x in (A, B, D):
            assert x.dtype == np.float64 and x.flags.c_contiguous
        check(lib().skgemm_dgemm_host(self.h, la, lb, ld_, M, N, K, alpha,
                                      _ptr(A), lda, _ptr(B), ldb, beta,
                                      _ptr(D), ldd, gamma))

    def dsyrk_host(self, la, N, K, alpha, A, lda, beta, C, ldc):
        for x in (A, C):
            assert x.dtype == np.float64 and x.flags.c_contiguous
        check(lib().skgemm_dsyrk_host(self.h, la, N, K, alpha, _ptr(A), lda,
                                      beta, _ptr(C), ldc))

    # -- fills / probes ----------------------------------------------------
    def fill_uniform(self, x, n, seed, lo=-1.0, hi=1.0):
        check(lib().skgemm_fill_uniform(self.h, _ptr(x), n, seed, lo, hi))

    def fill_normal(self, x, n, seed, sigma=1.0):
        check(lib().skgemm_fill_normal(self.h, _ptr(x), n, seed, sigma))

    def scale_columns_pow10(self, x, rows, cols, ld, seed, umax):
        check(lib().skgemm_scale_columns_pow10(self.h, _ptr(x), rows, cols, ld,
                                               seed, umax))

    def peak_probe(self, which):
        d = _out_double()
        check(lib().skgemm_peak_probe(self.h, which, ctypes.byref(d)))
        return d.value

    def copy_probe(self, nbytes):
        d = _out_double()
        check(lib().skgemm_copy_probe(self.h, nbytes, ctypes.byref(d)))
        return d.value

    def autotune(self, kind, M, N, K, la=COL_MAJOR, lb=ROW_MAJOR,
                 cache_path=None):
        t, g, kc, us = c_int(), c_int(), c_int(), _out_double()
        check(lib().skgemm_autotune(
            self.h, kind, la, lb, M, N, K,
            cache_path.encode() if cache_path else None, ctypes.byref(t),
            ctypes.byref(g), ctypes.byref(kc), ctypes.byref(us)))
        return dict(tile=t.value, grid=g.value, kchunks=kc.value, us=us.value)

    def load_tuning(self, path):
        check(lib().skgemm_load_tuning(self.h, path.encode()))


class DeviceArray:
    """A float64 device buffer allocated through skgemm_dev_alloc."""

    def __init__(self, ctx, n, zero=False):
        self.ctx, self.n = ctx, int(n)
        h = c_void_p()
        check(lib().skgemm_dev_alloc(ctx.h, max(self.n, 1) * 8,
                                     1 if zero else 0, ctypes.byref(h)))
        self.ptr = h.value

    @classmethod
    def from_host(cls, ctx, a):
        a = np.ascontiguousarray(a, np.float64).ravel()
        d = cls(ctx, a.size)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64).ravel()
        assert a.size <= self.n
        check(lib().skgemm_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data,
                                      a.size * 8))

    def download(self, n=None):
        n = self.n if n is None else n
        out = np.empty(n, np.float64)
        check(lib().skgemm_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr,
                                      n * 8))
        return out

    def free(self):
        if self.ptr:
            check(lib().skgemm_dev_free(self.ctx.h, self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def num_tiles():
    return lib().skgemm_num_tiles()


def tile_info(i):
    v = [c_int() for _ in range(4)]
    check(lib().skgemm_tile_info(i, *[ctypes.byref(x) for x in v]))
    return dict(bm=v[0].value, bn=v[1].value, bk=v[2].value,
                waves=v[3].value)


class Server:
    """Continuous-batching request loop (skgemm_server_*) over systems whose
    Psi stays resident in device memory."""

    def __init__(self, device=0, hbm_budget=64 << 30, max_batch=16,
                 max_inflight=2, precision=FP64, slices=0):
        h = c_void_p()
        check(lib().skgemm_server_create(device, hbm_budget, max_batch,
                                         max_inflight, precision, slices,
                                         ctypes.byref(h)))
        self.h = h

    def add_system(self, psi_ptr, Nd, Ns):
        i = ctypes.c_int()
        check(lib().skgemm_server_add_system(self.h, psi_ptr, Nd, Ns,
                                             ctypes.byref(i)))
        return i.value

    def submit(self, system, kind, out_ptr, q_ptr=None):
        t = ctypes.c_int64()
        check(lib().skgemm_server_submit(self.h, system, kind, q_ptr, out_ptr,
                                         ctypes.byref(t)))
        return t.value

    def wait(self, ticket):
        """Latency of the request in microseconds (submit -> completion)."""
        lat = _out_double()
        check(lib().skgemm_server_wait(self.h, ticket, ctypes.byref(lat)))
        return lat.value

    def stats(self):
        c, l, pk = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        f, b = _out_double(), _out_double()
        check(lib().skgemm_server_stats(self.h, ctypes.byref(c),
                                        ctypes.byref(l), ctypes.byref(f),
                                        ctypes.byref(b), ctypes.byref(pk)))
        return {"completed": c.value, "launches": l.value, "flops": f.value,
                "busy_s": b.value, "peak_inflight": pk.value}

    def close(self):
        if self.h:
            check(lib().skgemm_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Nloc:
    """Non-local projector set (skgemm_nloc_*): projectors concatenated in
    order (nrow[p] grid indices, chi_p column-major nrow x ncol, gamma_p)."""

    def __init__(self, ctx, Nd, nrow, ncol, atom_index, index_all, chi_all,
                 gamma_all):
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self._keep = (i32(nrow), i32(ncol), i32(atom_index), i32(index_all),
                      f64(chi_all), f64(gamma_all))
        nr, nc, at, ix, ch, ga = self._keep
        h = c_void_p()
        check(lib().skgemm_nloc_create(ctx.h, Nd, len(nr), nr.ctypes.data,
                                       nc.ctypes.data, at.ctypes.data,
                                       ix.ctypes.data, ch.ctypes.data,
                                       ga.ctypes.data, ctypes.byref(h)))
        self.h = h

    def apply(self, nb, dv, psi16, result, result_tile=32):
        check(lib().skgemm_nloc_apply(self.h, nb, dv, _ptr(psi16),
                                      _ptr(result), result_tile))

    def close(self):
        if self.h:
            check(lib().skgemm_nloc_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
