"""MI355X-native FP64 Stream-K GEMM / SYRK (HIP, gfx950).

The product is libskgemm.so and its C ABI (include/skgemm.h); skgemm.py is
a ctypes binding used by the tests and the benchmark.
"""
from .skgemm import *  # noqa: F401,F403
