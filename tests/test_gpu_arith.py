"""GPU check of the pair terms' square root and quotient (sqrt_pair,
div_pair in pbf-cuda_amd/csrc/pbf_device.h) against the IEEE operations:
every positive float from 2^-90 to 2^10 through the square root (the pair
terms use [0.99 KERNAL_EPS^2, 1.0001 h^2)), every float below 2^-90 staying
under KERNAL_EPS, and 2^28 quotients with numerators log-uniform in
[1e-18, 1e6] (both signs) over denominators log-uniform in [1e-4, 1e4].  The
pair terms' bit-exactness against the oracle's FP32 mirror (std::sqrt, '/')
rests on these being the IEEE results."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pbf-cuda_amd",
                   "pbf", "libpbf_arith_check.so")


def _bits(x):
    return int(np.float32(x).view(np.uint32))


def test_pair_sqrt_and_division_are_ieee(device):
    import torch  # noqa: F401  (the HIP runtime torch loaded)

    lib = ctypes.CDLL(LIB)
    fn = lib.pbf_check_pair_arith
    fn.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                   ctypes.c_ulonglong, ctypes.c_ulonglong * 4]
    fn.restype = ctypes.c_int
    out = (ctypes.c_ulonglong * 4)()
    lo, hi = _bits(2.0 ** -90), _bits(2.0 ** 10)
    assert lo <= _bits(0.99e-8) and _bits(0.1 * 0.1 * 1.0001) < hi
    rc = fn(lo, hi, lo, np.float32(1e-4), 1 << 28, out)
    assert rc == 0
    bad_sqrt, bad_small, bad_div, total = (int(v) for v in out)
    print(f"tested {total}: sqrt mismatches {bad_sqrt}, small-input failures {bad_small}, "
          f"quotient mismatches {bad_div}")
    assert total == (hi - lo) + lo + (1 << 28)
    assert bad_sqrt == 0 and bad_small == 0 and bad_div == 0
