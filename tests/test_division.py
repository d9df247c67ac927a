"""div_rn (clip_amd/csrc/node.hpp), the quotient the FP64 kernels use for
divisors that repeat, equals IEEE division: the constant denominators of the
WMRT filter and the stencils, and random variable divisors (density,
|grad phi|), over operands spanning many binades.  The FP64 parity with the
oracle (which divides plainly) depends on it."""
import ctypes
import os

import numpy as np
import pytest

LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libnode_check.so")
D_ = ctypes.POINTER(ctypes.c_double)


def lib():
    if not os.path.exists(LIBPATH):
        pytest.skip("tests/_build/libnode_check.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIBPATH)
    L.div_rn_batch.argtypes = [D_, ctypes.c_longlong, ctypes.c_double, D_]
    L.div_rn_pairs.argtypes = [D_, D_, ctypes.c_longlong, D_]
    return L


def random_doubles(rng, n, emin, emax):
    mant = rng.integers(0, 1 << 52, n, dtype=np.uint64)
    exp = rng.integers(emin, emax, n).astype(np.uint64) + np.uint64(1023)
    sign = rng.integers(0, 2, n).astype(np.uint64) << np.uint64(63)
    return ((exp << np.uint64(52)) | mant | sign).view(np.float64)


@pytest.mark.parametrize("d", [3.0, 6.0, 9.0, 12.0, 36.0, 48.0, 72.0, 144.0])
def test_constant_divisors(d):
    L = lib()
    rng = np.random.default_rng(int(d))
    x = random_doubles(rng, 2_000_000, -80, 10)
    out = np.empty_like(x)
    L.div_rn_batch(x.ctypes.data_as(D_), x.size, d, out.ctypes.data_as(D_))
    assert np.array_equal(out, x / d)


def test_variable_divisors():
    L = lib()
    rng = np.random.default_rng(5)
    x = random_doubles(rng, 2_000_000, -50, 10)
    d = np.abs(random_doubles(rng, 2_000_000, -28, 2))
    out = np.empty_like(x)
    L.div_rn_pairs(x.ctypes.data_as(D_), d.ctypes.data_as(D_), x.size, out.ctypes.data_as(D_))
    assert np.array_equal(out, x / d)
