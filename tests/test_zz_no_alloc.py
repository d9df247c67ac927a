"""clip_sum_noalloc (the allocation-free variant of clip_sum) through the
C-ABI: on two calls in a row, the second with a smaller grid, it returns the
same FP64 sum as clip_sum over the same phase field, and counts non-finite
values.  The test itself allocates the device buffers through the HIP
runtime the library links (the caller's side of the contract)."""
import ctypes

import numpy as np
import pytest

from clip_amd import Config, Solver, _abi
from tests import cases

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    # The HIP runtime libclip_amd.so is bound to: opened by soname after the
    # library, the dynamic linker returns the copy already loaded under that
    # name (torch's wheel bundles one with the same soname; whichever loaded
    # first serves both).
    _abi.load()
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    return lib


def _noalloc_sum(hip, arr, precision, dim, ext, nonfinite=True):
    lib = _abi.load()
    buf, out, bad = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(buf), arr.nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(out), 8) == 0
    assert hip.hipMalloc(ctypes.byref(bad), 4) == 0
    try:
        assert hip.hipMemcpy(buf, arr.ctypes.data, arr.nbytes, H2D) == 0
        e = (ctypes.c_uint * 3)(*ext)
        rc = lib.clip_sum_noalloc(buf, _abi.PRECISION[precision], dim, e, out, bad if nonfinite else None, None)
        _abi.check(rc)
        assert hip.hipDeviceSynchronize() == 0
        s = np.zeros(1, dtype=np.float64)
        n = np.zeros(1, dtype=np.uint32)
        assert hip.hipMemcpy(s.ctypes.data, out, 8, D2H) == 0
        assert hip.hipMemcpy(n.ctypes.data, bad, 4, D2H) == 0
        return float(s[0]), int(n[0])
    finally:
        for p in (buf, out, bad):
            hip.hipFree(p)


def _case(text, precision, steps):
    cfg = Config.from_string(text)
    sol = Solver(cfg, precision).init_fresh().step(steps)
    want = sol.sum("phi")
    phi = sol.get("phi", layout="soa")
    arr = np.ascontiguousarray(phi if precision == "f64" else phi.astype(np.float32))
    ext = cfg.extents()
    sol.close()
    return cfg.dim, ext, arr, want


def test_noalloc_sum_matches_clip_sum_on_two_calls():
    hip = _hip()
    # first call: 3D FP64 grid; second call: a smaller 2D FP32 grid
    for text, prec, steps in ((cases.bubble3d(n=24, radius=6), "f64", 5),
                              (cases.drop2d(n=20, radius=5), "f32", 5)):
        dim, ext, arr, want = _case(text, prec, steps)
        got, bad = _noalloc_sum(hip, arr, prec, dim, ext)
        print(f"dim {dim} ext {ext} {prec}: clip_sum {want!r} clip_sum_noalloc {got!r}", flush=True)
        assert bad == 0
        assert got == want


def test_noalloc_sum_counts_nonfinite_and_rejects_bad_sizes():
    hip = _hip()
    dim, ext, arr, want = _case(cases.drop2d(n=16, radius=4), "f64", 2)
    i = (ext[0] + 1) * 3  # an interior node
    arr = arr.copy()
    arr[i] = np.nan
    got, bad = _noalloc_sum(hip, arr, "f64", dim, ext)
    assert bad == 1
    assert np.isfinite(got)
    lib = _abi.load()
    e = (ctypes.c_uint * 3)(1 << 21, 4, 1)
    rc = lib.clip_sum_noalloc(ctypes.c_void_p(1), 0, 2, e, ctypes.c_void_p(1), None, None)
    assert rc == _abi.ERR_ARG and b"outside" in lib.clip_last_error()
