"""ctypes binding of the FP64 CPU oracle (oracle/_build/libclip_oracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg.
"""
import ctypes
import os
import subprocess

import numpy as np

from clip_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# CLIP_ORACLE_LIB selects another build of the same source (the mutation
# checks of tests/test_oracle_mutants.py load planted-mistake builds this way)
LIB = os.environ.get("CLIP_ORACLE_LIB") or os.path.join(ROOT, "oracle", "_build", "libclip_oracle.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.run(["sh", os.path.join(ROOT, "oracle", "build.sh")], check=True)
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int)
    lib.orc_create.restype = P
    lib.orc_create.argtypes = [ctypes.POINTER(_abi.ClipConfig)]
    for n in ("orc_init", "orc_collide", "orc_macroscopic"):
        getattr(lib, n).restype = ctypes.c_int
        getattr(lib, n).argtypes = [P]
    lib.orc_step.restype = ctypes.c_int
    lib.orc_step.argtypes = [P, ctypes.c_int]
    lib.orc_field_size.restype = ctypes.c_longlong
    lib.orc_field_size.argtypes = [P, ctypes.c_int]
    lib.orc_get.restype = ctypes.c_int
    lib.orc_get.argtypes = [P, ctypes.c_int, D]
    lib.orc_set.restype = ctypes.c_int
    lib.orc_set.argtypes = [P, ctypes.c_int, D]
    lib.orc_lattice.restype = ctypes.c_int
    lib.orc_lattice.argtypes = [ctypes.c_int, D, D, I, D, I, I]
    lib.orc_filter.restype = ctypes.c_int
    lib.orc_filter.argtypes = [ctypes.c_int, D, ctypes.c_double]
    lib.orc_sdf.restype = ctypes.c_double
    lib.orc_sdf.argtypes = [ctypes.POINTER(_abi.ClipConfig), ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.orc_phase.restype = ctypes.c_int
    lib.orc_phase.argtypes = [P, ctypes.c_char_p]
    lib.orc_get_prev.restype = ctypes.c_int
    lib.orc_get_prev.argtypes = [P, ctypes.c_int, D]
    lib.orc_set_prev.restype = ctypes.c_int
    lib.orc_set_prev.argtypes = [P, ctypes.c_int, D]
    lib.orc_gamma.restype = ctypes.c_double
    lib.orc_gamma.argtypes = [ctypes.c_int, ctypes.c_int, D]
    lib.orc_destroy.restype = None
    lib.orc_destroy.argtypes = [P]
    lib.orc_viscous_force.argtypes = [P, D, ctypes.c_double, D, D]
    lib.orc_side_map.restype = ctypes.c_int
    lib.orc_side_map.argtypes = [ctypes.c_int, ctypes.c_int, I]
    lib.orc_init_phi.argtypes = [P, D]
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """FP64 CPU restatement of CLIP's step in CLIP's AoS layout."""

    def __init__(self, config):
        self.config = config
        self.lib = load()
        self.h = self.lib.orc_create(ctypes.byref(config.raw))
        if not self.h:
            raise RuntimeError("oracle: bad config")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def init(self):
        self.lib.orc_init(self.h)
        return self

    def init_from_phi(self, phi_grid):
        """Initial state from phi on every node (grid shape, ghosts included)."""
        a = np.ascontiguousarray(phi_grid, dtype=np.float64).ravel()
        self.lib.orc_init_phi(self.h, _dptr(a))
        return self

    def step(self, n=1):
        self.lib.orc_step(self.h, int(n))
        return self

    def collide(self):
        self.lib.orc_collide(self.h)

    def macroscopic(self):
        self.lib.orc_macroscopic(self.h)

    def phase(self, *names):
        for n in names:
            if self.lib.orc_phase(self.h, n.encode()) != 0:
                raise ValueError(n)
        return self

    def get_prev(self, which):
        ex = self.config.extents()
        out = np.empty(ex[0] * ex[1] * ex[2] * self.config.q)
        if self.lib.orc_get_prev(self.h, which, _dptr(out)) != 0:
            raise ValueError("no outflow side")
        return out.reshape(ex[0], ex[1], ex[2], -1)

    def set_prev(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if self.lib.orc_set_prev(self.h, which, _dptr(a)) != 0:
            raise ValueError("no outflow side")

    def get(self, field):
        fid = _abi.FIELD[field]
        n = self.lib.orc_field_size(self.h, fid)
        out = np.empty(n, dtype=np.float64)
        self.lib.orc_get(self.h, fid, _dptr(out))
        ex = self.config.extents()
        return out.reshape(ex[0], ex[1], ex[2], -1)

    def set(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self.lib.orc_set(self.h, _abi.FIELD[field], _dptr(a))

    def interior(self, field):
        return self.get(field)[self.config.interior_slices()]


def lattice(dim):
    lib = load()
    Q = 9 if dim == 2 else 19
    M = np.zeros(Q * Q)
    Minv = np.zeros(Q * Q)
    e = np.zeros(Q * 3, dtype=np.int32)
    w = np.zeros(Q)
    inl = np.zeros(6 * 5, dtype=np.int32)
    nin = ctypes.c_int()
    lib.orc_lattice(dim, _dptr(M), _dptr(Minv), e.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dptr(w),
                    inl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(nin))
    k = nin.value
    return {"M": M.reshape(Q, Q), "Minv": Minv.reshape(Q, Q), "e": e.reshape(Q, 3), "w": w,
            "in": inl[: 2 * dim * k].reshape(2 * dim, k)}


def side_map(dim, slip):
    """The reference's wall map (slip False) or slip map (slip True) as the
    oracle holds it: one list of in-directions per side."""
    nin = 3 if dim == 2 else 5
    out = np.zeros(2 * dim * nin, dtype=np.int32)
    load().orc_side_map(dim, int(slip), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return out.reshape(2 * dim, nin)


def gamma(dim, q, u):
    uu = np.zeros(3)
    uu[: len(u)] = u
    return load().orc_gamma(dim, q, _dptr(uu))


def wmrt_filter(dim, v, s):
    a = np.ascontiguousarray(v, dtype=np.float64).copy()
    load().orc_filter(dim, _dptr(a), float(s))
    return a
