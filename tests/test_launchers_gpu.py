"""The reference Solver's boundary launchers (Solver.cuh:61-107,
Solver.cu:985-1062) through the C-ABI on caller-owned device arrays in CLIP's
AoS layout, against the FP64 oracle's passes on the same random arrays
(bit-identical: each copies values or applies the same operations in the same
order); and NSAllen::deviceInitializer from a caller's phase field
(clip_init_from_phi) against clip_init_fresh and the oracle."""
import ctypes

import numpy as np
import pytest

from clip_amd import ClipError, Config, Solver, _abi
from tests import cases
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Dev:
    """A device copy of a float64 array, allocated through the HIP runtime
    the library is bound to (the caller's side of the launcher contract)."""

    _hip = None

    def __init__(self, arr):
        if Dev._hip is None:
            _abi.load()
            lib = ctypes.CDLL("libamdhip64.so.7")
            lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            lib.hipFree.argtypes = [ctypes.c_void_p]
            Dev._hip = lib
        self.shape = arr.shape
        self.host = np.ascontiguousarray(arr, dtype=np.float64)
        self.buf = ctypes.c_void_p()
        assert Dev._hip.hipMalloc(ctypes.byref(self.buf), self.host.nbytes) == 0
        assert Dev._hip.hipMemcpy(self.buf, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    @property
    def ptr(self):
        return self.buf.value

    def get(self):
        out = np.empty_like(self.host)
        assert Dev._hip.hipMemcpy(out.ctypes.data, self.buf, out.nbytes, D2H) == 0
        return out

    def __del__(self):
        if self.buf:
            Dev._hip.hipFree(self.buf)


def _rand(cfg, ndof, seed):
    ex = cfg.extents()
    return np.random.default_rng(seed).uniform(0.1, 1.0, size=(ex[0], ex[1], ex[2], ndof))


def _setup(text):
    cfg = Config.from_string(text)
    print("launcher case dim", cfg.dim, "bc", list(cfg.bc_type)[: 2 * cfg.dim], flush=True)
    return cfg, Solver(cfg, "f64"), Oracle(cfg).init()


def _interior(cfg, a):
    return a[cfg.interior_slices()]


BOXES = {
    "box2d_walls": cases.box2d([("x-", "wall"), ("x+", "slip wall"), ("y-", "wall"), ("y+", "wall")], n=10, ny=8),
    "box2d_fc_neumann": cases.box2d([("x-", "neumann"), ("x+", "free convect"), ("y-", "free convect"),
                                     ("y+", "neumann")], n=10, ny=8),
    "box2d_periodic_x": cases.box2d([("x-", "periodic"), ("x+", "periodic"), ("y-", "slip wall"),
                                     ("y+", "free convect")], n=10, ny=8),
    "box3d_mixed": cases.box3d([("x-", "wall"), ("x+", "slip wall"), ("y-", "free convect"), ("y+", "neumann"),
                                ("z-", "periodic"), ("z+", "periodic")], n=(6, 7, 5)),
    "box3d_slip": cases.box3d([("x-", "slip wall"), ("x+", "slip wall"), ("y-", "slip wall"), ("y+", "wall"),
                               ("z-", "free convect"), ("z+", "neumann")], n=(6, 7, 5)),
    "bubble3d_periodic": cases.bubble3d(n=6, radius=2),
}


@pytest.mark.parametrize("name", sorted(BOXES))
def test_periodic_and_mirror_launchers(name):
    cfg, s, o = _setup(BOXES[name])
    f, phi = _rand(cfg, cfg.q, 1), _rand(cfg, 1, 2)
    df, dp = Dev(f), Dev(phi)
    s.periodic_boundary(cfg.q, df.ptr)
    s.periodic_boundary(1, dp.ptr)
    o.set("f", f)
    o.set("g", f)
    o.set("phi", phi)
    o.phase("periodic")
    assert np.array_equal(df.get(), o.get("f"))
    assert np.array_equal(dp.get(), o.get("phi"))
    s.mirror_boundary(dp.ptr)
    o.phase("mirror")
    assert np.array_equal(dp.get(), o.get("phi"))


@pytest.mark.parametrize("name", sorted(BOXES))
def test_wall_slip_neumann_launchers(name):
    cfg, s, o = _setup(BOXES[name])
    f, fp, g, gp = (_rand(cfg, cfg.q, k) for k in (3, 4, 5, 6))
    o.set("f", f)
    o.set("f_post", fp)
    o.set("g", g)
    o.set("g_post", gp)
    df, dfp, dg, dgp = Dev(f), Dev(fp), Dev(g), Dev(gp)
    s.wall_boundary(df.ptr, dfp.ptr, dg.ptr, dgp.ptr)
    o.phase("wall")
    assert np.array_equal(_interior(cfg, dfp.get()), o.interior("f_post"))
    assert np.array_equal(_interior(cfg, dgp.get()), o.interior("g_post"))
    s.slip_wall_boundary(df.ptr, dfp.ptr, dg.ptr, dgp.ptr)
    o.phase("slip_wall")
    assert np.array_equal(_interior(cfg, dfp.get()), o.interior("f_post"))
    assert np.array_equal(_interior(cfg, dgp.get()), o.interior("g_post"))
    s.neumann_boundary(dfp.ptr, dgp.ptr)
    o.phase("neumann")
    assert np.array_equal(_interior(cfg, dfp.get()), o.interior("f_post"))
    assert np.array_equal(_interior(cfg, dgp.get()), o.interior("g_post"))


@pytest.mark.parametrize("name", ["box2d_fc_neumann", "box2d_periodic_x", "box3d_mixed", "box3d_slip"])
def test_free_convect_launcher(name):
    cfg, s, o = _setup(BOXES[name])
    fp, gp, fprev, gprev = (_rand(cfg, cfg.q, k) for k in (7, 8, 9, 10))
    u = (_rand(cfg, cfg.dim, 11) - 0.55) * 0.2
    o.set("f_post", fp)
    o.set("g_post", gp)
    o.set_prev(0, fprev)
    o.set_prev(1, gprev)
    o.set("u", u)
    dfp, dgp, dfv, dgv, du = Dev(fp), Dev(gp), Dev(fprev), Dev(gprev), Dev(u)
    s.free_convect_boundary(du.ptr, dfp.ptr, dfv.ptr, dgp.ptr, dgv.ptr)
    o.phase("free_convect")
    assert np.array_equal(_interior(cfg, dfp.get()), o.interior("f_post"))
    assert np.array_equal(_interior(cfg, dgp.get()), o.interior("g_post"))


@pytest.mark.parametrize("text", [cases.jet2d(16, 20, radius=4), cases.jet3d((10, 12, 10), radius=3),
                                  cases.jet3d_c5(nx=12, ny=12, nz=10, radius=4.0, L=8.0, U=0.02)],
                         ids=["jet2d", "jet3d", "jet3d_modes"])
def test_velocity_launcher(text):
    cfg, s, o = _setup(text)
    f, g = _rand(cfg, cfg.q, 12), _rand(cfg, cfg.q, 13) * 0.1
    phi = _rand(cfg, 1, 14)
    o.set("f", f)
    o.set("g", g)
    o.set("phi", phi)
    df, dg, dp = Dev(f), Dev(g), Dev(phi)
    s.velocity_boundary(dp.ptr, df.ptr, dg.ptr)
    o.phase("velocity")
    # identical operations in the same order, without contraction in both
    assert np.array_equal(dp.get(), o.get("phi"))
    assert np.array_equal(df.get(), o.get("f"))
    assert np.array_equal(dg.get(), o.get("g"))


def test_launchers_reject_host_memory_and_wrong_pairs():
    cfg, s, _ = _setup(BOXES["box2d_walls"])
    host = np.zeros(cfg.extents()[0] * cfg.extents()[1] * cfg.q)
    with pytest.raises(ClipError):
        s.periodic_boundary(cfg.q, host.ctypes.data)
    a = Dev(_rand(cfg, cfg.q, 15))
    with pytest.raises(ClipError):
        s.wall_boundary(a.ptr, a.ptr, a.ptr, None)  # dev_b without dev_b_post
    with pytest.raises(ClipError):
        s.periodic_boundary(0, a.ptr)


@pytest.mark.parametrize("text", [cases.drop2d(n=24, radius=6), cases.jet3d((10, 12, 10), radius=3)],
                         ids=["drop2d", "jet3d"])
def test_init_from_phi(text):
    """clip_init_from_phi with initialCondition's field equals
    clip_init_fresh bit for bit; with a perturbed field it matches the
    oracle's deviceInitializer on that field, also after 20 steps."""
    cfg = Config.from_string(text)
    phi0 = cfg.initial_phi()
    a = Solver(cfg, "f64").init_fresh()
    b = Solver(cfg, "f64").init_from_phi(phi0)
    for f in ("phi", "p", "u", "f", "g"):
        assert np.array_equal(a.get(f), b.get(f)), f
    # another smooth field in [0, 1]: the profile shifted along x (periodic in
    # both cases) and sharpened
    pert = np.roll(phi0, 2, axis=0) ** 1.5
    c = Solver(cfg, "f64").init_from_phi(pert)
    o = Oracle(cfg).init_from_phi(pert)
    for f in ("phi", "p", "f", "g"):
        assert np.abs(c.interior(f) - o.interior(f)).max() <= 1e-14, f
    c.step(20)
    o.step(20)
    for f in ("phi", "u", "p"):
        x, y = c.interior(f), o.interior(f)
        e = np.linalg.norm(x - y) / np.linalg.norm(y)
        print(f"init_from_phi, 20 steps: {f} rel-L2 {e:.2e}", flush=True)
        assert e <= 1e-11, f
    with pytest.raises(ValueError):
        Solver(cfg, "f64").init_from_phi(phi0[:-1])
