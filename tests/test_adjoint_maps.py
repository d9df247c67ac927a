"""The host-built maps the reverse step transposes (clip_amd/csrc/
adjoint_maps.hpp, host build in tests/_build/libnode_check.so), checked on
the CPU against the FP64 oracle's own boundary phases on random states:

* SymPuller's terms, evaluated with the state's values, give the oracle's
  post-streaming populations (kernelStreaming + kernelHalfBounceBack +
  kernelFreeConvect + kernelNeumann after kernelPeriodicBoundary and
  JetBoundary, NsAllen.cu:881-892) at every interior node;
* ring_sources gives, for every ghost cell of the new phi, the cell of the
  old phi that the periodic, inlet and mirror copies leave there
  (Solver.cu:75-141, 550-921, 927-978)."""
import ctypes
import os

import numpy as np
import pytest

from clip_amd import Config, _abi
from tests import cases
from tests.oracle_lib import Oracle

LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libnode_check.so")

CASES = [
    ("jet2d", cases.jet2d(12, 10, radius=3.0)),
    ("box2d_mixed", cases.box2d([("x-", "wall"), ("x+", "do nothing"), ("y-", "free convect"), ("y+", "neumann")],
                                n=10, ny=9)),
    ("box2d_fc_x", cases.box2d([("x-", "free convect"), ("x+", "neumann"), ("y-", "wall"), ("y+", "periodic")],
                               n=9, ny=8)),
    ("rti2d", cases.rti2d(12, 16)),
    # a disc inside the periodic seams: where a disc crosses a seam the inlet
    # edge ghosts differ by design (resolved at their wrapped position)
    ("jet3d_c5", cases.jet3d_c5(nx=8, ny=8, nz=10, radius=3.0)),
    ("jet3d_neumann", cases.jet3d(n=(8, 10, 8), radius=2.5, inlet_side="z-", outlet_type="neumann")),
    ("box3d_walls", cases.bubble3d(n=6, radius=2, sides=[("x-", "wall"), ("x+", "free convect"), ("y-", "neumann"),
                                                         ("y+", "wall"), ("z-", "periodic"), ("z+", "periodic")])),
]


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIBPATH):
        pytest.skip("tests/_build/libnode_check.so not built (run __graft_entry__.build())")
    lb = ctypes.CDLL(LIBPATH)
    D = ctypes.POINTER(ctypes.c_double)
    lb.maps_pull_eval.argtypes = [ctypes.POINTER(_abi.ClipConfig), D, D, D, D]
    lb.maps_ring.argtypes = [ctypes.POINTER(_abi.ClipConfig), ctypes.POINTER(ctypes.c_longlong)]
    return lb


def soa(a):
    """Oracle grid array (ex, ey, ez, ndof) -> dense SoA (ndof, ez, ey, ex) flat."""
    return np.ascontiguousarray(np.transpose(a, (3, 2, 1, 0))).ravel()


def planes(cfg, arr):
    """Outflow plane sets (free-convect sides, ascending), entry q*plane + t,
    t over the two tangential axes (lower axis fastest) at the boundary node."""
    ex = cfg.extents()
    D, Q = cfg.dim, cfg.q
    out = []
    for s in range(2 * D):
        if cfg.raw.bc_type[s] != _abi.BC["free convect"]:
            continue
        a = s // 2
        t1 = 1 if a == 0 else 0
        t2 = 1 if a == 2 else 2
        b = ex[a] - 2 if s % 2 else 1
        n1, n2 = ex[t1], ex[t2]
        blk = np.zeros((Q, n2, n1))
        for i2 in range(n2):
            for i1 in range(n1):
                c = [0, 0, 0]
                c[a], c[t1], c[t2] = b, i1, i2
                blk[:, i2, i1] = arr[c[0], c[1], c[2], :]
        out.append(blk.ravel())
    return np.concatenate(out) if out else np.zeros(1)


@pytest.mark.parametrize("name,text", CASES, ids=[c[0] for c in CASES])
def test_symbolic_chain_matches_oracle_boundary_phases(lib, name, text):
    cfg = Config.from_string(text)
    o = Oracle(cfg).init()
    ex = cfg.extents()
    Q, D = cfg.q, cfg.dim
    rng = np.random.default_rng(5)
    f = rng.uniform(0.01, 0.2, size=(*ex, Q))
    prev = rng.uniform(0.01, 0.2, size=(*ex, Q))
    u = rng.uniform(-0.08, 0.08, size=(*ex, D))
    o.set("f", f)
    o.set("f_post", prev)  # streaming first copies f_post into f_prev
    o.set("u", u)
    o.phase("periodic", "velocity")
    F = o.get("f").copy()
    o.phase("stream", "wall", "free_convect", "neumann")
    want = o.interior("f_post")
    out = np.zeros(Q * ex[0] * ex[1] * ex[2])
    Fs, Ps, Us = soa(F), planes(cfg, prev), soa(u)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.maps_pull_eval(ctypes.byref(cfg.raw), Fs.ctypes.data_as(dp), Ps.ctypes.data_as(dp), Us.ctypes.data_as(dp),
                       out.ctypes.data_as(dp))
    got = np.transpose(out.reshape(Q, ex[2], ex[1], ex[0]), (3, 2, 1, 0))[cfg.interior_slices()]
    err = np.abs(got - want).max()
    print(name, "max |terms - oracle|", err)
    assert err < 1e-15


@pytest.mark.parametrize("name,text", CASES, ids=[c[0] for c in CASES])
def test_ring_map_matches_oracle_ghost_copies(lib, name, text):
    cfg = Config.from_string(text)
    o = Oracle(cfg).init()
    ex = cfg.extents()
    rng = np.random.default_rng(9)
    phi0 = rng.uniform(0.0, 1.0, size=(*ex, 1))
    o.set("phi", phi0)
    o.phase("periodic", "velocity", "mirror")
    phi1 = o.get("phi")[..., 0]
    src = np.zeros(ex[0] * ex[1] * ex[2], dtype=np.int64)
    lib.maps_ring(ctypes.byref(cfg.raw), src.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
    flat0 = soa(phi0)
    want = np.where(src >= 0, flat0[np.maximum(src, 0)], flat0)
    got = soa(phi1[..., None])
    assert np.array_equal(got, want)
    assert (src >= 0).sum() > 0
