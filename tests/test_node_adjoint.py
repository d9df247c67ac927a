"""The per-node reverse (node.hpp, host build tests/_build/libnode_check.so)
against central finite differences of the per-node forward: the moment-form
collision/macroscopic update of a full step, and the initial state with the
first collision.  Checked as directional derivatives along random
directions, per input group, in FP64."""
import ctypes
import os

import numpy as np
import pytest

from tests import oracle_lib

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "_build", "libnode_check.so")
D_ = ctypes.POINTER(ctypes.c_double)


def lib():
    if not os.path.exists(LIBPATH):
        pytest.skip("tests/_build/libnode_check.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIBPATH)
    dd = ctypes.c_double
    L.node_fwd.argtypes = [ctypes.c_int, D_, D_, D_, dd, D_, dd, D_, D_, D_, D_]
    L.node_rev.argtypes = [ctypes.c_int, D_, D_, D_, dd, D_, dd, D_, D_, D_, D_, D_, D_, D_]
    L.init_fwd.argtypes = [ctypes.c_int, D_, dd, dd, dd, D_, dd, D_, D_]
    L.init_rev.argtypes = [ctypes.c_int, D_, dd, dd, dd, D_, dd, D_, D_, D_]
    return L


def P(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(D_)


def phys_vec(case=1, sigma=1e-3, W=4.0, mob=0.02):
    return np.array([1.0, 0.01, 0.05, 0.2, 8 * sigma / W, 1.5 * sigma * W, mob, W, 0.0, -1e-5, 0.0, case])


def inputs(dim, rng):
    lat = oracle_lib.lattice(dim)
    w = lat["w"]
    Q = len(w)
    p = 0.02
    c0 = rng.uniform(0.2, 0.8)
    return {
        "fp": p * w + 1e-3 * rng.standard_normal(Q) * w,
        "gp": c0 * w + 1e-2 * rng.standard_normal(Q) * w,
        "c0": np.array([c0]),
        "dc": 0.05 * rng.standard_normal(3),
        "lap": np.array([0.01 * rng.standard_normal()]),
        "uo": 0.01 * rng.standard_normal(3),
    }


def fwd(L, dim, phys, x):
    Q = len(x["fp"])
    fs, gs, u = np.zeros(Q), np.zeros(Q), np.zeros(3)
    args = [P(phys)[1], P(x["fp"])[1], P(x["gp"])[1], float(x["c0"][0]), P(x["dc"])[1], float(x["lap"][0]),
            P(x["uo"])[1], fs.ctypes.data_as(D_), gs.ctypes.data_as(D_), u.ctypes.data_as(D_)]
    keep = [phys, x]  # noqa: F841
    L.node_fwd(dim, *args)
    return fs, gs, u[:dim]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("group", ["all", "fp", "gp", "c0", "dc", "lap", "uo", "beta", "kappa", "mobility"])
def test_step_node_reverse_matches_finite_differences(dim, group):
    L = lib()
    rng = np.random.default_rng(1 + dim)
    x = inputs(dim, rng)
    if dim == 2:
        x["dc"][2] = 0.0
        x["uo"][2] = 0.0
    phys = phys_vec()
    Q = len(x["fp"])
    lfs, lgs, lu = rng.standard_normal(Q), rng.standard_normal(Q), rng.standard_normal(3)
    if dim == 2:
        lu[2] = 0.0
    lfp, lgp, lin = np.zeros(Q), np.zeros(Q), np.zeros(11)
    L.node_rev(dim, P(phys)[1], P(x["fp"])[1], P(x["gp"])[1], float(x["c0"][0]), P(x["dc"])[1],
               float(x["lap"][0]), P(x["uo"])[1], P(lfs)[1], P(lgs)[1], P(lu)[1],
               lfp.ctypes.data_as(D_), lgp.ctypes.data_as(D_), lin.ctypes.data_as(D_))
    grad = {"fp": lfp, "gp": lgp, "c0": lin[0:1], "dc": lin[1:4], "lap": lin[4:5], "uo": lin[5:8]}
    pgrad = {"beta": lin[8], "kappa": lin[9], "mobility": lin[10]}
    pidx = {"beta": 4, "kappa": 5, "mobility": 6}

    def S(xx, ph):
        fs, gs, u = fwd(L, dim, ph, xx)
        return fs @ lfs + gs @ lgs + u @ lu[:dim]

    if group in pidx:
        k = pidx[group]
        h = 1e-4 * abs(phys[k])
        pp, pm = phys.copy(), phys.copy()
        pp[k] += h
        pm[k] -= h
        fd = (S(x, pp) - S(x, pm)) / (2 * h)
        an = pgrad[group]
    else:
        v = {k: (rng.standard_normal(a.shape) if (group == "all" or group == k) else np.zeros(a.shape))
             for k, a in x.items()}
        if dim == 2:
            v["dc"][2] = 0.0
            v["uo"][2] = 0.0
        scale = {k: max(np.abs(a).max(), 1e-3) for k, a in x.items()}
        h = 1e-5
        xp = {k: x[k] + h * scale[k] * v[k] for k in x}
        xm = {k: x[k] - h * scale[k] * v[k] for k in x}
        fd = (S(xp, phys) - S(xm, phys)) / (2 * h)
        an = sum(scale[k] * (grad[k] @ v[k]) for k in x)
    assert abs(fd - an) <= 2e-5 * max(abs(fd), abs(an)) + 1e-9, (group, fd, an)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("case", [1, 0])  # drop, bubble
@pytest.mark.parametrize("group", ["c0", "dc", "lap", "beta", "kappa", "mobility", "sigma"])
def test_init_node_reverse_matches_finite_differences(dim, case, group):
    L = lib()
    rng = np.random.default_rng(7 + dim + 10 * case)
    Q = 9 if dim == 2 else 19
    phys = phys_vec(case=case)
    sigma, half = 1e-3, 25.0
    c0 = rng.uniform(0.2, 0.8)
    dc = 0.05 * rng.standard_normal(3)
    if dim == 2:
        dc[2] = 0.0
    lap = 0.01 * rng.standard_normal()
    lfs, lgs = rng.standard_normal(Q), rng.standard_normal(Q)
    lin = np.zeros(9)
    L.init_rev(dim, P(phys)[1], sigma, half, c0, P(dc)[1], lap, P(lfs)[1], P(lgs)[1], lin.ctypes.data_as(D_))

    def S(ph, sg, c, d, lp):
        fs, gs = np.zeros(Q), np.zeros(Q)
        L.init_fwd(dim, P(ph)[1], sg, half, c, P(d)[1], lp, fs.ctypes.data_as(D_), gs.ctypes.data_as(D_))
        return fs @ lfs + gs @ lgs

    h = 1e-6
    if group == "c0":
        fd = (S(phys, sigma, c0 + h, dc, lap) - S(phys, sigma, c0 - h, dc, lap)) / (2 * h)
        an = lin[0]
    elif group == "dc":
        v = rng.standard_normal(3)
        if dim == 2:
            v[2] = 0
        fd = (S(phys, sigma, c0, dc + h * 0.05 * v, lap) - S(phys, sigma, c0, dc - h * 0.05 * v, lap)) / (2 * h)
        an = 0.05 * (lin[1:4] @ v)
    elif group == "lap":
        fd = (S(phys, sigma, c0, dc, lap + h) - S(phys, sigma, c0, dc, lap - h)) / (2 * h)
        an = lin[4]
    elif group == "sigma":
        hs = 1e-4 * sigma
        fd = (S(phys, sigma + hs, c0, dc, lap) - S(phys, sigma - hs, c0, dc, lap)) / (2 * hs)
        an = lin[8]
    else:
        k = {"beta": 4, "kappa": 5, "mobility": 6}[group]
        hk = 1e-4 * abs(phys[k])
        pp, pm = phys.copy(), phys.copy()
        pp[k] += hk
        pm[k] -= hk
        fd = (S(pp, sigma, c0, dc, lap) - S(pm, sigma, c0, dc, lap)) / (2 * hk)
        an = lin[{"beta": 5, "kappa": 6, "mobility": 7}[group]]
    assert abs(fd - an) <= 2e-5 * max(abs(fd), abs(an)) + 1e-9, (group, fd, an)


@pytest.mark.parametrize("dim", [2, 3])
def test_moment_form_equals_reference_filter_collision(dim):
    """The moment-form node update (node.hpp) equals the reference's
    two-pass WMRT collision computed with the oracle's filter
    (kernelCollisionMRTf NsAllen.cu:589-763, calculateVF 102-132)."""
    L = lib()
    rng = np.random.default_rng(3)
    lat = oracle_lib.lattice(dim)
    e, w = lat["e"].astype(float), lat["w"]
    Q = len(w)
    x = inputs(dim, rng)
    if dim == 2:
        x["dc"][2] = 0.0
        x["uo"][2] = 0.0
    phys = phys_vec()
    fs, gs, u = fwd(L, dim, phys, x)
    rho_h, rho_l, tau_h, tau_l, beta, kappa, M, W = phys[:8]
    c0, dc, lap, fp, uo = x["c0"][0], x["dc"][:dim], x["lap"][0], x["fp"], x["uo"][:dim]
    rho = rho_l + c0 * (rho_h - rho_l)
    drho = rho_h - rho_l
    tau = tau_l + c0 * (tau_h - tau_l)
    s = 1 / (tau + 0.5)
    mu = 4 * beta * c0 * (c0 - 1) * (c0 - 0.5) - kappa * lap
    fg = (rho - rho_l) * phys[8:8 + dim]
    p = fp.sum()

    def Gam(v):
        eu = e[:, :dim] @ v
        return w * (3 * eu + 4.5 * eu ** 2 - 1.5 * (v @ v))

    def visc(gneq_f):
        Pi = np.einsum("qa,qb,q->ab", e[:, :dim], e[:, :dim], gneq_f)
        return -tau * drho * (Pi @ dc)

    gneq = oracle_lib.wmrt_filter(dim, fp - (p * w + Gam(uo)), s)
    F = (mu - p * drho / 3) * dc + visc(gneq) + fg
    u_ref = e[:, :dim].T @ fp + 0.5 * F / rho
    gneq = oracle_lib.wmrt_filter(dim, fp - (p * w + Gam(u_ref)), s)
    F = (mu - p * drho / 3) * dc + visc(gneq) + fg
    hlp = 3 * w * (e[:, :dim] @ F) / rho
    gneq2 = oracle_lib.wmrt_filter(dim, fp - (p * w + Gam(u_ref) - 0.5 * hlp), s)
    f_ref = fp - gneq2 + hlp
    assert np.allclose(u, u_ref, rtol=1e-12, atol=1e-16)
    assert np.allclose(fs, f_ref, rtol=1e-11, atol=1e-16)
