"""Pins the FP64 CPU oracle with checks that do not restate its code:
closed forms, invariants, special cases and brute force on tiny grids.

The reference ships no tests or golden vectors for this path, so these
checks (plus the committed fixtures of test_golden.py) are what fix the
oracle; each names the reference code it pins.
"""
import itertools

import numpy as np
import pytest

from clip_amd import Config
from tests import cases
from tests import oracle_lib
from tests.oracle_lib import Oracle


def lat(dim):
    return oracle_lib.lattice(dim)


# ---------------------------------------------------------------- lattice
@pytest.mark.parametrize("dim", [2, 3])
def test_velocity_set_and_weights(dim):
    """WMRTvelSet (WMRT.cuh:189-206): isotropy moments of the weights."""
    L = lat(dim)
    e, w = L["e"][:, :dim].astype(float), L["w"]
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(np.einsum("q,qa->a", w, e), 0, atol=1e-16)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, e, e), np.eye(dim) / 3, atol=1e-16)
    # 4th-order isotropy: sum w e_a e_b e_c e_d = (d_ab d_cd + d_ac d_bd + d_ad d_bc)/9
    I = np.eye(dim)
    iso = (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I) + np.einsum("ad,bc->abcd", I, I)) / 9
    assert np.allclose(np.einsum("q,qa,qb,qc,qd->abcd", w, e, e, e, e), iso, atol=1e-16)
    # every direction has its opposite
    for q in range(len(w)):
        assert any((e[p] == -e[q]).all() for p in range(len(w)))


@pytest.mark.parametrize("dim", [2, 3])
def test_moment_matrix_inverse_and_orthogonality(dim):
    """convert*/reconvert* (WMRT.cuh:54-184): M^-1 M = I; rows orthogonal
    (plain inner product for D2Q9, w-weighted for D3Q19)."""
    L = lat(dim)
    M, Mi, w = L["M"], L["Minv"], L["w"]
    Q = len(w)
    assert np.abs(Mi @ M - np.eye(Q)).max() < 1e-14
    assert np.abs(M @ Mi - np.eye(Q)).max() < 1e-14
    G = M @ M.T if dim == 2 else M @ np.diag(w) @ M.T
    off = G - np.diag(np.diag(G))
    assert np.abs(off).max() < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_moment_rows_carry_mass_and_momentum(dim):
    L = lat(dim)
    M, e = L["M"], L["e"]
    assert np.all(M[0] == 1)
    mom_rows = [3, 5] if dim == 2 else [1, 2, 3]
    for a, r in enumerate(mom_rows):
        assert np.array_equal(M[r], e[:, a])


@pytest.mark.parametrize("dim", [2, 3])
def test_filter_relaxes_viscous_stress_and_keeps_hydrodynamic_modes(dim):
    """The WMRT filter (NsAllen.cu:416-431) scales the traceless second-order
    Hermite mode w_q (e_q.S.e_q) by s and leaves mass/momentum untouched."""
    L = lat(dim)
    e, w = L["e"][:, :dim].astype(float), L["w"]
    rng = np.random.default_rng(0)
    for _ in range(5):
        A = rng.normal(size=(dim, dim))
        S = A + A.T
        S -= np.trace(S) / dim * np.eye(dim)
        v = w * np.einsum("qa,ab,qb->q", e, S, e)
        s = rng.uniform(0.1, 1.9)
        out = oracle_lib.wmrt_filter(dim, v, s)
        assert np.allclose(out, s * v, atol=1e-15)
        u = rng.normal(size=dim) * 0.1
        h = w * (1 + 3 * e @ u)
        assert np.allclose(oracle_lib.wmrt_filter(dim, h, s), h, atol=1e-15)
        r = rng.normal(size=len(w))
        assert np.allclose(oracle_lib.wmrt_filter(dim, r, 1.0), r, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_equilibrium_moments(dim):
    """Equilibrium_new (NsAllen.cu:77-94): sum Gamma = 0, sum e Gamma = u,
    sum e e Gamma = u u (second order in u)."""
    L = lat(dim)
    e = L["e"][:, :dim].astype(float)
    u = np.array([0.031, -0.017, 0.011])[:dim]
    G = np.array([oracle_lib.gamma(dim, q, u) for q in range(len(L["w"]))])
    assert abs(G.sum()) < 1e-17
    assert np.allclose(e.T @ G, u, atol=1e-17)
    assert np.allclose(np.einsum("qa,qb,q->ab", e, e, G), np.outer(u, u), atol=1e-17)


@pytest.mark.parametrize("dim", [2, 3])
def test_side_lists_are_inward_directions(dim):
    """wallBCMap (WMRT.cuh:214-237): each side's list holds the directions
    pointing into the domain, the axis normal first."""
    L = lat(dim)
    e = L["e"]
    for s in range(2 * dim):
        a, sg = s // 2, (1 if s % 2 == 0 else -1)
        lst = list(L["in"][s])
        want = sorted(q for q in range(len(L["w"])) if e[q, a] == sg)
        assert sorted(lst) == want
        assert all(e[lst[0], b] == (sg if b == a else 0) for b in range(3))


# ---------------------------------------------------------------- geometry
def test_sdf_closed_forms():
    """Geometry::sdf (Geometry.cuh:102-161)."""
    lib = oracle_lib.load()
    c = Config.from_string(cases.drop2d(n=32, radius=8))
    cen = (16.5, 16.5)
    assert abs(lib.orc_sdf(c.raw, *cen, 0.0) + 8.0) < 1e-12
    assert abs(lib.orc_sdf(c.raw, cen[0] + 3.0, cen[1] + 4.0, 0.0) - (5.0 - 8.0)) < 1e-12
    b = Config.from_string(cases.bubble3d(n=16, radius=4))
    cc = (18 - 1) / 2.0  # N = n + 2 = 18, centre (N-1)/2
    assert abs(lib.orc_sdf(b.raw, cc + 2, cc + 3, cc + 6, ) - (7.0 - 4.0)) < 1e-12
    # square / cube: outside corner distance and inside distance to nearest face
    sq = c.copy()
    sq.raw.geom_type[0] = 2
    for d, v in enumerate((4.0, 6.0, 0.0)):
        sq.raw.geom_length[0][d] = v
    assert abs(lib.orc_sdf(sq.raw, cen[0], cen[1], 0) - (-2.0)) < 1e-12
    assert abs(lib.orc_sdf(sq.raw, cen[0] + 2 + 3, cen[1] + 3 + 4, 0) - 5.0) < 1e-12
    cu = b.copy()
    cu.raw.geom_type[0] = 3
    for d in range(3):
        cu.raw.geom_length[0][d] = 4.0
    assert abs(lib.orc_sdf(cu.raw, cc, cc, cc) + 2.0) < 1e-12
    assert abs(lib.orc_sdf(cu.raw, cc + 3, cc + 4, cc + 2) - np.sqrt(1 + 4)) < 1e-12
    # perturbation: y - (amp*L*cos(2 pi x / L) + y0)
    r = Config.from_string(cases.rti2d(32, 64, amp=0.1))
    for x in (0.0, 5.0, 8.0):
        want = 40.0 - (0.1 * 32 * np.cos(2 * np.pi * x / 32) + 32.0)
        assert abs(lib.orc_sdf(r.raw, x, 40.0, 0.0) - want) < 1e-12


# ---------------------------------------------------------------- stencils
def _set_phi(o, fn):
    ex = o.config.extents()
    i, j, k = np.meshgrid(np.arange(ex[0]), np.arange(ex[1]), np.arange(ex[2]), indexing="ij")
    phi = fn(i.astype(float), j.astype(float), k.astype(float))
    o.set("phi", phi.reshape(ex[0], ex[1], ex[2], 1))


@pytest.mark.parametrize("text", [cases.drop2d(n=12, radius=3), cases.bubble3d(n=8, radius=2)])
def test_gradient_exact_on_linear_fields(text):
    """Isotropic_Gradient (NsAllen.cu:280-322) is exact for linear phi."""
    cfg = Config.from_string(text)
    o = Oracle(cfg).init()
    a, b, c = 0.013, -0.021, 0.007
    cz = c if cfg.dim == 3 else 0.0
    _set_phi(o, lambda x, y, z: 0.3 + a * x + b * y + cz * z)
    o.phase("gradient")
    g = o.interior("dphi")
    want = [a, b, c][: cfg.dim]
    for d in range(cfg.dim):
        assert np.allclose(g[..., d], want[d], atol=1e-15)


@pytest.mark.parametrize("text", [cases.drop2d(n=12, radius=3), cases.bubble3d(n=8, radius=2)])
def test_laplacian_exact_on_quadratics(text):
    """Chemical_Potential's Laplacian (NsAllen.cu:219-231) is exact for
    quadratic phi: recovered from mu = 4 beta phi(phi-1)(phi-1/2) - kappa lap."""
    cfg = Config.from_string(text)
    o = Oracle(cfg).init()
    rng = np.random.default_rng(1)
    D = cfg.dim
    H = rng.normal(size=(3, 3)) * 1e-3
    H = H + H.T
    if D == 2:
        H[2, :] = H[:, 2] = 0
    lin = rng.normal(size=3) * 1e-2

    def fn(x, y, z):
        X = np.stack([x, y, z])
        return 0.4 + np.einsum("a...,a->...", X, lin) + 0.5 * np.einsum("a...,ab,b...->...", X, H, X)

    _set_phi(o, fn)
    o.phase("chem_pot")
    phi = o.interior("phi")[..., 0]
    mu = o.interior("mu")[..., 0]
    lap = (4 * cfg.beta * phi * (phi - 1) * (phi - 0.5) - mu) / cfg.kappa
    assert np.allclose(lap, np.trace(H), atol=1e-10)


def test_chemical_potential_on_tanh_profile():
    """For phi = 1/2 + 1/2 tanh(2x/W), the continuum mu with beta = 8 sigma/W and
    kappa = 3 sigma W / 2 (InputData.cu:113-114) is 16 sigma/W phi(1-phi)(phi-1/2);
    the 9-point stencil reproduces it to O(1/W^2)."""
    cfg = Config.from_string(cases.drop2d(n=96, radius=8, W=16))
    o = Oracle(cfg).init()
    W = cfg.interface_width
    _set_phi(o, lambda x, y, z: 0.5 + 0.5 * np.tanh(2 * (x - 48.3) / W))
    o.phase("chem_pot")
    phi = o.interior("phi")[..., 0]
    mu = o.interior("mu")[..., 0]
    want = 16 * cfg.sigma / W * phi * (1 - phi) * (phi - 0.5)
    scale = np.abs(want).max()
    assert np.abs(mu - want).max() < 0.05 * scale
    # the other common constant choice (beta = 12 sigma/W) would give mu ~ 0
    assert np.abs(mu).max() > 0.9 * scale


# ---------------------------------------------------------------- streaming / BCs
def _unique_pops(o, seed=0):
    ex = o.config.extents()
    rng = np.random.default_rng(seed)
    f = rng.uniform(0.1, 1.0, size=(ex[0], ex[1], ex[2], o.config.q))
    g = rng.uniform(0.1, 1.0, size=f.shape)
    o.set("f", f)
    o.set("g", g)
    return f, g


@pytest.mark.parametrize("text", [cases.drop2d(n=5, radius=1, ny=7), cases.bubble3d(n=4, radius=1)])
def test_periodic_streaming_is_a_roll(text):
    """kernelPeriodicBoundary + kernelStreaming (Solver.cu:75-141,
    NsAllen.cu:327-364) on a fully periodic grid equal np.roll of the interior
    by e_q, for every direction (brute force on a tiny grid)."""
    cfg = Config.from_string(text)
    o = Oracle(cfg).init()
    f, g = _unique_pops(o)
    o.phase("periodic", "stream")
    L = lat(cfg.dim)
    sl = cfg.interior_slices()
    fi = f[sl]
    fp = o.interior("f_post")
    for q in range(cfg.q):
        shift = tuple(int(L["e"][q, a]) for a in range(cfg.dim))
        axes = tuple(range(cfg.dim))
        want = np.roll(fi[..., q], shift, axis=axes)
        assert np.array_equal(fp[..., q], want), q


def test_half_bounce_back_reflects_and_has_zero_wall_flux():
    """kernelHalfBounceBack (Solver.cu:248-333): at a wall-adjacent node the
    incoming populations are the outgoing ones of the same node, reversed."""
    cfg = Config.from_string(cases.box2d([("x-", "periodic"), ("x+", "periodic"), ("y-", "wall"), ("y+", "wall")],
                                         n=6, ny=5))
    o = Oracle(cfg).init()
    f, _ = _unique_pops(o, 3)
    o.phase("periodic", "stream", "wall")
    fp = o.get("f_post")
    L = lat(2)
    e = L["e"]
    opp = [int(np.where((e == -e[q]).all(axis=1))[0][0]) for q in range(9)]
    for i in range(1, 7):
        for j, (s, sgn) in ((1, (2, 1)), (5, (3, -1))):
            for q in range(9):
                if e[q, 1] == sgn:
                    assert fp[i, j, 0, q] == f[i, j, 0, opp[q]]
            # normal mass flux through the wall link is zero
            out = sum(f[i, j, 0, q] for q in range(9) if e[q, 1] == -sgn)
            inn = sum(fp[i, j, 0, q] for q in range(9) if e[q, 1] == sgn)
            assert abs(out - inn) < 1e-15


def _slip_box(dim):
    if dim == 2:
        sides = [("x-", "slip wall"), ("x+", "slip wall"), ("y-", "wall"), ("y+", "slip wall")]
        return cases.specular(cases.box2d(sides, n=6, ny=5))
    sides = [("x-", "slip wall"), ("x+", "wall"), ("y-", "slip wall"), ("y+", "slip wall"), ("z-", "periodic"),
             ("z+", "periodic")]
    return cases.specular(cases.box3d(sides, n=(5, 6, 4)))


@pytest.mark.parametrize("dim", [2, 3])
def test_specular_slip_keeps_tangential_momentum_and_has_zero_wall_flux(dim):
    """Opt-in slip wall: slipWallBCMap (WMRT.cuh:243-264) applied by the
    half-way wall pass (kernelHalfBounceBack, Solver.cu:248-333).  At a node
    next to one slip side, the populations entering from the wall carry the
    mass and the tangential momentum of the same node's populations that left
    toward it, and the opposite normal momentum (no net flux through the
    wall).  A wall side of the same box reverses the tangential momentum
    instead (bounce-back).  Checked on random populations at every node that
    touches exactly one wall-type side."""
    cfg = Config.from_string(_slip_box(dim))
    assert cfg.raw.slip_specular == 1
    o = Oracle(cfg).init()
    f, g = _unique_pops(o, 5)
    o.phase("periodic", "stream", "wall")
    e = lat(dim)["e"]
    Q = cfg.q
    lo = [1, 1, 1 if dim == 3 else 0]
    hi = [cfg.extents()[a] - 2 if a < dim else 0 for a in range(3)]
    walls = [s for s in range(2 * dim) if cfg.raw.bc_type[s] in (0, 1)]
    checked = {0: 0, 1: 0}
    for name, src, post in (("f", f, o.get("f_post")), ("g", g, o.get("g_post"))):
        for c in itertools.product(*[range(lo[a], hi[a] + 1) for a in range(3)]):
            on = [s for s in walls if c[s // 2] == (hi if s % 2 else lo)[s // 2]]
            if len(on) != 1:
                continue
            s = on[0]
            a, sgn = s // 2, (1 if s % 2 == 0 else -1)
            slip = cfg.raw.bc_type[s] == 1
            inn = [q for q in range(Q) if e[q, a] == sgn]
            out = [q for q in range(Q) if e[q, a] == -sgn]
            vin = np.array([post[c][q] for q in inn])
            vout = np.array([src[c][q] for q in out])
            assert abs(vin.sum() - vout.sum()) < 1e-14, (name, c, "mass")
            for t in range(dim):
                pin = float(np.dot(e[inn, t], vin))
                pout = float(np.dot(e[out, t], vout))
                if t == a:
                    assert abs(pin + pout) < 1e-14, (name, c, "normal flux")
                elif slip:
                    assert abs(pin - pout) < 1e-14, (name, c, t, "tangential momentum kept")
                else:
                    assert abs(pin + pout) < 1e-14, (name, c, t, "tangential momentum reversed")
            checked[int(slip)] += 1
    assert checked[0] > 0 and checked[1] > 0, checked


@pytest.mark.parametrize("dim", [2, 3])
def test_slip_wall_launcher_is_the_specular_pass_without_wall_sides(dim):
    """Solver::slipWallBoundary (Solver.cu:1012-1016) applies slipWallBCMap
    to every wall-type side; on a box whose wall-type sides are all slip
    walls that is exactly the pinned specular pass."""
    if dim == 2:
        sides = [("x-", "slip wall"), ("x+", "slip wall"), ("y-", "free convect"), ("y+", "slip wall")]
        text = cases.box2d(sides, n=6, ny=5)
    else:
        sides = [("x-", "slip wall"), ("x+", "slip wall"), ("y-", "slip wall"), ("y+", "neumann"),
                 ("z-", "periodic"), ("z+", "periodic")]
        text = cases.box3d(sides, n=(5, 6, 4))
    a = Oracle(Config.from_string(text)).init()
    b = Oracle(Config.from_string(cases.specular(text))).init()
    for o in (a, b):
        _unique_pops(o, 9)
    a.phase("slip_wall")
    b.phase("wall")
    for f in ("f_post", "g_post"):
        assert np.array_equal(a.interior(f), b.interior(f))


def test_reference_slip_side_follows_the_wall_pass_gate():
    """Without the opt-in, a slip side bounces back exactly like a wall when
    some side is a wall (kernelHalfBounceBack takes Wall and SlipWall types,
    wallBoundary runs when isWall, Solver.cu:248-333, 998-1003) and is left as
    streamed otherwise (slipWallBoundary is never launched)."""
    walls = [("x-", "periodic"), ("x+", "periodic"), ("y-", "slip wall"), ("y+", "wall")]
    a = Oracle(Config.from_string(cases.box2d(walls, n=6, ny=5))).init()
    b = Oracle(Config.from_string(cases.box2d([walls[0], walls[1], ("y-", "wall"), ("y+", "wall")], n=6,
                                              ny=5))).init()
    for o in (a, b):
        _unique_pops(o, 7)
        o.phase("periodic", "stream", "wall")
    assert np.array_equal(a.get("f_post"), b.get("f_post"))
    lone = [("x-", "periodic"), ("x+", "periodic"), ("y-", "slip wall"), ("y+", "do nothing")]
    c = Oracle(Config.from_string(cases.box2d(lone, n=6, ny=5))).init()
    _unique_pops(c, 7)
    c.phase("periodic", "stream")
    before = c.get("f_post").copy()
    c.phase("wall")
    assert np.array_equal(before, c.get("f_post"))


def test_neumann_copies_inner_neighbour():
    """kernelNeumann (Solver.cu:462-542)."""
    cfg = Config.from_string(cases.box2d([("x-", "periodic"), ("x+", "periodic"), ("y-", "neumann"), ("y+", "neumann")],
                                         n=6, ny=6))
    o = Oracle(cfg).init()
    _unique_pops(o, 4)
    o.phase("periodic", "stream")
    before = o.get("f_post").copy()
    o.phase("neumann")
    after = o.get("f_post")
    e = lat(2)["e"]
    for q in range(9):
        if e[q, 1] == 1:
            assert np.array_equal(after[1:7, 1, 0, q], before[1:7, 2, 0, q])
        if e[q, 1] == -1:
            assert np.array_equal(after[1:7, 6, 0, q], before[1:7, 5, 0, q])


@pytest.mark.parametrize("un", [0.0, 0.03])
def test_free_convect_closed_form(un):
    """kernelFreeConvect (Solver.cu:342-453): a = (a_prev + |u_n| a_inner)/(1+|u_n|);
    with u_n = 0 it returns a_prev."""
    cfg = Config.from_string(cases.box2d([("x-", "periodic"), ("x+", "periodic"), ("y-", "free convect"), ("y+", "wall")],
                                         n=6, ny=6))
    o = Oracle(cfg).init()
    _unique_pops(o, 5)
    ex = cfg.extents()
    u = np.zeros((ex[0], ex[1], 1, 2))
    u[..., 1] = un
    o.set("u", u)
    prev_in = np.random.default_rng(6).uniform(size=(ex[0], ex[1], 1, 9))
    o.set("f_post", prev_in)
    o.phase("periodic", "stream")  # copies f_post into f_prev first
    fp = o.get("f_post").copy()
    o.phase("free_convect")
    out = o.get("f_post")
    e = lat(2)["e"]
    for q in range(9):
        if e[q, 1] == 1:
            want = (prev_in[1:7, 1, 0, q] + un * fp[1:7, 2, 0, q]) / (1 + un)
            assert np.allclose(out[1:7, 1, 0, q], want, rtol=0, atol=1e-16)


def test_velocity_inlet_returns_equilibrium_for_equilibrium_input():
    """JetBoundary (Solver.cu:550-921): if the ghost node holds the inlet
    equilibrium (f = Gamma(U), g = phi (Gamma(U) + w)), the tangential
    corrections vanish and the incoming populations are the equilibrium."""
    cfg = Config.from_string(cases.jet2d(12, 10, radius=30.0))  # disc covers the whole inlet
    o = Oracle(cfg).init()
    ex = cfg.extents()
    U = np.array(cfg.raw.bc_value[3][:2])
    L = lat(2)
    Gm = np.array([oracle_lib.gamma(2, q, U) for q in range(9)])
    phi = o.get("phi")
    f = o.get("f").copy()
    g = o.get("g").copy()
    jg = ex[1] - 1
    for i in range(ex[0]):
        f[i, jg, 0, :] = Gm
        g[i, jg, 0, :] = phi[i, jg, 0, 0] * (Gm + L["w"])
    o.set("f", f)
    o.set("g", g)
    o.phase("velocity")
    f2, g2 = o.get("f"), o.get("g")
    for q in L["in"][3]:
        assert np.allclose(f2[1:-1, jg, 0, q], Gm[q], atol=1e-17)
        assert np.allclose(g2[1:-1, jg, 0, q], phi[1:-1, jg, 0, 0] * (Gm[q] + L["w"][q]), atol=1e-16)


def _inlet_plane_nodes(cfg, side):
    """Ghost nodes of `side`'s plane over the interior tangential range, with
    the boundary node next to each and whether it lies inside geometry 0."""
    lib = oracle_lib.load()
    ex = cfg.extents()
    dim, ax = cfg.dim, side // 2
    ghost = ex[ax] - 1 if side % 2 else 0
    bnd = ghost - 1 if side % 2 else 1
    ranges = [range(1, ex[a] - 1) if a < dim else range(1) for a in range(3)]
    ranges[ax] = [ghost]
    out = []
    for c in itertools.product(*ranges):
        b = list(c)
        b[ax] = bnd
        z = float(c[2]) if dim == 3 else 0.0
        out.append((c, tuple(b), lib.orc_sdf(cfg.raw, float(c[0]), float(c[1]), z) <= 0))
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_velocity_inlet_invariants_on_random_ghosts(dim):
    """JetBoundary (Solver.cu:578-633) on arbitrary (non-equilibrium) ghost
    populations, inlet disc smaller than the plane.  Closed forms that follow
    from the update rules, not from restating them:
    inside the disc, with G = Gamma(U), geq = phi_ghost (G + w) and
    M_t = sum_{q>=1, e_n=0} e_t (f_q - G_q) before the update,
      * tangential and outgoing populations are untouched;
      * sum_{q>=1} (g_q - geq_q) = 0 (the -N/weight terms cancel N exactly:
        NIN - 1 = weight);
      * the normal non-equilibrium momentum sum_q e_n (f_q - G_q) = 0;
      * the tangential one sum_q e_t (f_q - G_q) = (1 - 2/weight) M_t:
        0 in 2D (weight 2), M_t / 2 in 3D (weight 4), since two incoming
        directions carry e_t = +-1;
    outside the disc the incoming populations are the boundary node's
    outgoing ones (bounce-back copy) and phi_ghost = phi_boundary."""
    side = 3  # y+
    if dim == 2:
        cfg = Config.from_string(cases.jet2d(12, 10, radius=3.0))
    else:
        cfg = Config.from_string(cases.jet3d(n=(10, 12, 10), radius=3.0, inlet_side="y+"))
    o = Oracle(cfg).init()
    ex = cfg.extents()
    L = lat(dim)
    e, w, Q = L["e"][:, :dim], L["w"], cfg.q
    opp = [int(np.where((e == -e[q]).all(axis=1))[0][0]) for q in range(Q)]
    ax, inward = side // 2, (-1 if side % 2 else 1)
    rng = np.random.default_rng(11)
    f0 = 0.1 * w + rng.uniform(-0.05, 0.05, size=(*ex, Q))
    g0 = rng.uniform(0.0, 0.1, size=(*ex, Q))
    phi0 = rng.uniform(0.1, 0.9, size=(*ex, 1))
    o.set("f", f0)
    o.set("g", g0)
    o.set("phi", phi0)
    o.phase("velocity")
    f1, g1, phi1 = o.get("f"), o.get("g"), o.get("phi")
    U = np.array(cfg.raw.bc_value[side][:dim])
    G = np.array([oracle_lib.gamma(dim, q, U) for q in range(Q)])
    incoming = [q for q in range(Q) if e[q, ax] == inward]
    tang = [a for a in range(dim) if a != ax]
    weight = len(incoming) - 1
    n_in = n_out = 0
    for c, b, inside in _inlet_plane_nodes(cfg, side):
        if inside:
            n_in += 1
            for q in range(Q):
                if q not in incoming:
                    assert f1[c][q] == f0[c][q] and g1[c][q] == g0[c][q]
            geq = phi0[c][0] * (G + w)
            assert abs((g1[c][1:] - geq[1:]).sum()) < 1e-15
            assert abs(((f1[c] - G) * e[:, ax]).sum()) < 1e-15
            for t in tang:
                Mt = sum(e[q, t] * (f0[c][q] - G[q]) for q in range(1, Q) if e[q, ax] == 0)
                after = ((f1[c] - G) * e[:, t]).sum()
                assert abs(after - (1.0 - 2.0 / weight) * Mt) < 1e-15, (c, t)
            assert phi1[c][0] == phi0[c][0]
        else:
            n_out += 1
            for q in incoming:
                assert f1[c][q] == f0[b][opp[q]] and g1[c][q] == g0[b][opp[q]]
            assert phi1[c][0] == phi0[b][0]
    assert n_in > 0 and n_out > 0


@pytest.mark.parametrize("text", [cases.drop2d(n=24, radius=6), cases.bubble3d(n=10, radius=3)])
def test_initial_g_moments(text):
    """KernelInitializeDistributions (NsAllen.cu:183-195):
    g = phi (Gamma(0) + w) - 1/2 w (1 - 4 (phi - 1/2)^2) / W (e . n), so by the
    lattice moments sum_q w = 1, sum_q w e = 0, sum_q w e e = I/3:
    sum_q g = phi and sum_q e g = -(1 - 4 (phi - 1/2)^2) n / (6 W)."""
    cfg = Config.from_string(text)
    o = Oracle(cfg).init()
    sl = cfg.interior_slices()
    L = lat(cfg.dim)
    e = L["e"][:, :cfg.dim].astype(float)
    g, phi, nr = o.get("g")[sl], o.get("phi")[sl][..., 0], o.get("normal")[sl]
    assert np.abs(np.linalg.norm(nr, axis=-1) - 1).max() < 1e-12  # the band matters: n is a unit vector
    assert np.allclose(g.sum(-1), phi, rtol=0, atol=1e-15)
    want = -(1 - 4 * (phi - 0.5) ** 2)[..., None] * nr / (6 * cfg.interface_width)
    got = np.einsum("...q,qa->...a", g, e)
    assert np.abs(want).max() > 1e-3
    assert np.allclose(got, want, rtol=0, atol=1e-16)


@pytest.mark.parametrize("text", [cases.drop2d(n=12, radius=3), cases.bubble3d(n=8, radius=2)])
def test_normal_is_unit_gradient_direction(text):
    """normal_FD (NsAllen.cu:241-275): on a linear phi (where the isotropic
    gradient is exact) n = grad phi / |grad phi| at every interior node, a
    unit vector; on a uniform phi n = 0 (the 1e-32 guard); ghost normals stay
    at their initial 0."""
    cfg = Config.from_string(text)
    o = Oracle(cfg).init()
    D = cfg.dim
    grad = np.array([0.013, -0.021, 0.007])[:D]
    if D == 2:
        grad = np.array([0.013, -0.021])
    _set_phi(o, lambda x, y, z: 0.3 + grad[0] * x + grad[1] * y + (grad[2] * z if D == 3 else 0.0))
    o.phase("gradient", "normals")
    nr = o.interior("normal")
    want = grad / np.linalg.norm(grad)
    for d in range(D):
        assert np.allclose(nr[..., d], want[d], rtol=0, atol=1e-14)
    assert np.allclose(np.linalg.norm(nr, axis=-1), 1.0, rtol=0, atol=1e-14)
    full = o.get("normal")
    ring = np.ones(full.shape[:3], dtype=bool)
    ring[cfg.interior_slices()] = False
    assert np.all(full[ring] == 0.0)
    _set_phi(o, lambda x, y, z: 0.25 + 0.0 * x)
    o.phase("gradient", "normals")
    assert np.all(o.interior("normal") == 0.0)


def test_mirror_copies_boundary_phi():
    """kernelMirrorBoundary (Solver.cu:927-978)."""
    cfg = Config.from_string(cases.box2d([("x-", "wall"), ("x+", "wall"), ("y-", "wall"), ("y+", "wall")], n=6, ny=5))
    o = Oracle(cfg).init()
    _set_phi(o, lambda x, y, z: 0.1 * x + 0.01 * y * y)
    o.phase("mirror")
    p = o.get("phi")[..., 0, 0]
    assert np.array_equal(p[0, 1:-1], p[1, 1:-1])
    assert np.array_equal(p[-1, 1:-1], p[-2, 1:-1])
    assert np.array_equal(p[:, 0], p[:, 1])
    assert np.array_equal(p[:, -1], p[:, -2])


# ---------------------------------------------------------------- collision / step
def test_collisions_conserve_phi_and_p_per_node():
    """kernelCollisionMRTg conserves phi (sum_q h_eq = phi) and
    kernelCollisionMRTf preserves p = sum f (row 0 of M, mass moment)."""
    cfg = Config.from_string(cases.drop2d(n=24, radius=6))
    o = Oracle(cfg).init().step(20)
    o.phase("normals", "collide_g", "collide_f")
    sl = cfg.interior_slices()
    gp, g = o.get("g_post")[sl], o.get("g")[sl]
    fp, f = o.get("f_post")[sl], o.get("f")[sl]
    assert np.allclose(g.sum(-1), gp.sum(-1), atol=1e-15)
    assert np.allclose(f.sum(-1), fp.sum(-1), atol=1e-15)


@pytest.mark.parametrize("text", [cases.drop2d(n=16, radius=40.0), cases.bubble3d(n=8, radius=30.0, g=0.0)])
def test_rest_state_is_a_fixed_point(text):
    """Uniform phase (no geometry carries id 0, so sdf = 1e10 everywhere), u = 0:
    nothing moves."""
    cfg = Config.from_string(text)
    cfg.raw.geom_id[0] = 7
    o = Oracle(cfg).init()
    f0, g0, p0 = o.interior("f").copy(), o.interior("g").copy(), o.interior("phi").copy()
    assert np.ptp(p0) == 0.0
    o.step(10)
    assert np.abs(o.interior("u")).max() < 1e-16
    assert np.allclose(o.interior("phi"), p0, atol=1e-15)
    assert np.allclose(o.interior("f"), f0, atol=1e-17)
    assert np.allclose(o.interior("g"), g0, atol=1e-16)


def test_uniform_flow_is_preserved():
    """Galilean check: uniform phase, uniform u0, equilibrium populations."""
    cfg = Config.from_string(cases.drop2d(n=16, radius=40.0))
    o = Oracle(cfg).init()
    ex = cfg.extents()
    u0 = np.array([0.02, -0.01])
    L = lat(2)
    Gm = np.array([oracle_lib.gamma(2, q, u0) for q in range(9)])
    p0 = o.get("p")[..., 0]
    phi = o.get("phi")[..., 0]
    f = p0[..., None] * L["w"] + Gm
    g = phi[..., None] * (Gm + L["w"])
    for name, arr in (("f", f), ("f_post", f), ("g", g), ("g_post", g)):
        o.set(name, arr)
    u = np.broadcast_to(u0, (ex[0], ex[1], 1, 2)).copy()
    o.set("u", u)
    o.step(5)
    uu = o.interior("u")
    assert np.allclose(uu[..., 0], u0[0], atol=1e-13)
    assert np.allclose(uu[..., 1], u0[1], atol=1e-13)


def test_drop_keeps_mirror_symmetry_and_mass():
    """A centred drop stays symmetric under x -> N+1-x, y -> N+1-y and x <-> y
    (pins direction/sign conventions of stencils, pull and forcing), and
    sum(phi) is conserved in a periodic box."""
    cfg = Config.from_string(cases.drop2d(n=32, radius=8))
    o = Oracle(cfg).init()
    s0 = o.interior("phi").sum()
    o.step(200)
    phi = o.interior("phi")[..., 0, 0]
    assert np.abs(phi - phi[::-1, :]).max() < 1e-12
    assert np.abs(phi - phi[:, ::-1]).max() < 1e-12
    assert np.abs(phi - phi.T).max() < 1e-12
    ux = o.interior("u")[..., 0, 0]
    uy = o.interior("u")[..., 0, 1]
    assert np.abs(ux + ux[::-1, :]).max() < 1e-14
    assert np.abs(ux - uy.T).max() < 1e-14
    assert abs(o.interior("phi").sum() - s0) < 1e-11 * s0


def test_init_pressure_is_laplace_jump():
    """KernelInitializeDistributions (NsAllen.cu:155-168): drop interior p =
    phi sigma / (L/2) / (rho/3); f = p w at u = 0."""
    cfg = Config.from_string(cases.drop2d(n=32, radius=8))
    o = Oracle(cfg).init()
    phi = o.get("phi")[..., 0]
    rho = cfg.rho_l + phi * (cfg.rho_h - cfg.rho_l)
    p = o.get("p")[..., 0]
    assert np.allclose(p, phi * cfg.sigma / (cfg.reference_length / 2) / (rho / 3), rtol=1e-15, atol=0)
    f = o.get("f")
    assert np.allclose(f, p[..., None] * lat(2)["w"], rtol=1e-15, atol=1e-30)


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_and_product_tables_agree(dim):
    """The oracle inverts M numerically (long double Gauss-Jordan); the
    product's tables come from exact rational inversion (gen_lattice.py).
    Both must give the same matrices."""
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        "gen_lattice", os.path.join(os.path.dirname(__file__), "..", "clip_amd", "csrc", "gen_lattice.py"))
    gl = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gl)
    if dim == 2:
        M = gl.build(gl.d2q9_rows(), gl.D2Q9_E)
    else:
        M = gl.build(gl.d3q19_rows(), gl.D3Q19_E)
    Mi = gl.inverse(M)
    L = lat(dim)
    assert np.array_equal(L["M"], np.array([[float(x) for x in r] for r in M]))
    assert np.abs(L["Minv"] - np.array([[float(x) for x in r] for r in Mi])).max() < 1e-16
    E = gl.D2Q9_E if dim == 2 else [e[:3] for e in gl.D3Q19_E]
    assert [tuple(r[:dim]) for r in L["e"]] == [tuple(e[:dim]) for e in E]
    # the kernels' wall / slip side lists (derived from the minus sides) equal
    # the reference's maps the oracle holds literally
    wall, slip = gl.side_lists(E, gl.D2Q9_IN_MINUS if dim == 2 else gl.D3Q19_IN_MINUS)
    assert np.array_equal(np.array(wall), oracle_lib.side_map(dim, False))
    assert np.array_equal(np.array(slip), oracle_lib.side_map(dim, True))
    for lists in (wall, slip):  # each side's list holds exactly its in-directions
        for s, lst in enumerate(lists):
            a, sgn = s // 2, (1 if s % 2 == 0 else -1)
            assert sorted(lst) == [q for q in range(len(E)) if E[q][a] == sgn]


# ------------------------------------------------------- viscous force
@pytest.mark.parametrize("dim,mode", [(2, "xy"), (2, "xx-yy"), (3, "xy"), (3, "yz"), (3, "xx-yy")])
def test_viscous_force_on_pure_stress_modes(dim, mode):
    """calculateVF after the WMRT filter (NsAllen.cu:102-132, 416-431): for
    gneq_q = A w_q P(e_q) with P a traceless second-order monomial, the
    filter scales gneq by s and sum_q e e gneq' follows from the lattice's
    fourth-order isotropy (sum w e_a^2 e_b^2 = 1/9, sum w e_a^4 = 1/3), so
    fv = -tau drho Pi' grad(phi) is known in closed form."""
    text = cases.drop2d(n=8, radius=2) if dim == 2 else cases.bubble3d(n=6, radius=2)
    cfg = Config.from_string(text)
    o = Oracle(cfg)
    L = lat(dim)
    e, w = L["e"], L["w"]
    A, s, tau = 0.37, 0.61, 0.23
    P = {"xy": e[:, 0] * e[:, 1], "yz": e[:, 1] * e[:, 2], "xx-yy": e[:, 0] ** 2 - e[:, 1] ** 2}[mode]
    gneq = A * w * P
    gf = oracle_lib.wmrt_filter(dim, gneq, s)
    assert np.allclose(gf, s * gneq, rtol=0, atol=1e-15)
    dc = np.array([0.3, -0.7, 0.5])
    fv = np.zeros(3)
    o.lib.orc_viscous_force(o.h, oracle_lib._dptr(gf), tau, oracle_lib._dptr(dc), oracle_lib._dptr(fv))
    Pi = np.zeros((3, 3))
    if mode == "xy":
        Pi[0, 1] = Pi[1, 0] = s * A / 9
    elif mode == "yz":
        Pi[1, 2] = Pi[2, 1] = s * A / 9
    else:
        Pi[0, 0], Pi[1, 1] = 2 * s * A / 9, -2 * s * A / 9
    dcd = dc.copy()
    if dim == 2:
        dcd[2] = 0.0
    expect = -tau * (cfg.rho_h - cfg.rho_l) * (Pi @ dcd)
    assert np.allclose(fv[:dim], expect[:dim], rtol=1e-13, atol=1e-18), (fv, expect)


# ------------------------------------------------------------ Laplace law
def test_static_drop_obeys_laplace_law():
    """The coupled step relaxes a 2D drop to the Young-Laplace jump
    P_in - P_out = sigma / R, with CLIP's pressure P = p rho / 3
    (kernelMacroscopicf, NsAllen.cu:369-511; the surface force mu grad phi,
    the pressure-gradient correction and the WMRT collision all enter)."""
    R = 10.0
    cfg = Config.from_string(cases.drop2d(n=48, radius=R, rho_ratio=10, mu_ratio=10))
    o = Oracle(cfg).init()
    o.step(3000)
    phi = o.interior("phi")[..., 0, 0]
    p = o.interior("p")[..., 0, 0]
    rho = cfg.rho_l + phi * (cfg.rho_h - cfg.rho_l)
    P = p * rho / 3.0
    inside = P[phi > 0.99].mean()
    outside = P[phi < 0.01].mean()
    jump = inside - outside
    assert abs(jump - cfg.sigma / R) < 0.08 * cfg.sigma / R, (jump, cfg.sigma / R)
    # the drop stays put and round: spurious currents are small
    assert np.abs(o.interior("u")).max() < 1e-2 * np.sqrt(cfg.sigma / R)
