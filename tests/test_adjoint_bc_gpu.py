"""Reverse mode on bounded domains (velocity inlet with azimuthal modes,
free-convect outflow, Neumann, walls, specular slip sides, do-nothing) checked against central
finite differences of the FP64 GPU forward: dL/dsigma, dL/dmobility,
dL/dA_m of the inlet modes and dL/dphi0 at interface-band nodes next to
the boundaries, for L = 1/2 |phi_T - phi_target|^2 with the target from a
run with other parameters."""
import numpy as np
import pytest

from clip_amd import Config, Solver
from tests import cases

pytestmark = pytest.mark.gpu

CASES = [
    ("jet2d", cases.jet2d(20, 24, radius=4), 40),
    ("box2d_mixed", cases.box2d([("x-", "wall"), ("x+", "do nothing"), ("y-", "free convect"), ("y+", "neumann")],
                                n=18, ny=20), 40),
    ("rti2d_walls", cases.rti2d(24, 40), 40),
    ("jet3d_c5_20x20x32", cases.jet3d_c5(nx=20, ny=20, nz=32), 30),
    ("jet3d_neumann", cases.jet3d(n=(14, 16, 14), radius=3.0, inlet_side="z-", outlet_type="neumann"), 30),
    ("box2d_specular_slip", cases.specular(cases.box2d([("x-", "slip wall"), ("x+", "slip wall"),
                                                        ("y-", "wall"), ("y+", "slip wall")], n=18, ny=20)), 40),
    ("box2d_specular_slip_only", cases.specular(cases.box2d([("x-", "periodic"), ("x+", "periodic"),
                                                             ("y-", "slip wall"), ("y+", "slip wall")],
                                                            n=18, ny=20)), 40),
]


def configured(text, sigma=None, mobility=None, amps=None):
    cfg = Config.from_string(text)
    if amps is not None:
        for m, a in enumerate(amps):
            cfg.raw.inlet_mode_amp[m] = float(a)
    return cfg.with_params(sigma=sigma, mobility=mobility)


def initial_phi(cfg):
    s = Solver(cfg, "f64").init_fresh()
    a = s.interior("phi")[..., 0]
    s.close()
    return a[..., 0] if cfg.dim == 2 else a


def forward(cfg, phi0, T, K=None, inner=0):
    s = Solver(cfg, "f64")
    phiT = s.adjoint_forward(phi0, T, K or T, inner_every=inner)
    return s, phiT


def loss(cfg, phi0, target, T):
    s, phiT = forward(cfg, phi0, T)
    s.close()
    return 0.5 * float(np.sum((phiT - target) ** 2))


@pytest.mark.parametrize("name,text,T", CASES, ids=[c[0] for c in CASES])
def test_bounded_adjoint_against_central_differences(name, text, T):
    base = Config.from_string(text)
    amps0 = np.array([base.raw.inlet_mode_amp[m] for m in range(base.raw.inlet_modes)])
    cfg = configured(text, sigma=base.sigma, mobility=base.mobility, amps=amps0)
    phi0 = initial_phi(cfg)
    tcfg = configured(text, sigma=1.2 * base.sigma, mobility=0.8 * base.mobility,
                      amps=1.5 * amps0 if len(amps0) else None)
    _, target = forward(tcfg, phi0, T)
    s, phiT = forward(cfg, phi0, T, K=10)
    g0, gs, gm = s.adjoint_backward(phiT - target)
    ga = s.adjoint_inlet_gradient()
    print(f"{name}: T={T} loss={0.5 * np.sum((phiT - target) ** 2):.3e} dL/dsigma={gs:.6e} dL/dM={gm:.6e}",
          flush=True)

    def fd(make, h):
        return (loss(make(+h), phi0, target, T) - loss(make(-h), phi0, target, T)) / (2 * h)

    checks = []
    hs = 1e-6 * base.sigma
    checks.append(("sigma", gs, fd(lambda d: configured(text, sigma=base.sigma + d, mobility=base.mobility,
                                                        amps=amps0), hs)))
    hm = 1e-6 * base.mobility
    checks.append(("mobility", gm, fd(lambda d: configured(text, sigma=base.sigma, mobility=base.mobility + d,
                                                           amps=amps0), hm)))
    for m in range(min(3, len(amps0))):
        def mk(d, m=m):
            a = amps0.copy()
            a[m] += d
            return configured(text, sigma=base.sigma, mobility=base.mobility, amps=a)
        checks.append((f"A_{m + 1}", ga[m], fd(mk, 1e-6)))
    # phi0 entries in the interface band, nearest the non-periodic boundaries first
    band = np.argwhere((phi0 > 0.1) & (phi0 < 0.9))
    shape = np.array(phi0.shape)
    edge = np.minimum(band, shape - 1 - band).min(axis=1)
    picks = band[np.argsort(edge, kind="stable")[:4]].tolist()
    rng = np.random.default_rng(2)
    picks += band[rng.choice(len(band), 2, replace=False)].tolist()
    for p in picks:
        def lf(d, p=tuple(p)):
            x = phi0.copy()
            x[p] += d
            return loss(cfg, x, target, T)
        checks.append((f"phi0{tuple(p)}", g0[tuple(p)], (lf(1e-6) - lf(-1e-6)) / 2e-6))
    worst = 0.0
    for what, adj, num in checks:
        err = abs(adj - num) / max(abs(num), 1e-300)
        print(f"  {what}: adjoint {adj:.9e} central difference {num:.9e} rel {err:.2e}", flush=True)
        if abs(num) > 1e-12:
            worst = max(worst, err)
            assert err < 2e-5, (what, adj, num)
    print(f"  worst relative difference {worst:.2e}", flush=True)


def test_nested_checkpointing_gives_identical_gradients():
    """Two checkpoint levels recompute the same states with the same kernels,
    so the gradients equal those of one level bit for bit."""
    text = cases.jet3d_c5(nx=20, ny=20, nz=32)
    cfg = Config.from_string(text)
    phi0 = initial_phi(cfg)
    T = 37
    out = []
    for K, inner in ((T, 0), (12, 4), (12, 5), (10, 0)):
        s, phiT = forward(cfg, phi0, T, K=K, inner=inner)
        g0, gs, gm = s.adjoint_backward(phiT - 0.9 * phiT)
        out.append((g0, gs, gm, s.adjoint_inlet_gradient()))
        s.close()
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0])
        assert o[1] == out[0][1] and o[2] == out[0][2]
        assert np.array_equal(o[3], out[0][3])
