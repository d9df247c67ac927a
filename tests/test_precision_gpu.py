"""Reduced-precision paths against the FP64 path (itself bit-identical to the
FP64 oracle on the parity cases) at full workload sizes:
  - 2D Rayleigh-Taylor, 256 x 1024, 20,000 steps: FP32 vs FP64;
  - 3D rising bubble, 256^3, 5,000 steps: FP16 / BF16 deviation storage vs FP64.
Reported: rel-L2 and max-abs of phi, u, p and the drift of sum(phi)."""
import numpy as np
import pytest

from clip_amd import Config, Solver
from tests import cases

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def compare(a, b, cfg):
    out = {}
    for f in ("phi", "u", "p"):
        x, y = a.interior(f), b.interior(f)
        assert np.all(np.isfinite(x)), f
        out[f] = (np.linalg.norm(x - y) / np.linalg.norm(y), np.abs(x - y).max())
    out["mass"] = abs(a.sum("phi") - b.sum("phi")) / abs(b.sum("phi"))
    return out


def report(name, e):
    print(name, " ".join(f"{k} rel-L2 {v[0]:.2e} max {v[1]:.2e}" for k, v in e.items() if k != "mass"),
          f"mass {e['mass']:.2e}", flush=True)


def test_rti2d_fp32_20000_steps():
    cfg = Config.from_string(cases.rti2d(256, 1024))
    print("RTI 256x1024, 20000 steps, FP32 vs FP64", flush=True)
    ref = Solver(cfg, "f64").init_fresh().step(20000)
    low = Solver(cfg, "f32").init_fresh().step(20000)
    e = compare(low, ref, cfg)
    report("rti2d fp32", e)
    t = TOL["rti2d_fp32"]
    for f in ("phi", "u", "p", "mass"):
        v = e[f] if f == "mass" else e[f][0]
        assert v <= t[f], (f, v, t[f])


@pytest.mark.parametrize("precision", ["mixed_f16", "mixed_bf16"])
def test_bubble3d256_mixed_5000_steps(precision):
    cfg = Config.from_string(cases.bubble3d(n=256, radius=32.0))
    print("bubble 256^3, 5000 steps,", precision, "vs FP64", flush=True)
    ref = Solver(cfg, "f64").init_fresh().step(5000)
    low = Solver(cfg, precision).init_fresh().step(5000)
    e = compare(low, ref, cfg)
    report("bubble3d256 " + precision, e)
    t = TOL["bubble3d256_" + precision]
    for f in ("phi", "u", "p", "mass"):
        v = e[f] if f == "mass" else e[f][0]
        assert v <= t[f], (f, v, t[f])


# Measured on one MI355X: RTI FP32 phi 1.5e-2 (the late, chaotic roll-up
# shifts the interface by a fraction of a node), u 2.6e-3, p 5.7e-3, mass
# 2.8e-4; bubble FP16 phi 4.3e-4, u 3.1e-3, p 5.8e-3, mass 1.6e-4; BF16 phi
# 6.7e-3, u 0.14, p 0.46, mass 1.4e-4 (BF16 keeps 8 significant bits: its
# pressure is mostly rounding noise at this horizon).
TOL = {
    "rti2d_fp32": {"phi": 5e-2, "u": 1e-2, "p": 2e-2, "mass": 1e-3},
    "bubble3d256_mixed_f16": {"phi": 2e-3, "u": 1e-2, "p": 2e-2, "mass": 1e-3},
    "bubble3d256_mixed_bf16": {"phi": 3e-2, "u": 0.3, "p": 0.8, "mass": 1e-3},
}
