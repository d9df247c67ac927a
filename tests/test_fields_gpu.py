"""Field uploads (clip_set_field), NSAllen::macroscopic (clip_macroscopic)
and checkpoints of the direct state, through the C-ABI against the FP64
oracle doing the same thing to the reference's arrays
(CheckPointer::load + NSAllen::macroscopic, main.cu:67-71)."""
import numpy as np
import pytest

from clip_amd import ClipError, Config, Solver
from tests import cases
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

CASES = [
    ("drop2d", cases.drop2d(n=24, radius=6)),
    ("jet2d", cases.jet2d(20, 24, radius=4)),
    ("rti2d_walls", cases.rti2d(24, 40)),
    ("jet3d", cases.jet3d(n=(12, 14, 12), radius=3.0, inlet_side="y+")),
]


def rel_l2(a, b):
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / den if den > 0 else np.linalg.norm(a.ravel())


def check(gpu, orc, fields=("phi", "u", "p"), tol=1e-11):
    errs = {f: rel_l2(gpu.interior(f), orc.interior(f)) for f in fields}
    print({k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for f, e in errs.items():
        assert e <= tol, (f, e)


@pytest.mark.parametrize("name,text", CASES, ids=[c[0] for c in CASES])
def test_set_populations_then_macroscopic_matches_oracle(name, text):
    """Perturbed f, g, f_post, g_post (ghost cells included) uploaded into
    both, then macroscopic(), then 15 steps."""
    cfg = Config.from_string(text)
    print("case", name, flush=True)
    gpu = Solver(cfg, "f64").init_fresh().step(12)
    orc = Oracle(cfg).init().step(12)
    rng = np.random.default_rng(3)
    for field in ("f", "g", "f_post", "g_post"):
        base = orc.get(field)
        new = base * (1 + 1e-3 * rng.standard_normal(base.shape))
        orc.set(field, new)
        gpu.set(field, new)
    for field in ("f_post", "g_post", "f", "g"):
        a, b = gpu.get(field)[cfg.interior_slices()], orc.get(field)[cfg.interior_slices()]
        assert np.array_equal(a, b), field
    orc.macroscopic()
    gpu.macroscopic()
    check(gpu, orc)
    orc.step(15)
    gpu.step(15)
    check(gpu, orc)


def test_macroscopic_twice_updates_u_like_the_reference():
    """kernelMacroscopicf builds gneq from u_old = the current u, so a second
    macroscopic() changes u (NsAllen.cu:400-412)."""
    cfg = Config.from_string(cases.drop2d(n=24, radius=6))
    gpu = Solver(cfg, "f64").init_fresh().step(20)
    orc = Oracle(cfg).init().step(20)
    for _ in range(2):
        gpu.macroscopic()
        orc.macroscopic()
        check(gpu, orc)
    gpu.step(10)
    orc.step(10)
    check(gpu, orc)


@pytest.mark.parametrize("field", ["p", "u", "phi"])
def test_set_macroscopic_field_feeds_the_next_collision(field):
    """An uploaded p / u / phi is what the next collision uses, and get()
    returns it until the next step (no silent recomputation)."""
    cfg = Config.from_string(cases.jet2d(20, 24, radius=4))
    gpu = Solver(cfg, "f64").init_fresh().step(10)
    orc = Oracle(cfg).init().step(10)
    v = orc.get(field)
    v = v + 1e-4 * np.random.default_rng(4).standard_normal(v.shape) * (np.abs(v).max() + 1e-3)
    if field == "phi":
        v = np.clip(v, 0.0, 1.0)
    orc.set(field, v)
    gpu.set(field, v)
    assert np.array_equal(gpu.get(field), v)
    if field == "phi":
        # the next collision's stencil inputs and density from the new phi
        orc.phase("chem_pot", "gradient")
        orc.set("rho", cfg.rho_l + v * (cfg.rho_h - cfg.rho_l))
    orc.step(8)
    gpu.step(8)
    check(gpu, orc)


def test_set_rejects_wrong_size_and_derived_fields():
    cfg = Config.from_string(cases.drop2d(n=16, radius=4))
    gpu = Solver(cfg, "f64").init_fresh()
    with pytest.raises(ValueError):
        gpu.set("f", np.zeros(10))
    with pytest.raises(ValueError):
        gpu.set("rho", gpu.get("rho"))


@pytest.mark.parametrize("precision", ["f64", "f32", "mixed_f16", "mixed_bf16"])
@pytest.mark.parametrize("name,text", [("drop2d", cases.drop2d(n=24, radius=6)),
                                       ("jet2d", cases.jet2d(20, 24, radius=4))], ids=["drop2d", "jet2d"])
def test_checkpoint_right_after_init_resumes_exactly(tmp_path, precision, name, text):
    """A checkpoint written right after initialisation (f_post held directly)
    resumes to the very same bits as the run that continued."""
    cfg = Config.from_string(text)
    print("case", name, precision, flush=True)
    a = Solver(cfg, precision).init_fresh()
    d = tmp_path / "Checkpoint"
    a.checkpoint_save(d)
    a.step(20)
    b = Solver(cfg, precision)
    b.checkpoint_load(d)
    assert b.info().step == 0
    b.step(20)
    for f in ("phi", "u", "p", "f", "g"):
        assert np.array_equal(b.get(f), a.get(f)), f


def test_checkpoint_after_upload_and_macroscopic_resumes_exactly(tmp_path):
    cfg = Config.from_string(cases.jet2d(20, 24, radius=4))
    a = Solver(cfg, "f64").init_fresh().step(7)
    fp = a.get("f_post")
    a.set("f_post", fp * 1.0001)
    a.macroscopic()
    d = tmp_path / "Checkpoint"
    a.checkpoint_save(d)
    a.step(12)
    b = Solver(cfg, "f64")
    b.checkpoint_load(d)
    b.step(12)
    for f in ("phi", "u", "p", "f", "g"):
        assert np.array_equal(b.get(f), a.get(f)), f
