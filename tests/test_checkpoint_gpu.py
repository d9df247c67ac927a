"""Checkpoint / resume in CLIP's on-disk layout (CheckPointer.cu:53-282)."""
import os
import struct

import numpy as np
import pytest

from clip_amd import ClipError, Config, Solver
from clip_amd import _abi
from tests import cases
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def _files(d):
    return sorted(os.listdir(d))


@pytest.mark.parametrize("text,precision", [
    (cases.drop2d(n=24, radius=6), "f64"),
    (cases.jet2d(20, 24, radius=4), "f64"),
    (cases.bubble3d(n=10, radius=3), "f32"),
])
def test_resume_is_bit_exact(tmp_path, text, precision):
    cfg = Config.from_string(text)
    a = Solver(cfg, precision).init_fresh().step(30)
    d = tmp_path / "Checkpoint"
    a.checkpoint_save(d)
    a.step(40)
    ref = {f: a.get(f) for f in ("phi", "u", "p", "f", "g", "f_post", "g_post")}
    b = Solver(cfg, precision)
    b.checkpoint_load(d)
    assert b.info().step == 30
    b.step(40)
    for f, v in ref.items():
        got = b.get(f)
        sl = cfg.interior_slices()
        assert np.array_equal(got[sl], v[sl]), f


@pytest.mark.parametrize("precision", ["mixed_f16", "mixed_bf16"])
def test_mixed_resume_is_bit_exact(tmp_path, precision):
    """16-bit storage: clip_amd_state.bin also carries the stored bytes of the
    pulled f and g, so the restart starts from the very same bits.  Without
    them the restart would re-encode decoded values, which is not always
    idempotent (the decode s/2^10 + w phi rounds in FP32): on this drop a
    one-ulp phi difference at 6 nodes grew to rel-L2 1.7e-5 in 40 steps."""
    cfg = Config.from_string(cases.drop2d(n=24, radius=6))
    a = Solver(cfg, precision).init_fresh().step(30)
    d = tmp_path / "Checkpoint"
    a.checkpoint_save(d)
    a.step(40)
    b = Solver(cfg, precision)
    b.checkpoint_load(d)
    b.step(40)
    for f in ("phi", "u", "f", "g"):
        x, y = b.get(f), a.get(f)
        print(precision, f, "max |diff| after resume", float(np.max(np.abs(x - y))), flush=True)
        assert np.array_equal(x, y), f


def test_mixed_state_block_checks(tmp_path):
    """The raw 16-bit block of clip_amd_state.bin: a truncated block is an I/O
    error; a block written by the other 16-bit type is not used (the decoded
    files are re-encoded), so the resume is close to the run but need not be
    exact."""
    cfg = Config.from_string(cases.drop2d(n=24, radius=6))
    a = Solver(cfg, "mixed_f16").init_fresh().step(30)
    d = tmp_path / "Checkpoint"
    a.checkpoint_save(d)
    # another 16-bit type: loads, and stays close to the FP16 run
    b = Solver(cfg, "mixed_bf16")
    b.checkpoint_load(d)
    a.step(20)
    b.step(20)
    x, y = b.interior("phi"), a.interior("phi")
    err = np.linalg.norm(x - y) / np.linalg.norm(y)
    print("bf16 from an fp16 checkpoint: phi rel-L2", err, flush=True)
    assert err <= 2e-2
    # truncated raw block
    sp = d / "clip_amd_state.bin"
    data = sp.read_bytes()
    sp.write_bytes(data[:-100])
    c = Solver(cfg, "mixed_f16")
    with pytest.raises(ClipError):
        c.checkpoint_load(d)


def test_file_set_and_layout(tmp_path):
    cfg = Config.from_string(cases.jet2d(20, 24, radius=4))
    s = Solver(cfg, "f64").init_fresh().step(5)
    d = tmp_path / "Checkpoint"
    s.checkpoint_save(d)
    names = _files(d)
    for n in ("dev_f.bin", "dev_g.bin", "dev_f_post.bin", "dev_g_post.bin", "dev_f_prev.bin", "dev_g_prev.bin",
              "timeInfo.bin", "domainSize.bin", "info.txt", "clip_amd_state.bin"):
        assert n in names
    ex = cfg.extents()
    nodes = ex[0] * ex[1] * ex[2]
    assert os.path.getsize(d / "dev_f.bin") == nodes * 9 * 8
    assert os.path.getsize(d / "timeInfo.bin") == 32
    t = open(d / "timeInfo.bin", "rb").read()
    assert struct.unpack_from("<i", t, 8)[0] == 5
    N = np.fromfile(d / "domainSize.bin", dtype=np.float64)
    assert list(N) == list(cfg.N)
    # AoS order ((i*ey + j)*ez + k)*Q + q, ghosts included
    f = np.fromfile(d / "dev_f_post.bin", dtype=np.float64).reshape(ex[0], ex[1], ex[2], 9)
    assert np.array_equal(f[cfg.interior_slices()], s.interior("f_post"))


def test_resume_from_reference_file_set_matches_oracle_restart(tmp_path):
    """Without the extension file the resume is the reference's
    (load f, g, f_post, g_post, then macroscopic() with u_old = 0)."""
    cfg = Config.from_string(cases.drop2d(n=20, radius=5))
    a = Solver(cfg, "f64").init_fresh().step(20)
    d = tmp_path / "ck"
    a.checkpoint_save(d)
    os.remove(d / "clip_amd_state.bin")
    b = Solver(cfg, "f64")
    b.checkpoint_load(d)
    o = Oracle(cfg).init()
    for f in ("f", "g", "f_post", "g_post"):
        o.set(f, a.get(f))
    o.set("u", np.zeros_like(a.get("u")))
    # phi from g_post, then the ghost ring refreshed from it (the resumed
    # state has no ghost ring of its own), then the rest of macroscopic()
    o.phase("macro_g", "periodic", "chem_pot", "gradient", "macro_f")
    for f in ("phi", "u", "p"):
        assert np.allclose(b.interior(f), o.interior(f), rtol=1e-12, atol=1e-16), f


def test_domain_mismatch_and_truncation_are_errors(tmp_path):
    cfg = Config.from_string(cases.drop2d(n=20, radius=5))
    a = Solver(cfg, "f64").init_fresh().step(2)
    d = tmp_path / "ck"
    a.checkpoint_save(d)
    other = Config.from_string(cases.drop2d(n=22, radius=5))
    with pytest.raises(ClipError) as e:
        Solver(other, "f64").checkpoint_load(d)
    assert "mismatch" in str(e.value)
    raw = open(d / "dev_g.bin", "rb").read()
    open(d / "dev_g.bin", "wb").write(raw[: len(raw) // 2])
    with pytest.raises(ClipError) as e:
        Solver(cfg, "f64").checkpoint_load(d)
    assert e.value.code == _abi.ERR_IO


def test_rotation(tmp_path):
    lib = _abi.load()
    base = tmp_path / "Checkpoint"
    for step in range(4):
        assert lib.clip_checkpoint_rotate(str(base).encode(), 2) == 0
        os.makedirs(base)
        open(base / "tag", "w").write(str(step))
    assert open(base / "tag").read() == "3"
    assert open(tmp_path / "Checkpoint_1" / "tag").read() == "2"
    assert open(tmp_path / "Checkpoint_2" / "tag").read() == "1"
    assert not os.path.exists(tmp_path / "Checkpoint_3")


def test_health_check_reports_divergence():
    cfg = Config.from_string(cases.drop2d(n=16, radius=4))
    s = Solver(cfg, "f64").init_fresh().step(2)
    s.check_health()
    phi = s.get("phi")
    phi[5, 5, 0, 0] = np.nan
    s.set("phi", phi)
    with pytest.raises(ClipError) as e:
        s.check_health()
    assert e.value.code == _abi.ERR_DIVERGED


def test_driver_runs_config(tmp_path):
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "..", "clip_amd", "clip")
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    (tmp_path / "config.txt").write_text(cases.drop2d(n=24, radius=6)
                                         + "checkpointInterval = 10\ncheckpointCopy = 2\noutputInterval = 10\n")
    r = subprocess.run([exe, "-steps", "20"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "MLUPS" in r.stdout
    assert os.path.exists(tmp_path / "Checkpoint" / "dev_f.bin")
    assert os.path.exists(tmp_path / "Checkpoint_1" / "dev_f.bin")
    for t in (0, 10, 20):
        assert os.path.exists(tmp_path / "results" / f"results_t{t}.vts")
    r = subprocess.run([exe, "-resume", "-steps", "5"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_vts_output_holds_the_fields(tmp_path):
    """VTSwriter layout (VTSwriter.cu:84-205): StructuredGrid over the interior
    extent, Velocity (DIM components), C, P, Rho, x fastest."""
    import xml.etree.ElementTree as ET
    for text in (cases.drop2d(n=12, radius=3), cases.bubble3d(n=8, radius=2.5)):
        cfg = Config.from_string(text)
        s = Solver(cfg, "f64").init_fresh().step(7)
        path = tmp_path / "out.vts"
        s.write_vts(path)
        root = ET.parse(path).getroot()
        grid = root.find("StructuredGrid")
        ex = cfg.extents()
        hi = [e - 2 for e in ex[: cfg.dim]] + [0] * (3 - cfg.dim)
        lo = [1] * cfg.dim + [0] * (3 - cfg.dim)
        assert grid.get("WholeExtent").split() == [str(v) for a in range(3) for v in (lo[a], hi[a])]
        arrays = {d.get("Name"): np.array(d.text.split(), dtype=float) for d in grid.iter("DataArray") if d.get("Name")}
        sl = cfg.interior_slices()

        def flat(a):  # grid (i, j, k, dof) -> k, j, i order, dof fastest
            return np.transpose(a[sl], (2, 1, 0, 3)).reshape(-1, a.shape[-1])

        assert np.allclose(arrays["C"], flat(s.get("phi"))[:, 0], rtol=1e-5, atol=1e-12)
        assert np.allclose(arrays["P"], flat(s.get("p"))[:, 0], rtol=1e-5, atol=1e-12)
        assert np.allclose(arrays["Rho"], flat(s.get("rho"))[:, 0], rtol=1e-5, atol=1e-12)
        assert np.allclose(arrays["Velocity"].reshape(-1, cfg.dim), flat(s.get("u")), rtol=1e-5, atol=1e-12)
