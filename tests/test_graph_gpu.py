"""Stepping through the captured two-step HIP graph (clip_enable_graph)
gives the same bits as plain stepping, on every replay: chunks of even and
odd step counts, with eager work between the replays (field exports, an
upload + macroscopic, synchronisation)."""
import numpy as np
import pytest

from clip_amd import Config, Solver
from tests import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,text,precision", [
    ("drop2d", cases.drop2d(n=32, radius=8), "f64"),
    ("jet2d", cases.jet2d(20, 24, radius=4), "f64"),
    ("rti2d", cases.rti2d(24, 40), "f32"),
    ("bubble3d", cases.bubble3d(n=12, radius=3), "mixed_f16"),
    ("jet3d", cases.jet3d(n=(12, 14, 12), radius=3.0, inlet_side="y+"), "mixed_bf16"),
], ids=["drop2d", "jet2d", "rti2d", "bubble3d", "jet3d"])
def test_graph_steps_equal_plain_steps(name, text, precision):
    cfg = Config.from_string(text)
    a = Solver(cfg, precision).init_fresh()
    b = Solver(cfg, precision).init_fresh().enable_graph(True)
    print("case", name, precision, flush=True)
    for chunk in (2, 5, 1, 8, 3):
        a.step(chunk)
        b.step(chunk)
        b.synchronize()
        for f in ("phi", "u", "p", "f", "g"):
            assert np.array_equal(a.get(f), b.get(f)), (chunk, f)
        assert a.info().step == b.info().step
    # an upload and macroscopic() between replays (the direct state)
    fp = a.get("f_post")
    for s in (a, b):
        s.set("f_post", fp * 1.00001)
        s.macroscopic()
    a.step(6)
    b.step(6)
    for f in ("phi", "u", "p", "f", "g"):
        assert np.array_equal(a.get(f), b.get(f)), ("after upload", f)
