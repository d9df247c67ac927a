"""Committed oracle fixtures (tests/golden/fixtures, written by
tests/golden/make_fixtures.py): the oracle must reproduce them bit for bit,
and the FP64 HIP path must match them to the FP64 parity bar."""
import os

import numpy as np
import pytest

from clip_amd import Config
from tests.oracle_lib import Oracle

FIX = os.path.join(os.path.dirname(__file__), "golden", "fixtures")
NAMES = sorted(f[:-4] for f in os.listdir(FIX) if f.endswith(".npz"))


def load(name):
    z = np.load(os.path.join(FIX, name + ".npz"))  # allow_pickle stays False
    return bytes(z["config"]).decode(), [int(s) for s in z["steps"]], z


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    text, steps, z = load(name)
    o = Oracle(Config.from_string(text)).init()
    done = 0
    for s in steps:
        o.step(s - done)
        done = s
        for f in ("phi", "u", "p"):
            assert np.array_equal(o.interior(f), z[f"{f}_{s}"]), (name, s, f)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_f64_matches_fixture(name):
    from clip_amd import Solver
    text, steps, z = load(name)
    print("fixture", name, steps, flush=True)
    s_ = Solver(Config.from_string(text), "f64").init_fresh()
    done = 0
    for s in steps:
        s_.step(s - done)
        done = s
        for f in ("phi", "u", "p"):
            a, b = s_.interior(f), z[f"{f}_{s}"]
            den = np.linalg.norm(b) or 1.0
            assert np.linalg.norm(a - b) / den <= 1e-11, (name, s, f)
