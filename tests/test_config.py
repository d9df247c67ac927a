"""config.txt surface (InputData / Boundary / Geometry parsers) against the
reference's own example configurations (tests/golden/examples, copied data
files) and the closed-form parameter derivation of InputData.cu:76-114."""
import math
import os

import pytest

from clip_amd import ClipError, Config
from clip_amd import _abi

EX = os.path.join(os.path.dirname(__file__), "golden", "examples")


def ex(path):
    return os.path.join(EX, path, "config.txt")


def close(a, b, rel=1e-14):
    return abs(a - b) <= rel * max(abs(a), abs(b), 1e-300)


def test_2d_drop_example():
    c = Config.from_file(ex("2D/drop"))
    assert c.dim == 2 and c.case_type == _abi.CASE["drop"]
    assert list(c.N) == [64, 256, 1] and c.final_step == 100000 and c.report_interval == 1000
    g, L, We, Re = 1e-6, 20.0, 10.0, 100.0
    rhoH, rhoL = 1.0, 1.0 / 100
    assert close(c.sigma, rhoH * g * L * L / We)
    muH = math.sqrt(g * rhoH * (rhoH - rhoL) * L ** 3) / Re
    assert close(c.mu_h, muH)
    assert close(c.mu_l, muH / 10)
    assert close(c.tau_h, 3 * muH / rhoH)
    assert close(c.tau_l, 3 * (muH / 10) / rhoL)
    assert close(c.kappa, 1.5 * c.sigma * 4)
    assert close(c.beta, 8 * c.sigma / 4)
    assert [c.bc_type[s] for s in range(4)] == [3, 3, 0, 0]
    assert c.num_geometries == 1 and c.geom_type[0] == 0 and c.geom_radius[0] == 10
    assert list(c.geom_center[0]) == [32, 192, 0]


def test_2d_bubble_example():
    c = Config.from_file(ex("2D/bubble"))
    assert c.case_type == _abi.CASE["bubble"] and c.mobility == 0.05


def test_2d_rti_example_derives_mobility():
    c = Config.from_file(ex("2D/RTI"))
    g, L, Re, Pe, Ca = 1e-6, 64.0, 3000.0, 1000.0, 0.26
    U = math.sqrt(g * L)
    assert close(c.mu_h, U * L / Re)
    assert close(c.sigma, (U * L / Re) * U / Ca)
    assert close(c.mobility, U * L / Pe)
    assert c.geom_type[0] == _abi.GEOM["perturbation"] and c.geom_amplitude[0] == 0.1
    assert list(c.geom_length[0]) == [128, 0, 0]
    # `amplitude` is not a top-level key of this file: the first line whose key
    # matches is the geometry block's (InputData.cu:294-317 scans every line)
    assert c.amplitude == 0.1


def test_3d_jet_example():
    c = Config.from_file(ex("3D/Jet"))
    assert c.dim == 3 and c.case_type == _abi.CASE["jet"]
    U, L, We, Re = 0.01, 30.0, 20000.0, 25000.0
    assert close(c.sigma, U * U * L / We)
    assert close(c.mu_h, U * L / Re)
    assert c.bc_type[3] == _abi.BC["velocity"] and list(c.bc_value[3]) == [0, -0.01, 0]
    # "free convect": the reference's whitespace stripping makes it Unknown
    # (Boundary.cu:282-288); accepted here.
    assert c.bc_type[2] == _abi.BC["free convect"]
    assert c.geom_type[0] == _abi.GEOM["sphere"]


def test_3d_rti_example():
    c = Config.from_file(ex("3D/RTI"))
    assert c.dim == 3 and list(c.N) == [64, 256, 64]


@pytest.mark.parametrize("name", ["3D/bubble", "3D/drop"])
def test_3d_examples_missing_z_sides_are_rejected(name):
    """These reference examples define only x and y sides; the reference
    throws 'Missing boundary conditions for sides: z-, z+' (Boundary.cu:238-276)."""
    with pytest.raises(ClipError) as e:
        Config.from_file(ex(name), dim=3)
    assert "z-, z+" in str(e.value)


def test_single_line_blocks_and_comments():
    text = """case = drop   # comment after value
N = [16, 16, 1]
referenceLength = 8
We = 5
Re = 100
gravity = [0, -1e-6, 0]
mobility = 0.02
interfaceWidth = 4
muRatio = 10
rhoRatio = 10
geometry = [ { type = "circle", center = [8, 8, 0], radius = 4, id = 0 } ]
boundary = [
  { side = "x-", type = "periodic" },
  { side = "x+", type = "periodic" },
  { side = "y-", type = "slip wall" },
  { side = "y+", type = "do nothing" }
]
"""
    c = Config.from_string(text)
    assert c.case_type == _abi.CASE["drop"]
    assert [c.bc_type[s] for s in range(4)] == [3, 3, 1, 6]
    assert c.num_geometries == 1 and c.geom_radius[0] == 4 and list(c.geom_center[0]) == [8, 8, 0]


@pytest.mark.parametrize("mutation,needle", [
    (lambda t: t.replace("case = drop", "case = foam"), "Unknown case type"),
    (lambda t: t.replace("We = 5\n", ""), "We"),
    (lambda t: t.replace('type = "periodic"', 'type = "sticky"', 1), "unknown type"),
    (lambda t: t.replace('side = "x+"', 'side = "x-"'), "Missing boundary"),
    (lambda t: t.replace('type = "circle"', 'type = "blob"'), "unknown type"),
])
def test_bad_configs_are_errors(mutation, needle):
    from tests import cases
    base = cases.drop2d(n=16, radius=4, We=5)
    with pytest.raises(ClipError) as e:
        Config.from_string(mutation(base))
    assert e.value.code == _abi.ERR_PARSE
    assert needle in str(e.value)


def test_checkpoint_interval_zero_with_copies_is_rejected():
    """CheckPointer::save computes step % checkpointInterval (CheckPointer.cu:59)."""
    from tests import cases
    t = cases.drop2d(n=16, radius=4) + "checkpointCopy = 2\ncheckpointInterval = 0\n"
    with pytest.raises(ClipError):
        Config.from_string(t)


def test_missing_file_is_io_error():
    with pytest.raises(ClipError) as e:
        Config.from_file("/nonexistent/config.txt")
    assert e.value.code == _abi.ERR_IO


def test_overrides():
    from tests import cases
    t = cases.drop2d(n=16, radius=4) + "sigma = 0.001\nmuH = 0.01\n"
    c = Config.from_string(t)
    assert c.sigma == 0.001 and c.mu_h == 0.01 and close(c.tau_h, 0.03)
