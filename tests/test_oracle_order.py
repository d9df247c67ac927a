"""What the FP64 bit-parity does and does not establish.

The FP64 kernels share one canonical operation order with the oracle (the
WMRT filter as identity plus its stress part, no FMA contraction), so the
two agree bit for bit.  The reference instead applies the dense transform
(m = M gneq, stress rows scaled, gneq = M^-1 m, WMRT.cuh:54-184) and nvcc
contracts FMAs by default.  These checks build the oracle in the literal
order (-DCLIP_ORACLE_LITERAL) and with FMA contraction and measure, on the
C1 drop (128^2, density ratio 1000), how far each departs from the
canonical build: ~1e-13 after 100 steps, ~1e-6 after 1000 (the drop
amplifies rounding), so a 1e-11 agreement after 1000 steps holds only
within one fixed operation order."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "clip_oracle.cpp")

RUN = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[3])
from clip_amd import Config
from tests import cases
from tests.oracle_lib import Oracle
cfg = Config.from_string(cases.drop2d(n=128, radius=25))
o = Oracle(cfg).init()
out = {}
done = 0
for n in [int(x) for x in sys.argv[2].split(",")]:
    o.step(n - done)
    done = n
    for f in ("phi", "u", "p"):
        out[f"{f}_{n}"] = o.interior(f).copy()
np.savez(sys.argv[1], **out)
"""


def has_fma():
    try:
        return " fma " in open("/proc/cpuinfo").read().replace("\n", " ")
    except OSError:
        return False


def build(tmp, name, flags):
    lib = os.path.join(str(tmp), f"liborc_{name}.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", *flags, "-fopenmp", "-fPIC", "-shared",
                    "-I", os.path.join(ROOT, "include"), SRC, "-o", lib], check=True)
    return lib


def run(tmp, lib, steps):
    out = os.path.join(str(tmp), os.path.basename(lib) + ".npz")
    env = dict(os.environ, CLIP_ORACLE_LIB=lib, OMP_NUM_THREADS="4")
    subprocess.run([sys.executable, "-c", RUN, out, ",".join(map(str, steps)), ROOT], check=True, env=env)
    return np.load(out)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def variants(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("oracle_order")
    libs = {"canonical": build(tmp, "canonical", []), "literal": build(tmp, "literal", ["-DCLIP_ORACLE_LITERAL"])}
    if has_fma():
        fma = ["-march=x86-64-v3", "-ffp-contract=fast"]
        libs["canonical_fma"] = build(tmp, "canonical_fma", fma)
        libs["literal_fma"] = build(tmp, "literal_fma", ["-DCLIP_ORACLE_LITERAL", *fma])
    return tmp, libs


@pytest.mark.parametrize("dim", [2, 3])
def test_literal_filter_is_the_same_linear_map(variants, dim):
    """The literal dense transform and the canonical stress-part form are the
    same map (M^-1 M = I) up to rounding."""
    _, libs = variants
    Q = 9 if dim == 2 else 19
    rng = np.random.default_rng(dim)
    out = {}
    for name in ("canonical", "literal"):
        lib = ctypes.CDLL(libs[name])
        lib.orc_filter.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_double]
        vals = []
        for trial in range(20):
            v = rng.standard_normal(Q) if name == "canonical" else out["canonical_in"][trial].copy()
            if name == "canonical":
                out.setdefault("canonical_in", []).append(v.copy())
            s = 0.3 + 0.05 * trial
            lib.orc_filter(dim, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), s)
            vals.append(v)
        out[name] = np.array(vals)
    err = np.abs(out["literal"] - out["canonical"]).max() / np.abs(out["canonical"]).max()
    assert err < 1e-14, err


def test_operation_order_drift_on_c1(variants):
    tmp, libs = variants
    res = {name: run(tmp, lib, [100, 1000]) for name, lib in libs.items()}
    ref = res["canonical"]
    for name, d in res.items():
        if name == "canonical":
            continue
        r100 = {f: rel(d[f"{f}_100"], ref[f"{f}_100"]) for f in ("phi", "u", "p")}
        r1000 = {f: rel(d[f"{f}_1000"], ref[f"{f}_1000"]) for f in ("phi", "u", "p")}
        print(f"{name} vs canonical: 100 steps " + " ".join(f"{f} {v:.1e}" for f, v in r100.items())
              + " | 1000 steps " + " ".join(f"{f} {v:.1e}" for f, v in r1000.items()), flush=True)
        assert max(r100.values()) < 1e-12, (name, r100)
        # the drop amplifies the rounding difference by many orders within
        # 1000 steps, yet stays a rounding-level trajectory difference
        assert 1e-9 < max(r1000.values()) < 1e-4, (name, r1000)
