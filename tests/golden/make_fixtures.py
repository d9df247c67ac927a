"""Writes the oracle fixtures under tests/golden/fixtures/ (run from the repo
root: python tests/golden/make_fixtures.py).

Each fixture holds the case text, the steps and phi, u, p (interior nodes,
FP64) written by the serial FP64 oracle.  test_golden.py re-runs the oracle
on the same case (a regression pin of the restatement) and the GPU tests
compare the HIP path against the stored vectors.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from clip_amd import Config  # noqa: E402
from tests import cases  # noqa: E402
from tests.oracle_lib import Oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures")

CASES = {
    "drop2d_16": (cases.drop2d(n=16, radius=4), [1, 10, 100]),
    "rti2d_12x20": (cases.rti2d(12, 20), [50]),
    "bubble3d_10": (cases.bubble3d(n=10, radius=3), [20]),
    "jet3d_10x10x16": (cases.jet3d((12, 12, 18), radius=2.5, inlet_side="z-"), [30]),
}


def main():
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    os.makedirs(OUT, exist_ok=True)
    for name, (text, steps) in CASES.items():
        cfg = Config.from_string(text)
        o = Oracle(cfg).init()
        done = 0
        arrays = {"config": np.frombuffer(text.encode(), dtype=np.uint8), "steps": np.array(steps)}
        for s in steps:
            o.step(s - done)
            done = s
            for f in ("phi", "u", "p"):
                arrays[f"{f}_{s}"] = o.interior(f)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **arrays)
        print(name, steps)


if __name__ == "__main__":
    main()
