"""Mutation checks of the oracle pins: each planted mistake (a sign, a weight,
a transposed index, a dropped condition) is compiled into its own build of
oracle/clip_oracle.cpp (-DCLIP_ORACLE_MUTANT=k) and the pin test of the
mutated function must fail against it, while passing against the real
oracle.  This shows the pins constrain what they claim to; it runs only on
the CPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "clip_oracle.cpp")
PINS = "tests/test_oracle_invariants.py"

# mutant id -> (what is planted, the pin test that must catch it)
MUTANTS = {
    1: ("inlet g: -N/weight also on the normal direction", "test_velocity_inlet_invariants_on_random_ghosts"),
    2: ("inlet f: tangential correction over weight + 1", "test_velocity_inlet_invariants_on_random_ghosts"),
    3: ("inlet f: tangential correction with the wrong sign", "test_velocity_inlet_invariants_on_random_ghosts"),
    4: ("inlet f: equilibrium difference reversed", "test_velocity_inlet_invariants_on_random_ghosts"),
    5: ("inlet outside the disc: copy without reversing the direction",
        "test_velocity_inlet_invariants_on_random_ghosts"),
    6: ("initial g: interface term with the wrong sign", "test_initial_g_moments"),
    7: ("initial g: interface term with a 3M factor", "test_initial_g_moments"),
    8: ("normal: divided by |grad phi|^2", "test_normal_is_unit_gradient_direction"),
    9: ("normal: gradient components permuted", "test_normal_is_unit_gradient_direction"),
    10: ("specular slip: slip sides bounce back through the wall map",
         "test_specular_slip_keeps_tangential_momentum_and_has_zero_wall_flux"),
    11: ("specular slip: each entering direction paired with the next list entry",
         "test_specular_slip_keeps_tangential_momentum_and_has_zero_wall_flux"),
}


def _run_pin(test, lib=None):
    env = dict(os.environ)
    if lib:
        env["CLIP_ORACLE_LIB"] = lib
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                           f"{PINS}::{test}"], cwd=ROOT, env=env, capture_output=True, text=True)


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("oracle_mutants")


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_pin_catches_planted_mistake(k, mutant_dir):
    what, test = MUTANTS[k]
    lib = os.path.join(str(mutant_dir), f"libclip_oracle_m{k}.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
                    f"-DCLIP_ORACLE_MUTANT={k}", "-I", os.path.join(ROOT, "include"), SRC, "-o", lib], check=True)
    r = _run_pin(test, lib)
    assert r.returncode != 0, f"mutant {k} ({what}) passed {test}:\n{r.stdout[-2000:]}"
    assert "failed" in r.stdout, r.stdout[-2000:]


def test_pins_pass_on_the_real_oracle():
    for test in sorted({t for _, t in MUTANTS.values()}):
        r = _run_pin(test)
        assert r.returncode == 0, r.stdout[-2000:]
