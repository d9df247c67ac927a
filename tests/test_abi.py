"""The C-ABI library loads without a GPU and exports every entry point that
include/clip_amd.h declares; the ctypes mirror matches the C struct layout."""
import ctypes
import os
import re
import subprocess
import tempfile

from clip_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "clip_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(clip_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers all of them
    assert set(names) <= set(_abi.SIGNATURES)


def test_signatures_match_header_prototypes():
    """Each prototype's parameter count equals its ctypes binding's, and the
    interior-layout constants agree."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    protos = dict(re.findall(r"\b(clip_[a-z_0-9]+)\s*\(([^)]*)\)\s*;", text))
    assert len(protos) >= 20
    for name, params in protos.items():
        params = params.strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        assert len(_abi.SIGNATURES[name][1]) == n, name
    defines = dict(re.findall(r"#define\s+(CLIP_INTERIOR_X_[A-Z]+)\s+(\d+)", text))
    assert int(defines["CLIP_INTERIOR_X_FASTEST"]) == _abi.INTERIOR_X_FASTEST
    assert int(defines["CLIP_INTERIOR_X_SLOWEST"]) == _abi.INTERIOR_X_SLOWEST


def test_ctypes_struct_layout_matches_header():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "clip_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(clip_config), offsetof(clip_config, t_final),
         offsetof(clip_config, sigma_override), offsetof(clip_config, bc_type),
         offsetof(clip_config, geom_center), offsetof(clip_config, inlet_mode_phase), sizeof(clip_info));
  printf("%zu\n", offsetof(clip_config, mobility_override));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        vals = [int(x) for x in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
    C = _abi.ClipConfig
    want = [ctypes.sizeof(C), C.t_final.offset, C.sigma_override.offset, C.bc_type.offset,
            C.geom_center.offset, C.inlet_mode_phase.offset, ctypes.sizeof(_abi.ClipInfo),
            C.mobility_override.offset]
    assert vals == want


def test_create_without_device_fails_loudly():
    """No silent CPU fallback: creating a solver needs a HIP device."""
    import pytest
    from clip_amd import Config, Solver, ClipError
    from tests import cases
    cfg = Config.from_string(cases.drop2d(n=16, radius=4))
    try:
        s = Solver(cfg)
    except ClipError as e:
        assert e.code in (_abi.ERR_HIP, _abi.ERR_NOMEM)
        return
    s.close()
    pytest.skip("a HIP device is present")


def test_error_paths_without_device():
    lib = _abi.load()
    assert lib.clip_step(None, 1) == _abi.ERR_ARG
    assert b"null" in lib.clip_last_error()
    assert lib.clip_checkpoint_rotate(b"/nonexistent/dir/for/rotation", 3) == _abi.OK


def test_geometry_sdf_through_the_abi():
    """Geometry::sdf (Geometry.cuh:102-161) is exported for host callers."""
    from clip_amd import Config
    from tests import cases
    cfg = Config.from_string(cases.drop2d(n=32, radius=8))
    c = (32 + 1) / 2.0
    assert abs(cfg.sdf(c + 3.0, c + 4.0) - (5.0 - 8.0)) < 1e-12
    assert cfg.sdf(c, c, geometry_id=5) == 1e10


def test_mobility_override_survives_derivation():
    """RTI derives the mobility from Pe (InputData.cu:76-114); an explicit
    override is kept."""
    from clip_amd import Config
    from tests import cases
    cfg = Config.from_string(cases.rti2d(32, 64))
    derived = cfg.mobility
    c2 = cfg.with_params(mobility=0.5 * derived)
    assert abs(c2.mobility - 0.5 * derived) < 1e-18
    assert abs(Config.from_string(cases.rti2d(32, 64)).derive().mobility - derived) < 1e-18


def test_set_field_rejects_derived_fields_and_bad_sizes_before_reading():
    """clip_set_field rejects RHO / MU / DPHI / NORMAL before touching the
    input (a null-backed buffer would fault if it were read); Solver.set
    checks the size."""
    lib = _abi.load()
    # null handle first, then derived fields on a null handle are ARG errors
    assert lib.clip_set_field(None, _abi.FIELD["rho"], 0, None) == _abi.ERR_ARG
