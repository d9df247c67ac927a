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


def test_ctypes_struct_layout_matches_header():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "clip_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(clip_config), offsetof(clip_config, t_final),
         offsetof(clip_config, sigma_override), offsetof(clip_config, bc_type),
         offsetof(clip_config, geom_center), offsetof(clip_config, inlet_mode_phase), sizeof(clip_info));
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
            C.geom_center.offset, C.inlet_mode_phase.offset, ctypes.sizeof(_abi.ClipInfo)]
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
