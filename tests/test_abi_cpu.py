"""CPU: the C-ABI library loads and exports every symbol include/vtrace.h declares.

No compute calls are made here (no GPU in this environment)."""
import ctypes
import os
import re

import pytest

from vtrace_amd import _native, abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vtrace.h")
SCENE_HEADER = os.path.join(ROOT, "include", "vtscene.h")


def header_functions(path=HEADER):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_header_declares_reference_exports():
    names = header_functions()
    for n in abi.REFERENCE_EXPORTS:
        assert n in names, n
    for n in abi.EXTENSION_EXPORTS:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_native.RENDER_LIB):
        pytest.fail("librender_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_native.RENDER_LIB)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_reference():
    # lib/common.h:145-151: u8 keys[6] then 4 doubles (8-byte aligned) -> 40 bytes
    assert ctypes.sizeof(abi.UserInput) == 40
    assert abi.UserInput.mouse_x.offset == 8
    # lib/common.h:89-92: two pointers
    assert ctypes.sizeof(abi.RenderTickInfo) == 2 * ctypes.sizeof(ctypes.c_void_p)
    assert ctypes.sizeof(abi.VtRenderParams) == 4 * 9 + 4 * 9 + 4


def test_python_binding_declares_all():
    lib = _native.render_lib()
    for n in abi.REFERENCE_EXPORTS + abi.EXTENSION_EXPORTS:
        assert getattr(lib, n).restype is not None or n in ("cleanup", "vt_serve_shutdown", "vt_dn_shutdown", "vt_set_cursor")


def test_scene_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.SCENE_LIB)
    names = header_functions(SCENE_HEADER)
    assert "vt_vox_model" in names and "vt_gen_volume" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
