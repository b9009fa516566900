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


_BEFORE_ENTRY = r"""
import ctypes, sys
sys.path.insert(0, sys.argv[1])
from vtrace_amd import _native, abi
lib = _native.render_lib()
P = (ctypes.c_float * 16)(*([1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]))
info = abi.RenderTickInfo(ctypes.cast(P, ctypes.c_void_p), ctypes.cast(P, ctypes.c_void_p))
w, h = ctypes.c_int32(0), ctypes.c_int32(0)
data = (ctypes.c_uint8 * 4)(255, 255, 255, 255)
pal = (ctypes.c_uint32 * 256)()
buf = (ctypes.c_uint8 * 64)()
out = [lib.render_tick(ctypes.byref(w), ctypes.byref(h), ctypes.byref(info)),
       lib.end_update_instances(1),
       0 if lib.start_update_instances(1) else -1,
       lib.add_texture(data, 1, 1, 1),
       lib.vt_add_volume_indexed(data, 1, 1, 1, pal),
       lib.vt_add_textures(data, 1, 1, 1, 1, None),
       lib.vt_read_hits(buf, 64),
       lib.vt_refit_instances(1)]
lib.cleanup()   # before entry: nothing to free, returns
print(" ".join(str(x) for x in out))
print(lib.vt_last_error().decode())
"""


def test_calls_before_entry_fail_with_the_reference_conventions():
    """Every export used before `entry` fails the way the reference reports failure (-1, or
    NULL for the mapped instance buffer: lib/memory.c:235-267, lib/entry.c:150-248) without
    touching the device; run in a fresh process so no earlier `entry` is in effect."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, "-c", _BEFORE_ENTRY, ROOT], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    codes = [int(x) for x in r.stdout.split("\n")[0].split()]
    assert codes == [-1] * 8, codes
    assert "not initialised" in r.stdout
