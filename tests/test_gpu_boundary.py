"""GPU: error conventions of the upload exports (lib/memory.c:286-385).

The reference's add_texture returns -1 once MAX_TEXTURES (65 536, lib/common.h:35) volumes
exist, after printing "ERROR: Tried allocating too many textures" (lib/memory.c:287-290);
here a zero-size volume is rejected the same way.  Each test starts from a fresh library
state and leaves a fresh one behind.
"""
import ctypes

import numpy as np
import pytest

from vtrace_amd import renderer as R

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_zero_size_and_null_volumes_are_rejected(fresh_renderer):
    lib = fresh_renderer.lib
    _say("case: zero-size add_texture")
    data = (ctypes.c_uint8 * 64)(*([255] * 64))
    pal = (ctypes.c_uint32 * 256)()
    for dims in ((0, 1, 1), (1, 0, 1), (1, 1, 0), (0, 0, 0)):
        assert lib.add_texture(data, *dims) == -1
        assert lib.vt_add_volume_indexed(data, *dims, pal) == -1
    assert lib.add_texture(None, 1, 1, 1) == -1
    assert lib.vt_add_textures(data, 0, 1, 1, 1, None) == -1
    assert lib.add_texture(data, 2, 2, 2) == 0      # the library still works afterwards
    R.reset_renderer()


def test_max_textures_then_minus_one(fresh_renderer, capfd):
    lib = fresh_renderer.lib
    _say("case: MAX_TEXTURES (65535 batched 1^3 volumes, then two add_texture)")
    n = 65535
    vols = np.full((n, 4), 255, dtype=np.uint8)
    ids = np.zeros(n, dtype=np.int32)
    assert lib.vt_add_textures(vols.ctypes.data, n, 1, 1, 1, ids.ctypes.data) == 0
    assert ids[0] == 0 and ids[-1] == n - 1
    one = (ctypes.c_uint8 * 4)(255, 255, 255, 255)
    assert lib.add_texture(one, 1, 1, 1) == 65535             # the last id below MAX_TEXTURES
    assert lib.add_texture(one, 1, 1, 1) == -1                # lib/memory.c:287-290
    assert lib.vt_add_textures(one, 1, 1, 1, 1, None) == -1
    pal = (ctypes.c_uint32 * 256)()
    assert lib.vt_add_volume_indexed(one, 1, 1, 1, pal) == -1
    assert "Tried allocating too many textures" in capfd.readouterr().err
    R.reset_renderer()
