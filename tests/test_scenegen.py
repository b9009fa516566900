"""CPU: the seeded synthetic scene generator (exact fill, determinism, committed digests)."""
import json
import os
import sys

import numpy as np
import pytest

from vtrace_amd import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)


@pytest.mark.parametrize("dims,fill", [((64, 64, 64), 0.3), ((50, 40, 30), 0.1), ((17, 9, 33), 0.77), ((8, 8, 8), 1.0)])
def test_exact_fill_and_index_range(dims, fill):
    v = scenes.gen_volume(dims, 5, fill)
    n = dims[0] * dims[1] * dims[2]
    assert int((v != 0).sum()) == round(fill * n)
    assert v.shape == (dims[2], dims[1], dims[0])


def test_occupied_are_lowest_noise_values():
    dims = (24, 20, 16)
    field = np.zeros(dims[0] * dims[1] * dims[2], dtype=np.float32)
    assert scenes.scene_lib().vt_gen_noise_field(*dims, 11, 4, field.ctypes.data) == 0
    v = scenes.gen_volume(dims, 11, 0.25).reshape(-1)
    occ = v != 0
    assert field[occ].max() <= field[~occ].min()


def test_boxes_forced_in():
    boxes = [[2, 3, 4, 10, 9, 12]]
    v = scenes.gen_volume((32, 32, 32), 3, 0.05, boxes=boxes)
    assert (v[4:12, 3:9, 2:10] != 0).all()


def test_determinism_and_digests():
    import make_scene_digests as M
    want = json.load(open(os.path.join(GOLDEN, "scene_digests.json")))
    assert M.digests() == want


def test_camera_matrices_are_gl_conventions():
    P = scenes.perspective(scenes.FOV, 4 / 3).reshape(4, 4)  # [c][r]
    assert P[2][3] == -1.0 and P[3][3] == 0.0
    C = scenes.look_at((1.0, 2.0, 3.0), (0.0, 0.0, 0.0)).reshape(4, 4).T.astype(np.float64)
    eye = np.linalg.inv(C) @ np.array([0, 0, 0, 1.0])
    assert np.allclose(eye[:3], [1, 2, 3], atol=1e-5)
    fwd = C @ np.array([0.0, 0.0, 0.0, 1.0])  # the target is on the -z axis of view space
    assert fwd[2] < 0 and abs(fwd[0]) < 1e-4 and abs(fwd[1]) < 1e-4
