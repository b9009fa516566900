"""Hit records of volumes wider than 1024 voxels and of walks longer than 65535 steps.

A hit record is 16 B: `voxel` holds bits 0-9 of the hit voxel's x, y, z, `vhi` bits 10-14
(x | y << 5 | z << 10); flags bit 2 marks a coordinate of 32768 or more (only its low 15
bits are held) and bit 3 a step count that saturated at 65535.

CPU: the oracle's records are pinned against the volume contents, not against its own packing
rule: in the wide volume every occupied voxel's palette index encodes its x (1 + x // 256), so
the decoded coordinate, through the texel NEAREST sampling picks for it, must land on an
occupied voxel whose index matches the coordinate's high bits (a dropped or shifted high part
moves it by 1024 or more, onto the wrong index); in the long walk the only
occupied voxels are the last x slab, so every hit took at least 65534 steps, the step count
must read 65535 exactly where it saturated (a wrapped count would be 0..2), and x = 65534 is
flagged as beyond 15 bits.
GPU: the same two frames, every field (vhi included) bit-exact against the oracle.
"""
import numpy as np
import pytest

import oracle_ffi as O
from parity_util import assert_hits_equal, gpu_render
from vtrace_amd import scenes
from vtrace_amd.renderer import hit_voxels

WIDE = (3000, 12, 12)
LONG = (65535, 3, 3)


def wide_scene():
    """3000 x 12 x 12 u8 volume, 10 % occupied, palette index 1 + x // 256, on a 3000 x 300 x 300
    box seen from above and in front in a 768 x 256 frame."""
    dx, dy, dz = WIDE
    rng = np.random.default_rng(41)
    occ = rng.random((dz, dy, dx)) < 0.10
    idx = (1 + np.arange(dx) // 256).astype(np.uint8)
    data = np.where(occ, idx[None, None, :], 0).astype(np.uint8)
    vol = scenes.Volume(data, scenes.gen_palette(41))
    rec = scenes.instance_record(scenes.model_matrix((0.0, 0.0, 0.0), (3000.0, 300.0, 300.0)), 0)[None, :]
    p = scenes.RenderParams(768, 256, viewport_square=0, mode=0)
    return scenes.Scene([vol], rec, scenes.perspective(scenes.FOV, 3.0),
                        scenes.look_at((1500.0, 900.0, 1400.0), (1500.0, 100.0, 150.0)), p, name="wide_3000")


def long_scene():
    """65535 x 3 x 3 u8 volume whose only occupied voxels are the slab x = 65534, on a
    1 x 30 x 30 box (centred on the origin), looked through from its -x side.  The walk runs in
    voxel space along the unscaled model-space direction (trace.frag:66), so y and z move by
    65534 * (dy / 30) / dx voxels over the walk: world slopes of about 1e-4 (a 2.3e-4 rad field
    of view) let every ray reach the far slab, crossing 0-2 y/z voxel boundaries on the way
    (65534-65536 steps)."""
    dx, dy, dz = LONG
    data = np.zeros((dz, dy, dx), dtype=np.uint8)
    data[:, :, dx - 1] = 5
    vol = scenes.Volume(data, scenes.gen_palette(42))
    rec = scenes.instance_record(scenes.model_matrix((0.0, 0.0, 0.0), (1.0, 30.0, 30.0)), 0)[None, :]
    p = scenes.RenderParams(64, 64, viewport_square=0, mode=0)
    s = 1.373e-4
    return scenes.Scene([vol], rec, scenes.perspective(2.3e-4, 1.0),
                        scenes.look_at((-1.5, 3.0, 2.5), (0.5, 3.0 + 2.0 * s, 2.5 + 2.2 * s)), p, name="long_65535")


def _check_wide(hits, vol):
    hit = hits["inst"] != 0xFFFFFFFF
    assert hit.sum() > 2000
    v = hit_voxels(hits[hit])
    assert v[:, 0].max() >= 2048 and v[:, 0].min() < 1024   # two high bits of x in use
    assert np.all(v < np.array(WIDE))
    # occupancy is read at the texel NEAREST sampling at voxel / size picks (trace.frag:77),
    # which for non-power-of-two dims can be the neighbour
    t = np.stack([[O.lib().orc_nearest_texel(int(c), WIDE[a]) for c in v[:, a]] for a in range(3)], axis=1)
    assert np.all(np.abs(t - v) <= 1)
    idx = vol.data[t[:, 2], t[:, 1], t[:, 0]]
    assert np.all(idx != 0)                                  # an occupied texel
    assert np.array_equal(idx, 1 + t[:, 0] // 256)           # ... at the coordinate recorded
    assert np.array_equal(hits["pal"][hit], idx)
    assert not np.any(hits["flags"][hit] & 12)


def _check_long(hits):
    hit = hits["inst"] != 0xFFFFFFFF
    assert hit.all()
    v = hit_voxels(hits)
    # x = 65534 does not fit 15 bits: flagged, its low 15 bits held
    assert np.all(v[..., 0] == (LONG[0] - 1) & 32767) and np.all(hits["flags"] & 4)
    assert np.all(v[..., 1:] < 3)
    steps = hits["steps"].astype(np.int64)
    sat = (hits["flags"] & 8) != 0
    assert sat.sum() > 100 and (~sat).sum() > 100
    assert np.all(steps[sat] == 65535)
    assert np.all(steps[~sat] >= 65534)                     # the whole length was walked


def test_oracle_records_of_a_volume_wider_than_1024():
    sc = wide_scene()
    hits, _, _ = O.render(sc)
    _check_wide(hits, sc.volumes[0])


def test_oracle_step_count_saturates_on_a_65535_voxel_walk():
    sc = long_scene()
    hits, _, _ = O.render(sc)
    _check_long(hits)


def test_hit_voxel_decoding_closed_form():
    rec = np.zeros(4, dtype=O.HIT_DTYPE)
    for i, (x, y, z) in enumerate(((0, 0, 0), (1023, 1024, 1025), (32767, 5000, 17), (2047, 31744, 1))):
        rec[i]["voxel"] = (x & 1023) | (y & 1023) << 10 | (z & 1023) << 20
        rec[i]["vhi"] = (x >> 10) | (y >> 10) << 5 | (z >> 10) << 10
        assert tuple(int(c) for c in hit_voxels(rec[i])) == (x, y, z)


@pytest.mark.gpu
@pytest.mark.parametrize("make", [wide_scene, long_scene], ids=["wide_3000", "long_65535"])
def test_gpu_hit_records_beyond_10_and_16_bits(make):
    sc = make()
    print(f"case: hit records, {sc.name}", flush=True)
    g_hits, _, _ = gpu_render(sc)
    o_hits, _, _ = O.render(sc)
    assert_hits_equal(g_hits, o_hits, sc.name)
    if sc.name == "wide_3000":
        _check_wide(g_hits, sc.volumes[0])
    else:
        _check_long(g_hits)
