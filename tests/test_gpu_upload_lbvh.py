"""GPU: volume upload (bricking) and LBVH build against the oracle, bit-exact."""
import sys

import numpy as np
import pytest

import oracle_ffi as O
from vtrace_amd import renderer as R
from vtrace_amd import scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dims,fmt,fill", [((64, 64, 64), 0, 0.3), ((50, 40, 30), 0, 0.1), ((33, 17, 9), 1, 0.4),
                                            ((8, 8, 8), 0, 0.0), ((24, 24, 24), 1, 1.0), ((200, 96, 128), 0, 0.02)])
def test_brick_roundtrip(dims, fmt, fill):
    _say(f"case: brick dims={dims} fmt={fmt} fill={fill}")
    r = R.reset_renderer()
    idx = scenes.gen_volume(dims, 21, fill)
    pal = scenes.gen_palette(21, opaque=(fmt == 0))
    vol = scenes.Volume(idx, pal) if fmt == 0 else scenes.Volume(scenes.expand_palette(idx, pal), None)
    tid = r.upload(vol)
    info = r.volume_info(tid)
    dense = r.volume_dense(tid)
    bmap, masks, payload = O.brick(vol)
    assert info["nbricks"] == int((bmap < 0xE0000000).sum())
    # brick map bit-exact: ids of non-empty bricks and the empty-space radius of empty ones
    assert np.array_equal(r.brick_map(tid), bmap)
    # voxel order after the nearest-texel remap; for these sizes the remap is the identity
    # except when size is not a power of two; compare against the oracle's remapped volume
    dz, dy, dx = vol.data.shape
    rx = np.array([O.lib().orc_nearest_texel(i, dx) for i in range(dx)])
    ry = np.array([O.lib().orc_nearest_texel(i, dy) for i in range(dy)])
    rz = np.array([O.lib().orc_nearest_texel(i, dz) for i in range(dz)])
    remapped = vol.data[np.ix_(rz, ry, rx)]
    occ = (remapped != 0) if fmt == 0 else ((remapped >> 24) != 0)
    # empty bricks read back as zero; occupied voxels must match exactly
    assert np.array_equal(dense[occ], remapped[occ])
    # popcount of the masks equals the occupied count
    pop = int(sum(bin(int(w)).count("1") for w in masks.reshape(-1)))
    assert pop == int(occ.sum())


def _records(n, seed, spread=4096.0, size=64.0):
    rng = scenes.Rng(seed)
    recs = []
    for i in range(n):
        t = tuple(rng.uniform(-spread, spread) for _ in range(3))
        s = (rng.uniform(0.5, 2.0) * size, rng.uniform(0.5, 2.0) * size, rng.uniform(0.5, 2.0) * size)
        recs.append(scenes.instance_record(scenes.model_matrix(t, s), i % 3))
    return np.stack(recs)


@pytest.mark.parametrize("n", [1, 2, 3, 16, 255, 1000, 4096, 5000, 9261, 40000])
def test_lbvh_matches_oracle(n):
    """Up to 4096 instances the sort runs in one workgroup, above it one workgroup per 1024-key
    chunk per pass: nodes, order and instance data equal the oracle's either way."""
    _say(f"case: lbvh n={n}")
    r = R.reset_renderer()
    for k in range(3):
        r.upload(scenes.Volume(scenes.gen_volume((8, 8, 8), k, 0.5), scenes.gen_palette(k)))
    recs = _records(n, 100 + n)
    if n >= 16:  # duplicate positions exercise the index tie-break of the Morton order
        recs[5] = recs[3]
    r.set_instances(recs)
    nodes, inst = r.lbvh()
    o_nodes, o_order, o_codes = O.lbvh(recs, 3)
    assert np.array_equal(nodes.view(np.uint32).reshape(-1, 8), o_nodes)
    inv, aabb, tex, valid = O.instance_prep(recs, 3)
    assert np.array_equal(inst["valid"], valid.astype(np.uint32))
    assert np.array_equal(inst["tex"], tex)
    # inverse rows 0..2 (column-major 3x4) bit-exact
    inv_rows = inv.reshape(n, 4, 4)[:, :, :3].reshape(n, 12)
    assert np.array_equal(inst["inv"].view(np.uint32), inv_rows.view(np.uint32))


@pytest.mark.parametrize("n", [1, 2, 37, 1000])
def test_refit_keeps_topology_and_bounds_moved_instances(n):
    """vt_refit_instances: same left/right links as the last build; leaf boxes equal the oracle's
    instance boxes of the moved records; every internal box is the union of its children's."""
    _say(f"case: lbvh refit n={n}")
    r = R.reset_renderer()
    for k in range(3):
        r.upload(scenes.Volume(scenes.gen_volume((8, 8, 8), k, 0.5), scenes.gen_palette(k)))
    recs = _records(n, 300 + n)
    r.set_instances(recs)
    before, _ = r.lbvh()
    builds = r.lbvh_builds()
    moved = recs.copy()
    rng = np.random.default_rng(n)
    moved[:, 12:15] += rng.uniform(-200.0, 200.0, (n, 3)).astype(np.float32)  # translation column
    r.refit_instances(moved)
    assert r.lbvh_builds() == builds
    after, inst = r.lbvh()
    assert np.array_equal(after["left"], before["left"]) and np.array_equal(after["right"], before["right"])
    _, aabb, _, _ = O.instance_prep(moved, 3)
    leaves = after[n - 1:]
    for k in range(n):
        i = int(leaves["left"][k])
        assert np.array_equal(leaves["lo"][k], aabb[i, :3]) and np.array_equal(leaves["hi"][k], aabb[i, 3:]), k
    for k in range(n - 1):
        lft, rgt = after[after["left"][k]], after[after["right"][k]]
        assert np.array_equal(after["lo"][k], np.minimum(lft["lo"], rgt["lo"]))
        assert np.array_equal(after["hi"][k], np.maximum(lft["hi"], rgt["hi"]))
    with pytest.raises(R.RendererError):
        r.refit_instances(moved[: max(1, n - 1)] if n > 1 else np.concatenate([moved, moved]))


def test_refit_frames_match_oracle_hits():
    """After a refit the frame's hits equal the oracle's (which builds its own tree for the moved
    instances): the closest hit does not depend on the traversal order."""
    sc = scenes.multi_volume_scene(count=8, n=64, world=512, seed=33, width=160, height=120, spp=1, max_bounces=2)
    r = R.reset_renderer()
    ids = R.load_scene(r, sc)
    r.render(sc.P, sc.C)
    moved = np.array(sc.records, np.float32).copy()
    rng = np.random.default_rng(5)
    moved[:, 12:15] += rng.uniform(-40.0, 40.0, (moved.shape[0], 3)).astype(np.float32)
    recs = moved.copy()
    recs.view(np.uint32)[:, 15] = np.array([ids[t] for t in moved.view(np.uint32)[:, 15]], dtype=np.uint32)
    _say("case: refit then render, 8 instances moved")
    r.refit_instances(recs)
    r.render(sc.P, sc.C)
    sc2 = scenes.Scene(sc.volumes, moved, sc.P, sc.C, sc.params, name="moved")
    o_hits, o_rad, _ = O.render(sc2, use_bvh=True)
    from parity_util import assert_hits_equal, radiance_error
    assert_hits_equal(r.hits(), o_hits, "refit")
    mx, _ = radiance_error(r.radiance(), o_rad)
    assert mx <= 1e-5
