"""CPU: the .vox reader (vtrace_amd/vox.py, vox.cpp) against the MagicaVoxel format and the
reference's asset path (src/voxel/magica_voxel.rs:18-44, render.rs:246-262, rawchunk.rs:287-296).

The mapping is pinned in closed form on files written here: file voxel (x, y, z, i) lands at
chunk (x, size.y - 1 - z, y), the chunk's raw array (index z + dz * (y + dy * x)) is what
add_texture receives as (w, h, d), and its colour is palette entry i - 1.  The reference's own
assets are checked against the voxel counts their SIZE/XYZI headers declare."""
import os
import struct

import numpy as np
import pytest

from vtrace_amd import vox

ASSETS = "/root/reference/assets"


def _chunk(cid: bytes, content: bytes, children: bytes = b"") -> bytes:
    return cid + struct.pack("<II", len(content), len(children)) + content + children


def _vox(models, palette=None, extra=b"") -> bytes:
    kids = b""
    for size, voxels in models:
        kids += _chunk(b"SIZE", struct.pack("<iii", *size))
        kids += _chunk(b"XYZI", struct.pack("<I", len(voxels)) + b"".join(struct.pack("<BBBB", *v) for v in voxels))
    kids += extra
    if palette is not None:
        kids += _chunk(b"RGBA", b"".join(struct.pack("<I", int(c)) for c in palette))
    return b"VOX " + struct.pack("<I", 150) + _chunk(b"MAIN", b"", kids)


def _palette(seed=3):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 2**24, 256, dtype=np.uint64) | (255 << 24)).astype(np.uint32)


def test_cubic_model_mapping_and_palette_offset():
    D = 5
    rng = np.random.default_rng(1)
    cells = {tuple(int(c) for c in rng.integers(0, D, 3)) for _ in range(30)}
    voxels = [(x, y, z, int(rng.integers(1, 256))) for (x, y, z) in cells]
    pal = _palette()
    (v,) = vox.parse(_vox([((D, D, D), voxels)], pal))
    assert v.dims == (D, D, D) and v.data.shape == (D, D, D)
    assert v.palette[0] == 0 and np.array_equal(v.palette[1:], pal[:255])
    want = np.zeros((D, D, D), dtype=np.uint8)  # [k][j][i]: texel (i, j, k) = chunk (z=i, y=j, x=k)
    for x, y, z, i in voxels:
        X, Y, Z = x, D - 1 - z, y
        flat = Z + D * (Y + D * X)          # chunk raw index
        want.reshape(-1)[flat] = i
    assert np.array_equal(v.data, want)
    # texel (i, j, k) = (y, D-1-z, x) for file voxel (x, y, z)
    for x, y, z, i in voxels:
        assert v.data[x, D - 1 - z, y] == i


def test_index_zero_and_models_in_order():
    pal = _palette(4)
    vs = vox.parse(_vox([((2, 2, 2), [(0, 0, 0, 0), (1, 1, 1, 7)]), ((3, 3, 3), [(2, 0, 1, 9)])], pal))
    assert [v.dims for v in vs] == [(2, 2, 2), (3, 3, 3)]
    assert vs[0].data[0, 1, 0] == 1          # index 0 takes palette entry 0, like index 1
    assert vs[0].palette[1] == pal[0]
    assert vs[1].data[2, 3 - 1 - 1, 0] == 9


def test_errors():
    pal = _palette()
    with pytest.raises(vox.VoxError):  # voxel z >= size.y lands outside the chunk (the reference panics)
        vox.parse(_vox([((4, 2, 6), [(0, 0, 5, 1)])], pal))
    with pytest.raises(vox.VoxError):  # no RGBA chunk
        vox.parse(_vox([((2, 2, 2), [(0, 0, 0, 1)])], None))
    good = _vox([((2, 2, 2), [(0, 0, 0, 1)])], pal)
    for cut in (10, 30, len(good) - 5):
        with pytest.raises(vox.VoxError):
            vox.parse(good[:cut])
    with pytest.raises(vox.VoxError):
        vox.parse(b"NOTVOX" + good[6:])
    # unknown chunks are skipped
    (v,) = vox.parse(_vox([((2, 2, 2), [(1, 0, 0, 3)])], pal, extra=_chunk(b"nTRN", b"\0" * 28)))
    assert int((v.data != 0).sum()) == 1


@pytest.mark.skipif(not os.path.isdir(ASSETS), reason="reference assets not present")
@pytest.mark.parametrize("name,size,count", [("Treasure.vox", 50, 37069), ("AncientTemple.vox", 40, 12225)])
def test_reference_assets(name, size, count):
    path = os.path.join(ASSETS, name)
    (v,) = vox.load(path)
    assert v.dims == (size, size, size)
    assert int((v.data != 0).sum()) == count
    raw = open(path, "rb").read()
    at = raw.index(b"RGBA")
    filepal = np.frombuffer(raw[at + 12:at + 12 + 1024], dtype="<u4")
    assert np.array_equal(v.palette[1:], filepal[:255])
