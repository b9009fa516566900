"""Writes scene_digests.json: SHA-256 digests of small seeded volumes and palettes from the
scene generator (vtrace_amd/csrc/scenegen.cpp), so an accidental change of the synthetic
workloads is caught.  Regenerate only on an intended change of the generator."""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from vtrace_amd import scenes  # noqa: E402

CASES = [((32, 32, 32), 1, 0.30), ((64, 64, 64), 2002, 0.25), ((50, 40, 30), 7, 0.10), ((16, 16, 16), 3, 0.0)]


def digests():
    out = {}
    for dims, seed, fill in CASES:
        v = scenes.gen_volume(dims, seed, fill)
        out[f"vol_{dims[0]}x{dims[1]}x{dims[2]}_s{seed}_f{fill}"] = hashlib.sha256(v.tobytes()).hexdigest()
    for seed in (1, 2002):
        out[f"pal_s{seed}"] = hashlib.sha256(scenes.gen_palette(seed).tobytes()).hexdigest()
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scene_digests.json")
    json.dump(digests(), open(path, "w"), indent=1)
    print("wrote", path)
