"""Writes rng_vectors.json: known answers of the hash RNG, computed here in pure Python from
the published algorithm (Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT 9(3)
2020, listing of `pcg`: one PCG-RXS-M-XS 32-bit round), independently of the oracle and of
the HIP kernels, plus the per-(seed, frame, pixel, sample) stream start this project uses:
pcg(pixel ^ pcg(sample ^ pcg(frame ^ pcg(seed))))."""
import json
import os

M = 0xFFFFFFFF


def pcg(v):
    state = (v * 747796405 + 2891336453) & M
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M
    return ((word >> 22) ^ word) & M


def seed(s, f, p, k):
    return pcg(p ^ pcg(k ^ pcg(f ^ pcg(s))))


if __name__ == "__main__":
    vals = [0, 1, 2, 3, 42, 255, 256, 65535, 65536, 123456789, 0x7FFFFFFF, 0x80000000, 0xDEADBEEF, 0xFFFFFFFF]
    out = {"pcg_hash": [[v, pcg(v)] for v in vals],
           "rng_seed": [[[s, f, p, k], seed(s, f, p, k)] for s, f, p, k in
                        [(0, 0, 0, 0), (1, 0, 0, 0), (2, 0, 12345, 3), (3, 7, 2073599, 15), (0xFFFFFFFF, 1, 1, 1)]]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_vectors.json")
    json.dump(out, open(path, "w"), indent=1)
    print("wrote", path)
