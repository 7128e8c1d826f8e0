"""Writes rng_kat.json: known-answer vectors for the MPD counter-based RNG.

Computed with Python integers (arbitrary precision, reduced mod 2^64 by
hand), a transcription independent of the C oracle and the HIP kernels, of:
  hash:          src/rdme/dev/xor_random_dev.cu:46-53
  timestep hash: src/rdme/MpdRdmeSolver.cu:987-994
  uniform:       xor_random_dev.cu:66-81, (float)u32 * 2.32830643654e-10f,
                 the float conversion done by numpy (round to nearest even).
Run: python tests/golden/make_rng_kat.py
"""
import json
import os

import numpy as np

M = (1 << 64) - 1


def hash32(v, seed):
    v = (v ^ seed) & M
    v = (v * 3935559000370003845 + 2691343689449507681) & M
    v ^= v >> 21
    v = (v ^ (v << 37)) & M
    v ^= v >> 4
    v = (v * 2685821657736338717) & M
    return v >> 32


def timestep_hash(seed32, t, s):
    h = ((((seed32 << 30) + t) << 2) + s) & M
    h = (h * 3202034522624059733 + 4354685564936845319) & M
    h ^= h >> 20
    h = (h ^ (h << 41)) & M
    h ^= h >> 5
    return (h * 7664345821815920749) & M


def uniform_bits(v1, shift, v2, seed):
    r = hash32(((v1 << shift) | v2) & M, seed)
    f = np.float32(r) * np.float32(2.32830643654e-10)
    return int(np.array([f], dtype=np.float32).view(np.uint32)[0])


def main():
    rng = np.random.default_rng(20261202)
    out = {"hash": [], "timestep_hash": [], "uniform": []}
    vals = [0, 1, 2, 0xFFFFFFFF, 1 << 32, M] + [int(x) for x in rng.integers(0, 2 ** 63, 40, dtype=np.int64)]
    seeds = [0, 1, M] + [int(x) for x in rng.integers(0, 2 ** 63, 10, dtype=np.int64)]
    for v in vals:
        for s in seeds[:5]:
            out["hash"].append([str(v), str(s), hash32(v, s)])
    for seed32 in [0, 1, (11 << 16) | 3, 0xFFFFFFFF, 0xDEAD0007]:
        for t in [0, 1, 2, 1000, 123456, 0xFFFFFFFF]:
            for s in range(4):
                out["timestep_hash"].append([seed32, t, s, str(timestep_hash(seed32, t, s))])
    for i in [0, 1, 5, 33554431, 4294967295 >> 4] + [int(x) for x in rng.integers(0, 1 << 25, 20)]:
        for shift, slots in ((3, 8), (4, 16), (2, 4), (1, 2)):
            for j in range(0, slots, max(1, slots // 4)):
                seed = timestep_hash((7 << 16) | 1, 42, j % 4)
                out["uniform"].append([i, shift, j, str(seed), uniform_bits(i, shift, j, seed)])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_kat.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print(path, {k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
