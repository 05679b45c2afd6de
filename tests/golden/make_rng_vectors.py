"""Writes rng_vectors.json: splitmix64, lowbias32, phase keys and block
hashes computed with Python big-int arithmetic, independently of the C code.

splitmix64 is S. Vigna's generator (public-domain reference
implementation, prng.di.unimi.it/splitmix64.c); lowbias32 is C. Wellons'
hash-prospector mixer.  The per-phase key and block hash follow the
definitions in oracle/fsg_oracle.c.
"""
import json
import os

M64 = (1 << 64) - 1
M32 = (1 << 32) - 1
GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(state):
    z = (state + GOLDEN) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def lowbias32(x):
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & M32
    x ^= x >> 16
    return x


def phase_key(seed, g):
    return splitmix64(splitmix64(seed) ^ g)


def block_hash(key, x0, y0):
    pos = ((x0 + 1) & M32) | (((y0 + 1) << 16) & M32)
    r = lowbias32(pos ^ (key & M32))
    s = lowbias32(r ^ (key >> 32))
    return r, s


def main():
    out = {
        # the generator's stream from state 0: next() returns splitmix64(k*GOLDEN)
        "splitmix64_stream_seed0": [splitmix64((k * GOLDEN) & M64) for k in range(8)],
        "splitmix64": [[x, splitmix64(x)] for x in
                       [0, 1, 2, 42, 0xDEADBEEF, M64, 1 << 63, 0x0123456789ABCDEF]],
        "lowbias32": [[x, lowbias32(x)] for x in [0, 1, 2, 0x80000000, M32, 0x12345678, 0xCAFEBABE]],
        "phase_key": [[s, g, phase_key(s, g)] for s in [0, 1, 3, 1000] for g in [0, 1, 2, 3, 299, 1 << 40]],
        "block_hash": [[k, x, y, *block_hash(k, x, y)] for k in [0, phase_key(1, 0), phase_key(7, 12)]
                       for (x, y) in [(-1, -1), (0, 0), (5, -1), (-1, 9), (511, 511), (16383, 100)]],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
