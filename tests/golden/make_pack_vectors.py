"""Writes the golden .pack images of tests/golden/ (run from the repository
root after `make -C oracle ref`).

Each image is what Chunk::ChunkWrite (chunk.cpp:219-330) writes for a chunk:
the header, then LZ4_compress_fast(records, ..., 10) of the tiles + layer2
records and of the background, compressed by the reference's own lz4
(oracle/_ref, built from source/libs/lz4/lz4.c).  The records come from the
oracle's restatement of chunk.cpp:238-242; the chunk contents from the
benchmark generator (ticked by the oracle where `ticks` is set, so that sand
grains carry MOVED) and tests/packcases.py, so the tests can rebuild them.
`map` "reference" writes the reference's own material indices
(falling_sand_engine_amd.reference_pack_map).
"""
import hashlib
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

import falling_sand_engine_amd as fsg  # noqa: E402
import lz4ref  # noqa: E402
import packcases  # noqa: E402
from oracle import pyoracle  # noqa: E402

CASES = {
    # name: generator kind, W, H, seed, chunk (cx, cy), phase, layer2 seed, background seed
    "basin": dict(kind=2, W=512, H=512, seed=11, cx=1, cy=2, phase=3, l2=None, bg=None),
    "striped_aux": dict(kind=3, W=256, H=256, seed=4, cx=1, cy=0, phase=-2, l2=5, bg=9),
    "air": dict(kind=0, W=128, H=128, seed=0, cx=0, cy=0, phase=0, l2=None, bg=None),
    # a ticked all-active chunk: falling sand with MOVED set, water, steam
    "ticked": dict(kind=3, W=256, H=256, seed=6, cx=1, cy=1, phase=1, l2=None, bg=None, ticks=24),
    "ticked_refmap": dict(kind=3, W=256, H=256, seed=6, cx=1, cy=1, phase=1, l2=5, bg=None, ticks=24,
                          map="reference"),
}


def pack_map_of(c):
    return fsg.reference_pack_map() if c.get("map") == "reference" else None


def chunk_parts(c):
    cells = np.zeros((c["H"], c["W"]), np.uint64) if c["kind"] == 0 else fsg.generate(c["kind"], c["W"], c["H"], c["seed"])
    if c.get("ticks"):
        cells = pyoracle.run(cells, pyoracle.default_config(), c["seed"], 0, c["ticks"])
    tiles = packcases.chunk_of(cells, c["cx"], c["cy"])
    l2 = None if c["l2"] is None else packcases.layer2_pattern(128, 128, c["l2"])
    bg = None if c["bg"] is None else packcases.background_pattern(128, 128, c["bg"])
    return tiles, l2, bg


def main():
    assert lz4ref.lib() is not None, "build the reference lz4 first: make -C oracle ref"
    meta = {"lz4_version": lz4ref.lib().LZ4_versionNumber(), "cases": {}}
    for name, c in CASES.items():
        tiles, l2, bg = chunk_parts(c)
        ra, rb = pyoracle.chunk_records(tiles, l2, bg, pack_map_of(c))
        za = lz4ref.compress_fast(ra, 10)
        zb = lz4ref.compress_fast(rb, 10)
        img = struct.pack("<biiii", c["phase"], len(ra), len(za), len(rb), len(zb)) + za + zb
        with open(os.path.join(HERE, f"pack_{name}.pack"), "wb") as f:
            f.write(img)
        meta["cases"][name] = dict(c, bytes=len(img), sha256=hashlib.sha256(img).hexdigest(),
                                   raw_a_sha256=hashlib.sha256(ra.tobytes()).hexdigest(),
                                   raw_b_sha256=hashlib.sha256(rb.tobytes()).hexdigest())
        print(name, len(img))
    with open(os.path.join(HERE, "pack_vectors.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
