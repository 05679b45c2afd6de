"""Random rigid bodies (pixel grids + fixed-point poses) for stamp tests."""
import math

import numpy as np

import falling_sand_engine_amd as fsg


def random_bodies(rng, W, H, n, max_side=16, angles=None, mats=(1, 7)):
    bodies, pixels = [], []
    off = 0
    for i in range(n):
        mw, mh = int(rng.integers(4, max_side + 1)), int(rng.integers(4, max_side + 1))
        x = float(rng.uniform(0, W))
        y = float(rng.uniform(0, H))
        ang = float(rng.uniform(0, 2 * math.pi)) if angles is None else angles[i % len(angles)]
        b = fsg.body_from_pose(x, y, ang, mw, mh, off, enabled=bool(rng.random() > 0.1))
        m = rng.choice(mats, size=(mh, mw))
        hole = rng.random((mh, mw)) < 0.15
        m = np.where(hole, 0, m)
        t = rng.integers(0, 4000, size=(mh, mw))
        px = (m.astype(np.uint64) | (t.astype(np.uint64) << np.uint64(16)) |
              (rng.integers(0, 2**32, size=(mh, mw), dtype=np.uint64) << np.uint64(32)))
        px[m == 0] = np.uint64(0)
        bodies.append(b)
        pixels.append(px.ravel())
        off += mw * mh
    return bodies, np.concatenate(pixels) if pixels else np.zeros(0, dtype=np.uint64)
