"""fsg_world_set_cells_noalloc (the batched setTile staged in a fixed device
array, no allocation) against fsg_world_set_cells, through the C-ABI: the
same grids after two calls in a row, the second with fewer edits, with
repeated positions (last edit wins), out-of-world edits, and the size and
material checks."""
import ctypes as C

import numpy as np
import pytest

import falling_sand_engine_amd as fsg
from falling_sand_engine_amd import _abi

pytestmark = pytest.mark.gpu

MAX = 65536  # FSG_SET_CELLS_NOALLOC_MAX


def make_edits(rng, n, W, H, nmat=8):
    """n edits: a third on a small pool of positions (many repeats), a few
    outside the world, the rest anywhere."""
    arr = (_abi.CellEdit * n)()
    pool = np.stack([rng.integers(0, W, 64), rng.integers(0, H, 64)], 1)
    for i in range(n):
        r = rng.random()
        if r < 0.33:
            x, y = pool[int(rng.integers(0, 64))]
        elif r < 0.36:
            x, y = int(rng.choice([-2, -1, W, W + 5])), int(rng.integers(-2, H + 2))
        else:
            x, y = rng.integers(0, W), rng.integers(0, H)
        arr[i].x, arr[i].y = int(x), int(y)
        arr[i].cell.mat = int(rng.integers(0, nmat)) | (_abi.MOVED if rng.random() < 0.2 else 0)
        arr[i].cell.temp = int(rng.integers(-3000, 3000))
        arr[i].cell.rgba = int(rng.integers(0, 2**32))
    return arr


def test_noalloc_set_cells_matches_set_cells(ctx):
    L = fsg.lib()
    rng = np.random.default_rng(7)
    W, H = 700, 300
    cells = fsg.generate(2, W, H, 3)
    print("set_cells_noalloc: two calls, 50000 then 3000 edits", flush=True)
    with fsg.World(ctx, W, H, 3) as a, fsg.World(ctx, W, H, 3) as b:
        a.upload(cells)
        b.upload(cells)
        for n in (50000, 3000):
            e = make_edits(rng, n, W, H)
            assert L.fsg_world_set_cells(a.handle, e, n) == 0
            assert L.fsg_world_set_cells_noalloc(b.handle, e, n) == 0
            ga, gb = a.download(), b.download()
            assert np.array_equal(ga, gb), int(np.count_nonzero(ga != gb))
            assert not np.array_equal(ga, cells)
        # then both tick alike (the edits count as new cells for chunk skipping too)
        a.tick(3)
        b.tick(3)
        assert np.array_equal(a.download(), b.download())


def test_noalloc_set_cells_limits(ctx):
    L = fsg.lib()
    rng = np.random.default_rng(8)
    W, H = 256, 128
    cells = fsg.generate(1, W, H, 5)
    print("set_cells_noalloc: limits", flush=True)
    with fsg.World(ctx, W, H, 5) as w:
        w.upload(cells)
        big = make_edits(rng, MAX + 1, W, H)
        assert L.fsg_world_set_cells_noalloc(w.handle, big, MAX + 1) == _abi.FSG_ERR_INVALID
        bad = make_edits(rng, 10, W, H)
        bad[4].cell.mat = 9  # no such material (8 in the table)
        assert L.fsg_world_set_cells_noalloc(w.handle, bad, 10) == _abi.FSG_ERR_INVALID
        assert np.array_equal(w.download(), cells)  # rejected calls change nothing
        assert L.fsg_world_set_cells_noalloc(w.handle, None, 0) == 0
        # exactly the maximum is accepted and equals the allocating entry point
        full = make_edits(rng, MAX, W, H)
        assert L.fsg_world_set_cells_noalloc(w.handle, full, MAX) == 0
        got = w.download()
        w.upload(cells)
        assert L.fsg_world_set_cells(w.handle, full, MAX) == 0
        assert np.array_equal(w.download(), got)
