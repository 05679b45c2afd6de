"""Device island check (fsg_world_physics_check) vs the CPU oracle: statuses,
counts, boxes, body images and the grid afterwards, bit-exact."""
import ctypes as C

import numpy as np
import pytest

import falling_sand_engine_amd as fsg
from falling_sand_engine_amd import _abi
from test_islands import random_fragments

pytestmark = pytest.mark.gpu


def compare(got, ref):
    assert len(got) == len(ref)
    for i, (g, r) in enumerate(zip(got, ref)):
        assert (g["status"], g["count"], g["box"]) == (r["status"], r["count"], r["box"]), (i, g, r)
        if r["status"] == _abi.ISLAND_BODY:
            assert np.array_equal(g["pixels"], r["pixels"]), i


@pytest.mark.parametrize("W,H,p,nseeds", [(64, 64, 0.5, 50), (200, 130, 0.55, 400), (131, 77, 0.6, 300),
                                          (1024, 1024, 0.56, 2000)])
def test_islands_match_oracle(ctx, orc, W, H, p, nseeds):
    cfg = orc.default_config()
    c = random_fragments(W, H, W + H, p=p)
    rng = np.random.default_rng(W * H)
    seeds = [(int(x), int(y)) for x, y in zip(rng.integers(-3, W + 3, nseeds), rng.integers(-3, H + 3, nseeds))]
    seeds += seeds[:20]  # repeated seeds: found removed or attached already
    print(f"islands W={W} H={H} p={p} seeds={len(seeds)}", flush=True)
    with fsg.World(ctx, W, H, 1, cfg) as w:
        w.upload(c)
        got = w.physicsCheck(seeds)
        grid = w.download()
    ref = orc.physics_check(c, cfg, seeds)
    compare(got, ref)
    assert np.array_equal(grid, c)
    assert any(d["status"] == _abi.ISLAND_BODY for d in ref)


def test_pixel_copy_and_argument_checks(ctx, orc):
    cfg = orc.default_config()
    W, H = 96, 64
    c = random_fragments(W, H, 5, p=0.5)
    seeds = [(x, y) for y in range(0, H, 7) for x in range(0, W, 5)]
    arr = (_abi.Island * len(seeds))()
    for i, (x, y) in enumerate(seeds):
        arr[i].x, arr[i].y = x, y
    need = C.c_size_t()
    lib = fsg.lib()
    with fsg.World(ctx, W, H, 1, cfg) as w:
        w.upload(c)
        assert lib.fsg_world_physics_check(w.handle, arr, len(seeds), C.byref(need)) == _abi.FSG_OK
        assert need.value > 0
        small = np.empty(need.value - 1, dtype=np.uint64)
        assert lib.fsg_world_island_pixels(w.handle, small.ctypes.data, small.size) == _abi.FSG_ERR_INVALID
        pix = np.empty(need.value, dtype=np.uint64)
        assert lib.fsg_world_island_pixels(w.handle, pix.ctypes.data, pix.size) == _abi.FSG_OK
        ref = orc.physics_check(c, cfg, seeds)
        for i, r in enumerate(ref):
            assert (arr[i].status, arr[i].count) == (r["status"], r["count"])
            if r["status"] == _abi.ISLAND_BODY:
                o, n = arr[i].pixel_offset, arr[i].w * arr[i].h
                assert np.array_equal(pix[o:o + n].reshape(arr[i].h, arr[i].w), r["pixels"])
        assert np.array_equal(w.download(), c)
        assert lib.fsg_world_physics_check(w.handle, None, 3, None) == _abi.FSG_ERR_INVALID
        assert lib.fsg_world_physics_check(w.handle, arr, 0, None) == _abi.FSG_OK
        # a check with no bodies leaves no images to copy
        assert lib.fsg_world_island_pixels(w.handle, None, 0) == _abi.FSG_OK


@pytest.mark.parametrize("skip", [0, 1])
def test_ticks_after_a_check_stay_bit_exact(ctx, orc, skip):
    # removal happens outside a tick: with chunk skipping every chunk must
    # run at the next tick
    cfg = orc.default_config(chunk_skip=skip)
    W, H = 192, 160
    c = fsg.generate(4, W, H, 3)
    c[20:40, 30:60] = fsg.make_cell(fsg.MAT_STONE, 320, 0xFF808080)  # a floating 600-cell slab
    c[5:8, 100:103] = fsg.make_cell(fsg.MAT_STONE, 320, 0xFF808080)  # and a 9-cell one
    with fsg.World(ctx, W, H, 7, cfg) as w:
        w.upload(c)
        w.tick(30)
        c_ref = w.download()
        got = w.physicsCheck([(45, 30), (101, 6), (0, H - 1)])
        w.tick(30)
        grid = w.download()
    ref = orc.physics_check(c_ref, cfg, [(45, 30), (101, 6), (0, H - 1)])
    compare(got, ref)
    assert got[0]["status"] == _abi.ISLAND_BODY and got[1]["status"] == _abi.ISLAND_REMOVED
    orc.run(c_ref, cfg, 7, 30, 30)
    assert np.array_equal(grid, c_ref)
