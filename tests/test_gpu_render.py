"""Device texture pass (fsg_world_render) vs the CPU oracle, bit-exact on
every texel and on the movingTiles census."""
import ctypes as C

import numpy as np
import pytest

import falling_sand_engine_amd as fsg
from falling_sand_engine_amd import _abi
from falling_sand_engine_amd._abi import LAYER_CELLS, LAYER_EMISSION, LAYER_FIRE, RenderParams

pytestmark = pytest.mark.gpu

ALL = LAYER_CELLS | LAYER_EMISSION | LAYER_FIRE


def params(layers=ALL, rect=None, fire=255, count=1):
    p = RenderParams()
    p.layers = layers
    p.x0, p.y0, p.w, p.h = rect if rect is not None else (0, 0, 0, 0)
    p.fire_mat = fire
    p.count_moved = count
    for m, e in enumerate(fsg.EMIT_DEFAULT):
        p.emit[m] = e
    return p


def gpu_layers(w, layers=(LAYER_CELLS, LAYER_EMISSION, LAYER_FIRE)):
    return {b: w.layer(b) for b in layers}


def assert_layers_equal(got, ref, what):
    for b, a in got.items():
        if not np.array_equal(a, ref[b]):
            idx = np.argwhere(a != ref[b])
            y, x = idx[0]
            raise AssertionError(f"{what}: layer {b}: {len(idx)} texels differ; first (x={x}, y={y}) "
                                 f"gpu={int(a[y, x]):#010x} cpu={int(ref[b][y, x]):#010x}")


@pytest.mark.parametrize("W,H", [(64, 64), (130, 66), (65, 33), (257, 190), (4, 3)])
@pytest.mark.parametrize("kind", [3, 4])
def test_render_after_ticks_matches_oracle(ctx, orc, W, H, kind):
    cfg = orc.default_config()
    cells = fsg.generate(kind, W, H, W + H)
    fire = fsg.MAT_SAND
    print(f"render W={W} H={H} kind={kind}", flush=True)
    with fsg.World(ctx, W, H, 5, cfg) as w:
        w.upload(cells)
        w.tick(7)
        prev = w.download()
        w.tick(1)
        cur = w.download()
        st = w.render(ALL, fire_mat=fire, stats=True)
        got = gpu_layers(w)
    ref, rst = orc.render(cur, cfg, params(fire=fire), prev=prev)
    assert_layers_equal(got, ref, f"{W}x{H} kind {kind}")
    assert st["moved"] == [int(v) for v in rst.moved]
    assert st["fire_cells"] == rst.fire_cells and st["had_dirty"] == bool(rst.had_dirty)
    if W * H >= 1024:
        assert sum(st["moved"]) > 0  # the census saw the tick's moves


def test_rectangles_update_only_their_texels_and_the_fire_layer_keeps(ctx, orc):
    cfg = orc.default_config()
    W, H = 203, 131
    cells = fsg.generate(4, W, H, 9)
    rects = [(0, 0, 0, 0), (-7, 5, 40, 30), (13, 2, 51, 77), (190, 120, 50, 50), (1, 1, 3, 1),
             (64, 64, 64, 64), (300, 10, 5, 5), (5, 5, 1, 200)]
    ref = {b: np.zeros((H, W), dtype=np.uint32) for b in (1, 2, 4)}
    with fsg.World(ctx, W, H, 3, cfg) as w:
        w.upload(cells)
        for i, rect in enumerate(rects):
            w.tick(3)
            cur = w.download()
            fire = (fsg.MAT_SAND, fsg.MAT_WATER)[i % 2]  # alternate, so earlier fire texels are kept
            print(f"rect {rect} fire={fire}", flush=True)
            w.render(ALL, rect=rect, fire_mat=fire)
            orc.render(cur, cfg, params(rect=rect, fire=fire), layers=ref)
            assert_layers_equal(gpu_layers(w), ref, f"after rect {rect}")


def test_layers_subset_and_census_edge_cases(ctx, orc):
    cfg = orc.default_config()
    W, H = 96, 80
    cells = fsg.generate(3, W, H, 2)
    with fsg.World(ctx, W, H, 1, cfg) as w:
        w.upload(cells)
        # before any tick: no previous state, the census is empty
        st = w.render(LAYER_EMISSION, stats=True)
        assert st == {"moved": [0] * 16, "fire_cells": 0, "had_dirty": False}
        assert w.layer_ptr(LAYER_CELLS) == 0 and w.layer_ptr(LAYER_EMISSION) != 0
        lib = fsg.lib()
        tex = np.empty((H, W), dtype=np.uint32)
        assert lib.fsg_world_layer_download(w.handle, LAYER_CELLS, 0, 0, W, H, tex.ctypes.data) == _abi.FSG_ERR_STATE
        assert lib.fsg_world_layer_download(w.handle, LAYER_EMISSION, 0, 0, W + 1, H,
                                            tex.ctypes.data) == _abi.FSG_ERR_INVALID
        ref, _ = orc.render(cells, cfg, params(layers=LAYER_EMISSION))
        assert np.array_equal(w.layer(LAYER_EMISSION), ref[LAYER_EMISSION])
        # the census covers the last tick and the setTile edits since
        prev = w.download()
        w.tick(1)
        w.setTile(3, 3, fsg.make_cell(fsg.MAT_LAVA, 320, 0x01020304))
        w.setTile(4, 3, fsg.make_cell(fsg.MAT_LAVA, 320, 0x01020305))
        st = w.render(LAYER_CELLS, stats=True)
        cur = w.download()
        _, rst = orc.render(cur, cfg, params(layers=LAYER_CELLS), prev=prev)
        assert st["moved"] == [int(v) for v in rst.moved] and st["moved"][fsg.MAT_LAVA] >= 2
        # count_moved off: no census even with stats
        st = w.render(LAYER_CELLS, stats=True, count_moved=False)
        assert sum(st["moved"]) == 0
        bad = params(layers=8)
        assert lib.fsg_world_render(w.handle, C.byref(bad), None) == _abi.FSG_ERR_INVALID


def test_layers_count_against_the_budget(orc):
    cfg = orc.default_config()
    W = H = 256
    # the world's device bytes: two states, chunk flags and awake list, counts, stats, class tables
    world_bytes = 2 * W * H * 8 + 3 * 16 * 4 + 16 + 18 * 8 + 3 * 512 * 4
    with fsg.Context(0, hbm_budget=world_bytes + W * H * 4 + 1024) as c, fsg.World(c, W, H, 1, cfg) as w:
        w.upload(fsg.generate(1, W, H, 1))
        w.render(LAYER_CELLS)  # fits
        with pytest.raises(fsg.FsgError):
            w.render(LAYER_EMISSION)  # a second layer does not
        used = c.usage()
        assert used <= world_bytes + W * H * 4 + 1024


def test_render_of_a_batched_world_at_full_size(ctx, orc):
    # 4096^2 striped world stepped in a batch; one full render of all layers
    cfg = orc.default_config(reactions=0)
    W = H = 4096
    cells = fsg.generate(3, W, H, 3)
    with fsg.World(ctx, W, H, 3, cfg) as w:
        w.upload(cells)
        fsg.batch_step(ctx, [w], 4)
        prev = w.download()
        fsg.batch_step(ctx, [w], 1)
        cur = w.download()
        st = w.render(ALL, fire_mat=fsg.MAT_WATER, stats=True)
        got = gpu_layers(w)
    ref, rst = orc.render(cur, cfg, params(fire=fsg.MAT_WATER), prev=prev)
    assert_layers_equal(got, ref, "4096^2")
    assert st["moved"] == [int(v) for v in rst.moved]
