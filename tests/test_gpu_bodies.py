"""Rigid-body stamp/unstamp on the GPU vs the oracle's claim rounds."""
import math

import numpy as np
import pytest

import falling_sand_engine_amd as fsg
from bodies import random_bodies

pytestmark = pytest.mark.gpu


# stone/wood bodies, and bodies with SAND and SOUP typed pixels (a stamped
# SAND/SOUP pixel leaves its cell a stamping target; later rounds must not
# land on it)
SOLID_MATS, MIXED_MATS = (1, 7), (1, 7, 2, 3)


@pytest.mark.parametrize("seed,nb,skip,W,H,max_side,mats", [
    (1, 30, 0, 320, 256, 16, SOLID_MATS), (2, 120, 0, 320, 256, 16, SOLID_MATS),
    (3, 60, 1, 320, 256, 16, SOLID_MATS), (4, 160, 1, 320, 256, 16, MIXED_MATS),
    # over 2^20 body pixels
    (5, 1100, 1, 1024, 1024, 64, SOLID_MATS), (6, 1100, 0, 1024, 1024, 64, MIXED_MATS)])
def test_stamp_tick_unstamp_matches_oracle(ctx, orc, seed, nb, skip, W, H, max_side, mats):
    rng = np.random.default_rng(seed)
    cells = fsg.generate(1, W, H, seed)
    bodies, pixels = random_bodies(rng, W, H, nb, max_side=max_side, mats=mats)
    assert (pixels.size > 1 << 20) == (max_side == 64)
    print(f"bodies seed={seed} nb={nb} skip={skip} {W}x{H} pixels={pixels.size} mats={mats}", flush=True)
    cfg = orc.default_config(chunk_skip=skip)
    # oracle: stamp, 3 ticks, unstamp
    ref = cells.copy()
    rs = orc.stamp(ref, cfg, bodies, pixels, displaced_cap=100000)
    orc.run(ref, cfg, seed, 0, 3)
    ref_px = orc.unstamp(ref, cfg, bodies, pixels)
    with fsg.World(ctx, W, H, seed, cfg) as w:
        w.upload(cells)
        gs = w.stamp_bodies(bodies, pixels, displaced_cap=100000)
        after_stamp = w.download()
        w.tick(3)
        got_px = w.unstamp_bodies(pixels.size)
        got = w.download()
    cpu_after_stamp = cells.copy()
    orc.stamp(cpu_after_stamp, cfg, bodies, pixels)
    assert np.array_equal(after_stamp, cpu_after_stamp)
    assert gs["placed"] == rs["placed"] and gs["displaced"] == rs["displaced"]
    assert gs["displaced_cells"] == rs["displaced_cells"]
    assert gs["hits_sand"] == rs["hits_sand"] and gs["hits_soup"] == rs["hits_soup"]
    assert np.array_equal(got, ref), int(np.count_nonzero(got != ref))
    assert np.array_equal(got_px, ref_px)


def test_round_trip_without_tick_is_identity(ctx, orc):
    # Non-overlapping bodies in open air, quarter-turn poses: stamp then
    # unstamp returns every pixel unchanged and leaves only air behind; the
    # vacated cells keep the temperature the pixel carried.
    W, H = 256, 256
    rng = np.random.default_rng(9)
    cells = np.full((H, W), fsg.make_cell(0, 320), dtype=np.uint64)
    bodies, pixels = [], []
    off = 0
    for i in range(20):
        mw, mh = int(rng.integers(4, 17)), int(rng.integers(4, 17))
        ang = [0.0, math.pi / 2, math.pi, -math.pi / 2][i % 4]
        bodies.append(fsg.body_from_pose(24 + (i % 5) * 45, 24 + (i // 5) * 45, ang, mw, mh, off))
        m = rng.choice([1, 7], size=mh * mw)
        t = rng.integers(0, 4000, size=mh * mw)
        pixels.append(m.astype(np.uint64) | (t.astype(np.uint64) << np.uint64(16)))
        off += mw * mh
    pixels = np.concatenate(pixels)
    with fsg.World(ctx, W, H, 0, orc.default_config()) as w:
        w.upload(cells)
        res = w.stamp_bodies(bodies, pixels)
        assert res["placed"] == pixels.size and res["unplaced"] == 0
        px = w.unstamp_bodies(pixels.size)
        grid = w.download()
    assert np.array_equal(px, pixels)
    assert (fsg.cell_mat(grid) == 0).all() and (fsg.cell_rgba(grid) == 0).all()
    ref = cells.copy()
    orc.stamp(ref, orc.default_config(), bodies, pixels)
    ref_px = orc.unstamp(ref, orc.default_config(), bodies, pixels)
    assert np.array_equal(grid, ref) and np.array_equal(ref_px, pixels)
    assert sorted(fsg.cell_temp(grid)[fsg.cell_temp(grid) != 320].tolist()) == \
        sorted(t for t in fsg.cell_temp(pixels).tolist() if t != 320)


def test_stamp_state_errors(ctx, orc):
    with fsg.World(ctx, 64, 64, 0) as w:
        with pytest.raises(fsg.FsgError) as e:
            w.unstamp_bodies(0)
        assert e.value.status == -5
        b = fsg.body_from_pose(10, 10, 0, 4, 4, 0)
        with pytest.raises(fsg.FsgError):   # pixel array too short for the body
            w.stamp_bodies([b], np.zeros(3, dtype=np.uint64))


def test_tick_after_unstamp_and_census(ctx, orc):
    """The reference's order, stamp -> tick -> unstamp -> texture pass
    (game.cpp:1760-2067), then more ticks with chunk skipping: the grid
    equals the oracle's, and the census compares with the state before the
    tick (the stamp and unstamp put the claim slots they borrow from the
    spare buffer back)."""
    W, H, seed = 384, 256, 12
    rng = np.random.default_rng(seed)
    cells = fsg.generate(2, W, H, seed)
    bodies, pixels = random_bodies(rng, W, H, 70, max_side=16, mats=MIXED_MATS)
    cfg = orc.default_config(chunk_skip=1)
    print(f"stamp/tick/unstamp/render {W}x{H} pixels={pixels.size}", flush=True)
    ref = cells.copy()
    orc.stamp(ref, cfg, bodies, pixels)
    pre_tick = ref.copy()
    orc.run(ref, cfg, seed, 0, 1)
    ref_px = orc.unstamp(ref, cfg, bodies, pixels)
    _layers, ref_st = orc.render(ref, cfg, _render_params(), prev=pre_tick)
    after = orc.run(ref.copy(), cfg, seed, 1, 20)
    with fsg.World(ctx, W, H, seed, cfg) as w:
        w.upload(cells)
        w.stamp_bodies(bodies, pixels)
        w.tick(1)
        got_px = w.unstamp_bodies(pixels.size)
        st = w.render(fsg.LAYER_CELLS, stats=True)
        assert np.array_equal(w.download(), ref)
        w.tick(20)
        got_after = w.download()
    assert np.array_equal(got_px, ref_px)
    assert st["moved"] == [int(v) for v in ref_st.moved] and sum(st["moved"]) > 0
    assert np.array_equal(got_after, after), int(np.count_nonzero(got_after != after))


def _render_params():
    from falling_sand_engine_amd._abi import RenderParams
    p = RenderParams()
    p.layers = fsg.LAYER_CELLS
    p.fire_mat = 255
    p.count_moved = 1
    for m, e in enumerate(fsg.EMIT_DEFAULT):
        p.emit[m] = e
    return p
