"""GPU tick vs the CPU oracle, bit-exact on every cell (det mode)."""
import numpy as np
import pytest

import falling_sand_engine_amd as fsg

pytestmark = pytest.mark.gpu


def run_gpu(ctx, cells, cfg, seed, ticks, tick0=0):
    H, W = cells.shape
    with fsg.World(ctx, W, H, seed, cfg) as w:
        if tick0:
            w.set_tick(tick0)
        w.upload(cells)
        w.tick(ticks)
        return w.download()


def first_diff(a, b):
    idx = np.argwhere(a != b)
    if idx.size == 0:
        return None
    y, x = idx[0]
    return f"{len(idx)} cells differ; first at (x={x}, y={y}): gpu={int(a[y, x]):#x} cpu={int(b[y, x]):#x}"


# up to 64 chunks: the quarter-chunk kernel; 640x576 (90 chunks): one
# 1024-thread workgroup per chunk; 1088x1088 (289 chunks, ids below 8): the
# 256-thread chunk kernel with its class tables
@pytest.mark.parametrize("W,H", [(8, 8), (64, 64), (70, 50), (130, 66), (128, 192), (65, 33), (201, 140),
                                 (640, 576), (1088, 1088)])
@pytest.mark.parametrize("seed", [0, 1, 7])
def test_small_worlds_random_mix(ctx, orc, W, H, seed):
    rng = np.random.default_rng(seed * 1000 + W * 7 + H)
    mats = rng.choice([0, 1, 2, 3, 4, 5, 6, 7], size=(H, W), p=[.35, .1, .2, .12, .08, .07, .05, .03])
    temps = rng.integers(0, 700 * 16, size=(H, W))
    moved = rng.integers(0, 2, size=(H, W)) * (mats == 2)
    cells = (mats.astype(np.uint64) | (moved.astype(np.uint64) << np.uint64(8)) |
             ((temps.astype(np.uint64) & np.uint64(0xFFFF)) << np.uint64(16)) |
             (rng.integers(0, 2**32, size=(H, W), dtype=np.uint64) << np.uint64(32)))
    cfg = orc.default_config()
    ref = orc.run(cells.copy(), cfg, seed, 0, 25)
    got = run_gpu(ctx, cells, cfg, seed, 25)
    assert first_diff(got, ref) is None, first_diff(got, ref)


# up to 64 chunks a world ticks as quarter chunks (4 workgroups per chunk),
# up to 256 chunks as one 1024-thread workgroup per chunk (640x640: 100)
@pytest.mark.parametrize("kind,W,H,seed,ticks", [(1, 512, 512, 1, 1000), (2, 256, 256, 2, 200),
                                                 (3, 512, 256, 3, 100), (1, 640, 640, 5, 200)])
def test_generated_worlds(ctx, orc, kind, W, H, seed, ticks):
    cells = fsg.generate(kind, W, H, seed)
    cfg = orc.default_config(reactions=0 if kind == 3 else 1)
    ref = orc.run(cells.copy(), cfg, seed, 0, ticks)
    got = run_gpu(ctx, cells, cfg, seed, ticks)
    assert first_diff(got, ref) is None, first_diff(got, ref)
    # sand/stone are never created or destroyed without reactions involving them
    assert np.count_nonzero(fsg.cell_mat(got) == 2) == np.count_nonzero(fsg.cell_mat(cells) == 2)


def test_last_round_as_quarters(ctx, orc):
    # 36 x 36 = 1296 chunks: more than a round of resident workgroups (five per
    # CU, 1280 on 256 CUs), so the chunks past the last full round tick as
    # quarters in the same launch
    W = H = 2304
    cells = fsg.generate(2, W, H, 21)
    cfg = orc.default_config()
    ref = orc.run(cells.copy(), cfg, 21, 0, 12)
    got = run_gpu(ctx, cells, cfg, 21, 12)
    assert first_diff(got, ref) is None, first_diff(got, ref)


def test_last_round_as_quarters_random_mix(ctx, orc):
    # every material, hot cells (burning wood, reactions), MOVED flags, in a
    # world whose last 16 of 1296 chunks tick as 256-thread quarters
    W = H = 2304
    rng = np.random.default_rng(77)
    mats = rng.choice([0, 1, 2, 3, 4, 5, 6, 7], size=(H, W), p=[.35, .1, .2, .12, .08, .07, .05, .03])
    temps = rng.integers(0, 700 * 16, size=(H, W))
    moved = rng.integers(0, 2, size=(H, W)) * (mats == 2)
    cells = (mats.astype(np.uint64) | (moved.astype(np.uint64) << np.uint64(8)) |
             ((temps.astype(np.uint64) & np.uint64(0xFFFF)) << np.uint64(16)) |
             (rng.integers(0, 2**32, size=(H, W), dtype=np.uint64) << np.uint64(32)))
    cfg = orc.default_config()
    ref = orc.run(cells.copy(), cfg, 77, 0, 10)
    got = run_gpu(ctx, cells, cfg, 77, 10)
    assert first_diff(got, ref) is None, first_diff(got, ref)


def test_resume_at_tick(ctx, orc):
    cells = fsg.generate(2, 192, 128, 5)
    cfg = orc.default_config(temp_every=4, temp_phase=2)
    ref = orc.run(cells.copy(), cfg, 5, 37, 23)
    got = run_gpu(ctx, cells, cfg, 5, 23, tick0=37)
    assert first_diff(got, ref) is None, first_diff(got, ref)


def test_cell_iter_1_and_2(ctx, orc):
    cells = fsg.generate(2, 128, 128, 9)
    for ci in (1, 2):
        cfg = orc.default_config(cell_iter=ci)
        ref = orc.run(cells.copy(), cfg, 9, 0, 40)
        got = run_gpu(ctx, cells, cfg, 9, 40)
        assert first_diff(got, ref) is None, (ci, first_diff(got, ref))


@pytest.mark.parametrize("kind,W,H,seed,ticks", [(1, 384, 320, 11, 600), (4, 640, 384, 12, 500)])
def test_chunk_skip_is_exact(ctx, orc, kind, W, H, seed, ticks):
    cells = fsg.generate(kind, W, H, seed)
    cfg_on = orc.default_config(chunk_skip=1)
    cfg_off = orc.default_config(chunk_skip=0)
    ref = orc.run(cells.copy(), cfg_off, seed, 0, ticks)
    with fsg.World(ctx, W, H, seed, cfg_on) as w:
        w.upload(cells)
        w.tick(ticks)
        got = w.download()
        st = w.stats()
    assert first_diff(got, ref) is None, first_diff(got, ref)
    print(f"processed {st.tiles_processed} of {st.tiles_offered} chunk updates")
    assert st.tiles_processed < st.tiles_offered  # something was actually skipped


def test_tick_temperature_standalone(ctx, orc):
    cells = fsg.generate(2, 200, 100, 4)
    cfg = orc.default_config(temp_every=0)
    ref = orc.run(cells.copy(), cfg, 4, 0, 3)
    orc.temperature(ref, cfg)
    ref = orc.run(ref, cfg, 4, 3, 2)
    with fsg.World(ctx, 200, 100, 4, cfg) as w:
        w.upload(cells)
        w.tick(3)
        w.tickTemperature()
        w.tick(2)
        got = w.download()
    assert first_diff(got, ref) is None, first_diff(got, ref)


def test_set_tile_and_get_tile(ctx, orc):
    cells = fsg.generate(1, 128, 128, 3)
    cfg = orc.default_config()
    ref = cells.copy()
    with fsg.World(ctx, 128, 128, 3, cfg) as w:
        w.upload(cells)
        w.setTile(5, 5, fsg.make_cell(2, 320, 0xFF00FFFF))
        w.setTile(5, 5, fsg.make_cell(3, 320, 0x80E06020))   # last edit wins
        w.setTile(-1, 3, fsg.make_cell(2))                    # ignored, like setTile
        w.tick(10)
        got = w.download()
        assert w.getTile(-5, 0) == 15
    ref[5, 5] = fsg.make_cell(3, 320, 0x80E06020)
    orc.run(ref, cfg, 3, 0, 10)
    assert first_diff(got, ref) is None, first_diff(got, ref)


def test_get_tile_mirror(ctx, orc):
    """getTile reads a host mirror: every cell equals the oracle's, a setTile
    is seen at once, and ticks, stamps and batched steps drop the mirror."""
    W, H = 200, 130  # pages clipped at the right and bottom edges
    cells = fsg.generate(2, W, H, 4)
    cfg = orc.default_config()
    ref = cells.copy()
    with fsg.World(ctx, W, H, 4, cfg) as w:
        w.upload(cells)
        w.tick(6)
        orc.run(ref, cfg, 4, 0, 6)
        assert all(w.getTile(x, y) == int(ref[y, x]) for y in range(H) for x in range(W))
        sand = fsg.make_cell(fsg.MAT_SAND, 320, 0x11223344)
        w.setTile(150, 100, sand)
        ref[100, 150] = sand
        assert w.getTile(150, 100) == sand
        fsg.batch_step(ctx, [w], 4)
        orc.run(ref, cfg, 4, 6, 4)
        w.prefetch(100, 50, 100, 80)
        assert all(w.getTile(x, y) == int(ref[y, x]) for y in range(50, H) for x in range(100, W, 3))
        w.tick(3)
        orc.run(ref, cfg, 4, 10, 3)
        assert all(w.getTile(x, y) == int(ref[y, x]) for y in range(0, H, 2) for x in range(W))


def test_stats_match_oracle(ctx, orc):
    cells = fsg.generate(2, 320, 192, 8)
    cfg = orc.default_config()
    ref = orc.run(cells.copy(), cfg, 8, 0, 30)
    with fsg.World(ctx, 320, 192, 8, cfg) as w:
        w.upload(cells)
        w.tick(30)
        st = w.stats()
    assert list(st.count) == [int(v) for v in orc.counts(ref)]
    assert st.checksum == orc.checksum(ref)


def test_cpp_facade_program(orc):
    """The C++ facade (include/fsg_world.hpp) driven like ME::world."""
    import json
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "build", "world_facade_test")
    assert os.path.exists(exe), "build it with falling_sand_engine_amd.build.build_cpp_tests()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    cells = fsg.generate(2, 256, 192, 7)
    for x in range(100, 110):
        cells[5, x] = fsg.make_cell(2, 320, 0xFF50C8E6)
    orc.run(cells, orc.default_config(), 7, 0, 10)
    assert got["checksum"] == orc.checksum(cells)
    assert got["wall"] == 15 and got["tile_50_50"] == int(cells[50, 50])
    assert got["files_ok"] == 1  # writeChunks / readChunks round trip over edited cells
    assert got["mirror_ok"] == 1  # getTile's host mirror against region(), setTile, tick
    assert got["counts"] == [int(v) for v in orc.counts(cells)[:8]]
    layers, _ = orc.render(cells, orc.default_config(), _cells_layer_params())
    assert got["texel_50_50"] == int(layers[1][50, 50])


def _cells_layer_params():
    from falling_sand_engine_amd._abi import LAYER_CELLS, RenderParams
    p = RenderParams()
    p.layers = LAYER_CELLS
    p.fire_mat = 255
    return p
