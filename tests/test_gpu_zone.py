"""The tick zone on the device (fsg_world_set_tick_zone) against the oracle's
zone-limited tick, with and without chunk skipping, and the listed skip path
(k_wake + k_tick_list) across tick-counter jumps and batched ticks."""
import numpy as np
import pytest

import falling_sand_engine_amd as fsg

pytestmark = pytest.mark.gpu

W, H = 320, 256
ZONES = [(64, 64, W - 128, H - 128),  # the game's zone: the world less a one-chunk frame
         (37, 21, 101, 77),           # odd edges inside chunks
         (0, 100, W, 3),              # a thin band
         (64, 64, 64, 64)]            # one chunk


@pytest.mark.parametrize("skip", [0, 1])
@pytest.mark.parametrize("zone", ZONES)
@pytest.mark.parametrize("kind", [1, 2])
def test_zone_matches_oracle(ctx, orc, kind, zone, skip):
    seed = 100 + kind
    cfg = orc.default_config(chunk_skip=skip)
    cells = fsg.generate(kind, W, H, seed)
    print(f"zone kind={kind} zone={zone} skip={skip}", flush=True)
    ref = orc.run(cells.copy(), cfg, seed, 0, 40, zone=zone)
    with fsg.World(ctx, W, H, seed, cfg) as w:
        w.upload(cells)
        w.tickZone = zone
        assert w.tickZone == zone
        w.tick(40)
        assert np.array_equal(w.download(), ref)


def test_zone_one_workgroup_per_chunk(ctx, orc):
    # W x H above keeps to the quarter-chunk kernel (20 chunks); 640x640 has
    # 100, one 1024-thread workgroup per chunk; the reference's zone
    W2 = H2 = 640
    seed = 9
    cfg = orc.default_config()
    cells = fsg.generate(2, W2, H2, seed)
    zone = (128, 128, W2 - 256, H2 - 256)
    ref = orc.run(cells.copy(), cfg, seed, 0, 40, zone=zone)
    with fsg.World(ctx, W2, H2, seed, cfg) as w:
        w.upload(cells)
        w.tickZone = zone
        w.tick(40)
        assert np.array_equal(w.download(), ref)


@pytest.mark.parametrize("skip", [0, 1])
def test_zone_changes_mid_run(ctx, orc, skip):
    seed = 7
    cfg = orc.default_config(chunk_skip=skip)
    cells = fsg.generate(2, W, H, seed)
    steps = [((64, 64, 64, 64), 15), ((30, 40, 250, 150), 15), ((0, 0, W, H), 10), ((100, 0, 64, H), 10)]
    ref = cells.copy()
    t = 0
    for z, n in steps:
        orc.run(ref, cfg, seed, t, n, zone=z)
        t += n
    with fsg.World(ctx, W, H, seed, cfg) as w:
        assert w.tickZone == (0, 0, W, H)
        w.upload(cells)
        for z, n in steps:
            w.tickZone = z
            w.tick(n)
        assert np.array_equal(w.download(), ref)


def test_zone_temperature_and_particles(ctx, orc):
    zone = (37, 21, 101, 77)
    cfg = orc.default_config()
    cells = fsg.generate(2, W, H, 3)
    ref = orc.temperature(cells.copy(), cfg, zone=zone)
    rng = np.random.default_rng(3)
    ps = [fsg.make_particle(fsg.make_cell(fsg.MAT_SAND, 320, i), float(rng.uniform(1, W - 1)),
                            float(rng.uniform(1, H - 1)), float(rng.uniform(-1, 1)), float(rng.uniform(-1, 2)),
                            0.0, 0.1, target=(50.0, 50.0, 0.2) if i % 7 == 0 else None) for i in range(400)]
    ref_ps = list(ps)
    with fsg.World(ctx, W, H, 3, cfg) as w:
        w.upload(cells)
        w.tickZone = zone
        w.tickTemperature()
        assert np.array_equal(w.download(), ref)
        w.particles_add(ps)
        for t in range(20):
            alive, landed = w.particles_tick()
            ref_ps, ref_landed = orc.particles_tick(ref, cfg, ref_ps, zone=zone)
            assert (alive, landed) == (len(ref_ps), ref_landed), t
        got = w.particles()
        assert [(p.x, p.y, p.vx, p.vy, p.lifetime) for p in got] == [(p.x, p.y, p.vx, p.vy, p.lifetime)
                                                                      for p in ref_ps]
        assert np.array_equal(w.download(), ref)


def test_listed_skip_path_across_jumps_and_batches(ctx, orc):
    """The awake list's counts are kept per tick parity: ticks run by the batch
    path, or a jump of the tick counter, in between listed ticks."""
    seed = 9
    cfg = orc.default_config(chunk_skip=1)
    cells = fsg.generate(2, W, H, seed)
    ref = cells.copy()
    orc.run(ref, cfg, seed, 0, 10)
    orc.run(ref, cfg, seed, 10, 5)
    orc.run(ref, cfg, seed, 15, 7)
    orc.run(ref, cfg, seed, 100, 6)   # after set_tick(100)
    orc.run(ref, cfg, seed, 106, 3)
    with fsg.World(ctx, W, H, seed, cfg) as w:
        w.upload(cells)
        w.tick(10)
        fsg.batch_step(ctx, [w], 5)
        w.tick(7)
        w.set_tick(100)
        w.tick(6)
        fsg.batch_step(ctx, [w], 3)
        assert np.array_equal(w.download(), ref)


def test_zone_arguments(ctx, orc):
    with fsg.World(ctx, 128, 64, 1, orc.default_config()) as w:
        for bad in [(-1, 0, 4, 4), (0, 0, 0, 4), (0, 0, 129, 4), (100, 8, 29, 1), (0, 60, 16, 5)]:
            with pytest.raises(fsg.FsgError):
                w.tickZone = bad
        assert w.tickZone == (0, 0, 128, 64)
        w.tickZone = (127, 63, 1, 1)
        assert w.tickZone == (127, 63, 1, 1)
