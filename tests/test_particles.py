"""Pins of the oracle's restatement of world::tickCells (world.cpp:2030-2189)
on hand-built cases: the landing spiral's order, landing in the start cell,
the spiral search, the throw-up when the spiral finds nothing, the order of
particles in one pass, lifetime, phase and leaving the world."""
import math

import numpy as np

import falling_sand_engine_amd as fsg
from oracle import pyoracle as O

AIR, STONE, SAND, WATER = fsg.MAT_AIR, fsg.MAT_STONE, fsg.MAT_SAND, fsg.MAT_WATER


def cell(m, t=320, rgba=0):
    return np.uint64(fsg.make_cell(m, t, rgba))


def world(W, H, m=AIR):
    return np.full((H, W), cell(m), np.uint64)


def test_spiral_order():
    s = O.spiral()
    # from (0,0) one step right, then around counter-clockwise in y-down cells
    assert s[:12] == [(0, 0), (1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1), (0, -1), (1, -1), (2, -1),
                      (2, 0), (2, 1)]
    # 1024 distinct offsets filling a 32x32 square, each ring after the one inside it
    assert len(s) == 1024 and len(set(s)) == 1024
    xs, ys = [a for a, _ in s], [b for _, b in s]
    assert (min(xs), max(xs), min(ys), max(ys)) == (-15, 16, -15, 16)
    ring = [max(abs(a - 0.5) + 0.5, abs(b - 0.5) + 0.5) for a, b in s]
    assert all(r1 <= r2 for r1, r2 in zip(ring, ring[1:]))


def test_falling_particle_lands_in_its_start_cell():
    """Gravity 0.1 from rest at (10.5, 2.5) over a stone floor at y = 20: the
    pass that carries it into the floor row writes its cell into the cell it
    started that pass in, which is air, and removes it."""
    cfg = O.default_config()
    g = world(32, 32)
    g[20:, :] = cell(STONE)
    tile = cell(SAND, 77, 0x55667788)
    ps = [fsg.make_particle(tile, 10.5, 2.5, 0.0, 0.0, 0.0, 0.1)]
    starts = []
    for _ in range(40):
        if not ps:
            break
        starts.append((int(ps[0].x), int(ps[0].y)))
        ps, landed = O.particles_tick(g, cfg, ps)
    assert not ps and landed == 1
    ys, xs = np.nonzero(g == tile)
    assert list(zip(xs, ys)) == [starts[-1]] and starts[-1][1] < 20
    # the free fall before it: y grows by the accumulated velocity each pass
    vy, y = np.float32(0), np.float32(2.5)
    for sx, sy in starts:
        assert int(y) == sy
        vy = np.float32(vy + np.float32(0.1))
        div = int(abs(vy) + 1)
        for _ in range(div):
            y = np.float32(y + np.float32(vy / np.float32(div)))


def test_spiral_search_when_the_start_cell_is_taken():
    """Inside stone moving into stone: the start cell is not air, so the cell
    goes to the first air cell of the spiral around the particle's position
    at the collision -- offset (-1, 1) (5th in spiral order) before (2, -1)
    (10th)."""
    cfg = O.default_config()
    g = world(40, 40, STONE)
    g[22, 19] = cell(AIR)  # (20,21) + (-1, 1)
    g[20, 22] = cell(AIR)  # (20,21) + (2, -1)
    tile = cell(WATER, 5, 1)
    ps, landed = O.particles_tick(g, cfg, [fsg.make_particle(tile, 20.5, 20.5, 0.0, 0.5)])
    assert not ps and landed == 1
    assert g[22, 19] == tile and g[20, 22] == cell(AIR)


def test_thrown_up_when_the_spiral_finds_nothing():
    cfg = O.default_config()
    g = world(64, 64, STONE)
    p = fsg.make_particle(cell(SAND), 30.5, 40.5, 0.0, 0.5)
    ps, landed = O.particles_tick(g, cfg, [p])
    assert landed == 0 and len(ps) == 1
    assert ps[0].vy == -4.0 and ps[0].y == np.float32(np.float32(40.5) + np.float32(0.5)) - 16


def test_later_particles_see_earlier_landings():
    """Two particles with the same start in one pass: the first lands in the
    start cell, the second finds it taken and goes to the spiral's first air
    cell around its collision point."""
    cfg = O.default_config()
    g = world(32, 32)
    g[11:, :] = cell(STONE)
    a, b = cell(SAND, 1, 1), cell(SAND, 2, 2)
    ps = [fsg.make_particle(a, 10.5, 10.5, 0.0, 0.6), fsg.make_particle(b, 10.5, 10.5, 0.0, 0.6)]
    ps, landed = O.particles_tick(g, cfg, ps)
    assert not ps and landed == 2
    assert g[10, 10] == a
    # b collides at (10.5, 11.1); spiral from there: (10,11) stone, (11,11) stone, (11,12) stone,
    # (10,12) stone, (9,12) stone, (9,11) stone, (9,10) air -> lands there
    assert g[10, 9] == b


def test_lifetime_phase_and_leaving_the_world():
    cfg = O.default_config()
    g = world(16, 16)
    g[8, :] = cell(STONE)
    T = fsg._abi.PARTICLE_TEMPORARY
    ps = [fsg.make_particle(cell(SAND), 4.5, 4.5, 0, 0, 0, 0, lifetime=2, flags=T),   # dies after 2 passes
          fsg.make_particle(cell(SAND), 5.5, 6.5, 0, 1.0, flags=fsg._abi.PARTICLE_PHASE),  # passes the stone row
          fsg.make_particle(cell(SAND), 15.5, 3.5, 2.0, 0.0),                             # leaves the world
          fsg.make_particle(cell(SAND), 2.5, 6.5, 0.0, 1.0, flags=T)]                     # temporary: dies on contact
    ps, landed = O.particles_tick(g, cfg, ps)
    assert landed == 0 and [p.lifetime for p in ps] == [1, 0]
    assert ps[1].y == np.float32(7.5) and (g == cell(AIR)).sum() == 16 * 15
    ps, _ = O.particles_tick(g, cfg, ps)
    ps, _ = O.particles_tick(g, cfg, ps)
    assert len(ps) == 1 and ps[0].flags == fsg._abi.PARTICLE_PHASE and ps[0].y == np.float32(9.5)


def test_target_force_pulls_toward_the_target():
    cfg = O.default_config()
    g = world(64, 64)
    p = fsg.make_particle(cell(SAND), 10.5, 10.5, target=(40.5, 50.5, 0.5))
    (q,), _ = O.particles_tick(g, cfg, [p])
    d = math.hypot(30.0, 40.0)  # 50: under 100, so the pull is damped by 0.95 (world.cpp:2046-2049)
    assert q.vx == np.float32(np.float32(np.float32(30.0) / np.float32(d) * np.float32(0.5)) * np.float32(0.95))
    assert q.vx > 0 and q.vy > 0 and q.x > p.x and q.y > p.y
