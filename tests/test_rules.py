"""Oracle rule pinning: closed-form micro-worlds, conservation, temperature
closed forms, the dirty-rect predicate's soundness and an independent
restatement of the phase schedule (shuffled block order)."""
import numpy as np
import pytest

from worlds import AIR, LAVA, OIL, SAND, STEAM, STONE, WATER, WOOD, cell, mats, parse, temps


def box(w, h):
    g = np.full((h, w), cell(AIR), dtype=np.uint64)
    g[-1, :] = cell(STONE)
    g[:, 0] = cell(STONE)
    g[:, -1] = cell(STONE)
    return g


# --- single-block rules --------------------------------------------------------

def test_block_sand_falls_into_air(orc):
    cfg = orc.default_config()
    # sand over air, other column stone: the diagonal is blocked, so it always falls
    for r in range(64):
        ch, q = orc.block_rule(cfg, 0, r, 0, [cell(SAND), cell(STONE), cell(AIR), cell(STONE)])
        assert ch and (q[0] & 0xFF) == AIR and (q[2] & 0xFF) == SAND
        assert bool(q[2] & 0x100) == bool((r >> 8) & 1)   # MOVED set with the r[8] coin


def test_block_sand_skips_fall_one_in_sixteen(orc):
    cfg = orc.default_config()
    fell = 0
    for r in range(16):
        _, q = orc.block_rule(cfg, 0, r, 0, [cell(SAND), cell(STONE), cell(AIR), cell(AIR)])
        fell += (q[2] & 0xFF) == SAND
    assert fell == 15  # rand() % 20 == 0 -> 1/16 skip when a diagonal is free


def test_block_sand_iteration_gate(orc):
    cfg = orc.default_config()  # sand iterations = 2
    ch, q = orc.block_rule(cfg, 2, 0xFFFF, 0, [cell(SAND), cell(STONE), cell(AIR), cell(STONE)])
    assert not ch


def test_block_density_order(orc):
    cfg = orc.default_config()
    pairs = [(SAND, WATER, True), (WATER, OIL, True), (OIL, WATER, False), (LAVA, WATER, True),
             (WATER, STEAM, True), (OIL, AIR, True), (SAND, STONE, False), (WATER, LAVA, False)]
    for top, bottom, sinks in pairs:
        _, q = orc.block_rule(cfg, 0, 0xFF00, 0, [cell(top), cell(STONE), cell(bottom), cell(STONE)])
        assert ((q[2] & 0xFF) == top) == sinks, (top, bottom)


def test_block_gas_rises_only_into_air(orc):
    cfg = orc.default_config()
    _, q = orc.block_rule(cfg, 0, 0, 0, [cell(AIR), cell(STONE), cell(STEAM), cell(STONE)])
    assert (q[0] & 0xFF) == STEAM
    ch, q = orc.block_rule(cfg, 0, 0, 0, [cell(WATER, 0), cell(STONE), cell(STEAM), cell(STONE)])
    # water sinks into steam (steam rises through it) - the swap is the water's move
    assert (q[0] & 0xFF) == STEAM and (q[2] & 0xFF) == WATER


def test_block_water_over_lava_reaction(orc):
    cfg = orc.default_config()
    ch, q = orc.block_rule(cfg, 0, 0, 0, [cell(WATER, 111), cell(AIR), cell(LAVA, 222), cell(LAVA, 333)])
    assert ch
    assert [v & 0xFF for v in q] == [STEAM, AIR, STONE, STONE]
    assert [(v >> 16) & 0xFFFF for v in q] == [111, 320, 222, 333]   # products keep temperature
    assert (q[0] >> 32) == cfg.materials[STEAM].rgba and (q[2] >> 32) == cfg.materials[STONE].rgba
    cfg0 = orc.default_config(reactions=0)
    _, q = orc.block_rule(cfg0, 0, 0, 0, [cell(WATER), cell(AIR), cell(LAVA), cell(LAVA)])
    assert [v & 0xFF for v in q] != [STEAM, AIR, STONE, STONE]


def test_block_wood_burns_to_steam_and_air(orc):
    cfg = orc.default_config()
    hot = 300 * 16 + 1
    ch, q = orc.block_rule(cfg, 0, 0, 0, [cell(WOOD, hot), cell(AIR, 5), cell(STONE), cell(AIR)])
    assert [v & 0xFF for v in q] == [AIR, STEAM, STONE, AIR]
    assert ((q[1] >> 16) & 0xFFFF) == hot and ((q[0] >> 16) & 0xFFFF) == 5
    ch, q = orc.block_rule(cfg, 0, 0, 0, [cell(WOOD, 300 * 16), cell(AIR), cell(STONE), cell(AIR)])
    assert not ch  # exactly at the threshold: no burn ("above")
    ch, q = orc.block_rule(cfg, 0, 0, 0, [cell(WOOD, hot), cell(STONE), cell(STONE), cell(STONE)])
    assert [v & 0xFF for v in q] == [STEAM, STONE, STONE, STONE]  # no air: becomes steam in place


def test_block_sand_slide_needs_moved(orc):
    cfg = orc.default_config()
    base = [cell(SAND), cell(STONE), cell(STONE), cell(AIR)]
    ch, q = orc.block_rule(cfg, 0, 0x00010000, 0xFFFFFFFF, base)
    assert not ch  # stuck (MOVED clear) and r[16:19] != 0: no slide
    ch, q = orc.block_rule(cfg, 0, 0x00F00000, 0xFFFFFFFF, base)
    assert [v & 0xFF for v in q] == [AIR, STONE, STONE, SAND]  # r[16:19] == 0: unstuck, slides
    assert q[3] & 0x100  # and moves on with MOVED set
    ch, q = orc.block_rule(cfg, 0, 0, 0xFFF007FF, base)  # s[11:19] == 0 still clears MOVED after the slide
    assert [v & 0xFF for v in q] == [AIR, STONE, STONE, SAND] and not (q[3] & 0x100)
    moving = [cell(SAND, moved=True), cell(STONE), cell(STONE), cell(AIR)]
    ch, q = orc.block_rule(cfg, 0, 0x00010000, 0xFFFFFFFF, moving)
    assert [v & 0xFF for v in q] == [AIR, STONE, STONE, SAND]
    ch, q = orc.block_rule(cfg, 0, 0x00010000, 0xFFFFFFE0, moving)    # s[0:4] == 0: refuses, stops
    assert [v & 0xFF for v in q] == [SAND, STONE, STONE, AIR] and not (q[0] & 0x100)
    # the right column draws from r[20:23]
    right = [cell(STONE), cell(SAND), cell(AIR), cell(STONE)]
    ch, q = orc.block_rule(cfg, 0, 0x00100000, 0xFFFFFFFF, right)
    assert not ch
    ch, q = orc.block_rule(cfg, 0, 0x000F0000, 0xFFFFFFFF, right)
    assert [v & 0xFF for v in q] == [STONE, AIR, SAND, STONE] and q[2] & 0x100


def test_block_slide_three_cycle(orc):
    cfg = orc.default_config()
    # slider TL, side TR air, target BR water: water goes up beside, TL becomes the air
    q0 = [cell(SAND, 1, moved=True), cell(AIR, 2), cell(STONE, 3), cell(WATER, 4)]
    _, q = orc.block_rule(cfg, 0, 0, 0xFFFFFFFF, q0)
    assert [v & 0xFF for v in q] == [AIR, WATER, STONE, SAND]
    assert [(v >> 16) & 0xFFFF for v in q] == [2, 4, 3, 1]


def test_block_liquid_flows_along_its_heading(orc):
    """A SOUP cell's MOVED bit is its heading (set: right).  Facing the air
    beside it, it flows whatever the draw; facing away it turns only when
    its row's 4 turn bits (r[24:27] top, r[28:31] bottom) are all zero, and
    then carries the new heading."""
    cfg = orc.default_config()
    NOTURN = 0x11000000  # one bit set in each row's turn field
    # bottom row: water left, air right
    right = [cell(STONE), cell(STONE), cell(WATER, moved=True), cell(AIR)]
    _, q = orc.block_rule(cfg, 0, NOTURN, 0, right)
    assert [v & 0xFF for v in q] == [STONE, STONE, AIR, WATER] and q[3] & 0x100
    left = [cell(STONE), cell(STONE), cell(WATER), cell(AIR)]
    ch, q = orc.block_rule(cfg, 0, NOTURN, 0xFFFFFFFF, left)
    assert not ch  # heading left, away from the air: stays
    _, q = orc.block_rule(cfg, 0, 0x01000000, 0, left)  # bottom row's turn bits zero
    assert [v & 0xFF for v in q] == [STONE, STONE, AIR, WATER] and q[3] & 0x100  # turned right and moved
    # top row: air left, water right heading left: flows; its heading stays left
    top = [cell(AIR), cell(WATER), cell(STONE), cell(STONE)]
    _, q = orc.block_rule(cfg, 0, NOTURN, 0, top)
    assert [v & 0xFF for v in q] == [WATER, AIR, STONE, STONE] and not q[0] & 0x100
    # gas still spreads with the s[23]/s[24] coins
    gas = [cell(STONE), cell(STONE), cell(STEAM), cell(AIR)]
    moved = [orc.block_rule(cfg, 0, 0, s << 24, gas)[1][3] & 0xFF for s in (0, 1)]
    assert moved == [AIR, STEAM]


def test_dam_break_front_advances(orc):
    """A wide column of water released against the left wall of an empty
    basin: the front carries its heading across the floor, so within 300
    ticks water reaches beyond the middle of the basin; a fair coin per pair
    (a random walk of the front) would still be near the dam."""
    cfg = orc.default_config(temp_every=0)
    W, H = 192, 96
    g = box(W, H)
    g[20:H - 1, 1:40] = cell(WATER)
    orc.run(g, cfg, 1, 0, 300)
    cols = np.nonzero((mats(g) == WATER).any(axis=0))[0]
    assert cols.max() > W // 2


# --- small worlds ---------------------------------------------------------------

def test_grain_falls_to_floor(orc):
    cfg = orc.default_config(temp_every=0)
    g = box(16, 40)
    g[0, 8] = cell(SAND)
    orc.run(g, cfg, 1, 0, 200)
    m = mats(g)
    assert (m == SAND).sum() == 1
    ys, xs = np.nonzero(m == SAND)
    assert ys[0] == 38  # resting on the floor (row 39 is stone)


def test_sand_heap_is_supported(orc):
    cfg = orc.default_config(temp_every=0)
    g = box(48, 40)
    g[2:22, 20:24] = cell(SAND)
    orc.run(g, cfg, 2, 0, 600)
    m = mats(g)
    assert (m == SAND).sum() == 80
    # settled: no sand above a lighter non-solid cell
    below = m[1:, :]
    assert not np.any((m[:-1, :] == SAND) & ((below == AIR) | (below == WATER)))
    # a heap, not a column: it spread to more than the 4 source columns
    assert len(set(np.nonzero(m == SAND)[1])) > 4


def test_water_levels_out(orc):
    cfg = orc.default_config(temp_every=0)
    g = box(24, 30)
    g[5:25, 1:5] = cell(WATER)  # a 4-wide column of water in one corner
    orc.run(g, cfg, 3, 0, 3000)
    m = mats(g)
    assert (m == WATER).sum() == 80
    heights = (m[:, 1:-1] == WATER).sum(axis=0)
    assert heights.max() - heights.min() <= 1, heights


def test_oil_floats_on_water(orc):
    cfg = orc.default_config(temp_every=0)
    g = box(12, 30)
    g[10:15, 1:-1] = cell(WATER)
    g[15:20, 1:-1] = cell(OIL)
    orc.run(g, cfg, 4, 0, 600)
    m = mats(g)
    ys_oil = np.nonzero(m == OIL)[0]
    ys_water = np.nonzero(m == WATER)[0]
    assert ys_oil.max() < ys_water.min()


def test_steam_rises_to_ceiling(orc):
    cfg = orc.default_config(temp_every=0)
    g = box(10, 30)
    g[0, :] = cell(STONE)
    g[20:28, 1:9] = cell(STEAM)  # 64 steam cells
    orc.run(g, cfg, 5, 0, 800)
    m = mats(g)
    assert (m == STEAM).sum() == 64
    assert (m[1:9, 1:9] == STEAM).all()


def test_conservation_with_reactions(orc):
    import falling_sand_engine_amd as fsg
    cells = fsg.generate(2, 256, 192, 12)
    c0 = orc.counts(cells)
    g = cells.copy()
    cfg = orc.default_config()
    orc.run(g, cfg, 12, 0, 150)
    c1 = orc.counts(g)
    d = c1.astype(np.int64) - c0.astype(np.int64)
    assert d[AIR] == 0 and d[SAND] == 0 and d[OIL] == 0
    assert d[LAVA] + d[STONE] == 0       # stone only comes from lava
    assert d[WATER] + d[STEAM] + d[WOOD] == 0   # steam only from water and wood
    assert d[LAVA] < 0                    # the reaction did fire


def test_conservation_without_reactions(orc):
    import falling_sand_engine_amd as fsg
    cells = fsg.generate(3, 256, 128, 3)
    g = cells.copy()
    orc.run(g, orc.default_config(reactions=0), 3, 0, 100)
    assert (orc.counts(g) == orc.counts(cells)).all()


# --- temperature ------------------------------------------------------------------

def test_temperature_uniform_is_fixed_point(orc):
    g = parse("\n".join(["#s.w~lWo"] * 8), temp=1234)
    before = g.copy()
    orc.temperature(g, orc.default_config())
    assert (g == before).all()


@pytest.mark.parametrize("D", [800, 7, -7, -800, 1])
def test_temperature_spike_closed_form(orc, D):
    T0 = 320
    g = parse("\n".join(["." * 9] * 9), temp=T0)
    g[4, 4] = cell(AIR, T0 + D)
    orc.temperature(g, orc.default_config())
    t = temps(g)
    assert t[4, 4] == T0 + D + ((-4 * D) >> 3)
    for (y, x) in [(3, 4), (5, 4), (4, 3), (4, 5)]:
        assert t[y, x] == T0 + (D >> 3)
    assert t[3, 3] == T0 and t[0, 0] == T0


def test_temperature_edge_neighbour_is_self(orc):
    g = parse("\n".join(["." * 5] * 5), temp=0)
    g[0, 0] = cell(AIR, 800)
    orc.temperature(g, orc.default_config())
    t = temps(g)
    # two in-world neighbours at 0, two out-of-world neighbours count as 800
    assert t[0, 0] == 800 + ((0 + 0 + 800 + 800 - 3200) >> 3)
    assert t[0, 1] == 100 and t[1, 0] == 100


def test_temperature_saturates(orc):
    g = parse("...\n...\n...", temp=32767)
    g[1, 1] = cell(AIR, 32767)
    cfg = orc.default_config()
    cfg.materials[AIR].add_temp = 100
    orc.temperature(g, cfg)
    assert (temps(g) == 32767).all()


# --- phase schedule: independent restatement with shuffled block order -------------

def py_tick(orc, g, cfg, seed, tick, rng):
    H, W = g.shape
    OX, OY = [0, 1, 1, 0], [0, 1, 0, 1]
    WALL = 15
    for it in range(cfg.cell_iter):
        gph = tick * cfg.cell_iter + it
        key = orc.phase_key(seed, gph)
        ox, oy = OX[gph & 3], OY[gph & 3]
        blocks = [(x0, y0) for y0 in range(-oy, H, 2) for x0 in range(-ox, W, 2)]
        rng.shuffle(blocks)
        for x0, y0 in blocks:
            pos = [(x0, y0), (x0 + 1, y0), (x0, y0 + 1), (x0 + 1, y0 + 1)]
            q = [int(g[y, x]) if 0 <= x < W and 0 <= y < H else WALL for x, y in pos]
            r, s = orc.block_hash(key, x0, y0)
            _, q2 = orc.block_rule(cfg, it, r, s, q)
            for (x, y), v in zip(pos, q2):
                if 0 <= x < W and 0 <= y < H:
                    g[y, x] = v
    if cfg.temp_every and tick % cfg.temp_every == cfg.temp_phase:
        t = temps(g)
        p = np.pad(t, 1, mode="edge")
        s4 = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:]
        nt = t + ((s4 - 4 * t) >> 3)
        nt = np.clip(nt, -32768, 32767).astype(np.int16).view(np.uint16).astype(np.uint64)
        g[:] = (g & np.uint64(0xFFFFFFFF0000FFFF)) | (nt << np.uint64(16))
    return g


@pytest.mark.parametrize("W,H,seed", [(9, 7, 0), (16, 16, 1), (21, 12, 2)])
def test_schedule_order_independent(orc, W, H, seed):
    import random
    rng = np.random.default_rng(seed)
    mats_ = rng.choice(8, size=(H, W), p=[.35, .1, .2, .12, .08, .07, .05, .03])
    g = (mats_.astype(np.uint64) | (rng.integers(0, 6000, size=(H, W)).astype(np.uint64) << np.uint64(16)))
    cfg = orc.default_config()
    a = g.copy()
    b = g.copy()
    shuf = random.Random(seed)
    for t in range(6):
        orc.run(a, cfg, seed, t, 1)
        py_tick(orc, b, cfg, seed, t, shuf)
        assert (a == b).all(), t


# --- dirty-rect predicate -------------------------------------------------------------

def test_predicate_stable_worlds_do_not_change(orc):
    cfg = orc.default_config()
    g = parse("""
##########
#........#
#........#
#..####..#
#ssss#www#
#ssss#www#
##########""")
    assert not orc.region_unstable(g, cfg, 0, 0, 10, 7, 0, 0, 10, 7)
    for seed in range(30):
        h = g.copy()
        orc.run(h, cfg, seed, seed * 3, 2)
        assert (h == g).all()


def test_predicate_flags_each_kind_of_instability(orc):
    cfg = orc.default_config()
    cases = ["s\n.", "w.", ".\n~", "~.", "s.\n#.", "w\nl", "o\n."]
    for art in cases:
        g = parse(art)
        H, W = g.shape
        assert orc.region_unstable(g, cfg, 0, 0, W, H, 0, 0, W, H), art
    g = parse("#")
    g[0, 0] = cell(SAND, moved=True)
    assert orc.region_unstable(g, cfg, 0, 0, 1, 1, 0, 0, 1, 1)
    g = parse("..\n..")
    g[0, 0] = cell(AIR, 999)
    assert orc.region_unstable(g, cfg, 0, 0, 2, 2, 0, 0, 2, 2)


def test_predicate_is_sound_on_settled_worlds(orc):
    # The kernel skips a chunk C when its region C+4 (4 = cell_iter + 1) is
    # stable.  Whenever the predicate says so, a tick with any seed must leave
    # C unchanged.
    import falling_sand_engine_amd as fsg
    stable_seen = 0
    for trial in range(6):
        cfg = orc.default_config(temp_every=0 if trial % 2 == 0 else 1)
        W, H = 96, 64
        g = fsg.generate(1, W, H, 100 + trial)
        orc.run(g, cfg, trial, 0, 400)
        for ty in range(0, H, 16):
            for tx in range(0, W, 16):
                if orc.region_unstable(g, cfg, tx - 4, ty - 4, tx + 20, ty + 20, tx, ty, tx + 16, ty + 16):
                    continue
                stable_seen += 1
                for seed in range(3):
                    h = g.copy()
                    orc.run(h, cfg, seed, 400 + seed, 1)
                    assert (h[ty:ty + 16, tx:tx + 16] == g[ty:ty + 16, tx:tx + 16]).all(), (trial, tx, ty)
    assert stable_seen > 20
