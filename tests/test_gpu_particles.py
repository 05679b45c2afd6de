"""Device particle pass (fsg_world_particles_tick) against the oracle's
world::tickCells: the same particle list (every field, bit for bit) and the
same grid after each pass, with sparse sprays and dense clusters whose
landings depend on each other's order."""
import numpy as np
import pytest

import falling_sand_engine_amd as fsg
from falling_sand_engine_amd import _abi

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "vx", "vy", "ax", "ay", "target_x", "target_y", "target_force", "lifetime", "flags",
          "in_object_state")


def key(p):
    return tuple(getattr(p, f) for f in FIELDS) + (p.tile.mat, p.tile.temp, p.tile.rgba)


def spray(rng, W, H, n, cluster=None):
    out = []
    mats = [fsg.MAT_SAND, fsg.MAT_WATER, fsg.MAT_STONE, fsg.MAT_OIL]
    for i in range(n):
        if cluster is None:
            x, y = rng.uniform(1, W - 1), rng.uniform(1, H * 0.6)
        else:
            cx, cy = cluster
            x, y = cx + rng.normal(0, 1.5), cy + rng.normal(0, 1.5)
        c = fsg.make_cell(int(rng.choice(mats)), int(rng.integers(-100, 900)), int(rng.integers(0, 2**32)))
        flags = 0
        r = rng.random()
        if r < 0.05:
            flags = _abi.PARTICLE_PHASE
        elif r < 0.12:
            flags = _abi.PARTICLE_TEMPORARY
        target = (float(rng.uniform(0, W)), float(rng.uniform(0, H)), 0.3) if rng.random() < 0.05 else None
        out.append(fsg.make_particle(c, float(x), float(y), float(rng.uniform(-1.5, 1.5)),
                                     float(rng.uniform(-1.0, 2.5)), 0.0, 0.1,
                                     lifetime=int(rng.integers(0, 40)), flags=flags, target=target))
    return out


@pytest.mark.parametrize("seed,n,clusters", [(1, 3000, 0), (2, 2000, 6), (3, 600, 1)])
def test_particles_match_oracle(ctx, orc, seed, n, clusters):
    rng = np.random.default_rng(seed)
    W, H = 384, 256
    cfg = orc.default_config()
    cells = fsg.generate(1, W, H, seed)
    ps = spray(rng, W, H, n)
    for k in range(clusters):
        ps += spray(rng, W, H, 150, cluster=(float(rng.uniform(40, W - 40)), float(rng.uniform(60, 160))))
    print(f"particles seed={seed} n={len(ps)} clusters={clusters}", flush=True)
    ref = cells.copy()
    ref_ps = list(ps)
    with fsg.World(ctx, W, H, seed, cfg) as w:
        w.upload(cells)
        w.particles_add(ps)
        total = 0
        for t in range(60):
            alive, landed = w.particles_tick()
            ref_ps, ref_landed = orc.particles_tick(ref, cfg, ref_ps)
            assert (alive, landed) == (len(ref_ps), ref_landed), t
            got = w.particles()
            assert [key(p) for p in got] == [key(p) for p in ref_ps], t
            total += landed
            if t % 10 == 9 or not alive:
                assert np.array_equal(w.download(), ref), t
            if not alive:
                break
    assert total > 100


def test_particles_then_tick_with_chunk_skipping(ctx, orc):
    """Landings wake their chunks: ticking afterwards with skipping on equals
    the oracle, which has no skipping."""
    rng = np.random.default_rng(5)
    W, H = 320, 256
    cfg = orc.default_config(chunk_skip=1)
    cells = fsg.generate(1, W, H, 5)
    ps = spray(rng, W, H, 1500)
    ref = cells.copy()
    orc.run(ref, cfg, 5, 0, 30)
    ref_ps = list(ps)
    for _ in range(25):
        ref_ps, _ = orc.particles_tick(ref, cfg, ref_ps)
    orc.run(ref, cfg, 5, 30, 20)
    with fsg.World(ctx, W, H, 5, cfg) as w:
        w.upload(cells)
        w.tick(30)
        w.particles_add(ps)
        for _ in range(25):
            w.particles_tick()
        w.tick(20)
        assert np.array_equal(w.download(), ref)


def test_particle_arguments(ctx, orc):
    with fsg.World(ctx, 64, 64, 1, orc.default_config()) as w:
        bad = fsg.make_particle(fsg.make_cell(12), 3.0, 3.0)  # no material 12
        with pytest.raises(fsg.FsgError):
            w.particles_add([bad])
        nan = fsg.make_particle(fsg.make_cell(2), float("nan"), 3.0)
        with pytest.raises(fsg.FsgError):
            w.particles_add([nan])
        assert w.particles_tick() == (0, 0) and w.particles() == []
