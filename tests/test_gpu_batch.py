"""Many worlds in shared launches (chunk work-list + HIP graph of 16 ticks)
vs each world stepped alone, and vs the oracle; host-forced failures."""
import numpy as np
import pytest

import falling_sand_engine_amd as fsg

pytestmark = pytest.mark.gpu


def uneven_worlds(ctx, n, seed0, skip=0, lo=64, hi=384):
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(n):
        W = int(rng.integers(lo // 64, hi // 64 + 1)) * 64
        H = int(rng.integers(lo // 64, hi // 64 + 1)) * 64
        cells = fsg.generate(4, W, H, seed0 + i)
        cfg = fsg.default_config(chunk_skip=skip)
        w = fsg.World(ctx, W, H, seed0 + i, cfg)
        w.upload(cells)
        out.append((w, cells, cfg))
    return out


@pytest.mark.parametrize("skip", [0, 1])
def test_batch_equals_alone_and_oracle(ctx, orc, skip):
    worlds = uneven_worlds(ctx, 10, 500, skip)
    alone = uneven_worlds(ctx, 10, 500, skip)
    fsg.batch_step(ctx, [w for w, _, _ in worlds], 37)   # 2 graph replays + 5 direct ticks
    for w, _, _ in alone:
        w.tick(37)
    for (wb, cells, cfg), (wa, _, _) in zip(worlds, alone):
        assert wb.tick_count == 37
        assert np.array_equal(wb.download(), wa.download())
    for (w, cells, cfg) in worlds[:3]:
        ref = orc.run(cells.copy(), orc.default_config(chunk_skip=skip), w.seed, 0, 37)
        assert np.array_equal(w.download(), ref)
    for w, _, _ in worlds + alone:
        w.close()


def test_batch_graph_reuse_and_mixed_ticks(ctx, orc):
    worlds = uneven_worlds(ctx, 6, 900)
    ws = [w for w, _, _ in worlds]
    fsg.batch_step(ctx, ws, 16)
    ws[0].tick(3)                 # one world moves ahead alone
    fsg.batch_step(ctx, ws, 32)   # replays the cached graph with per-world ticks
    for i, (w, cells, cfg) in enumerate(worlds):
        n = 48 + (3 if i == 0 else 0)
        assert w.tick_count == n
        if i < 2:
            assert np.array_equal(w.download(), orc.run(cells.copy(), orc.default_config(), w.seed, 0, n))
    for w in ws:
        w.close()


def test_budget_and_alloc_failures(ctx):
    c = fsg.Context(0, hbm_budget=3 * 1024 * 1024)
    try:
        w = fsg.World(c, 256, 256, 0)           # 1 MiB of state
        used = c.usage()
        assert used > 0
        with pytest.raises(fsg.FsgError) as e:
            fsg.World(c, 512, 512, 0)           # 4 MiB: over the budget
        assert e.value.status == -4
        assert c.usage() == used
        c.inject_alloc_failures(1)
        with pytest.raises(fsg.FsgError) as e:
            fsg.World(c, 128, 128, 0)
        assert e.value.status == -2
        assert c.usage() == used                # a failed create releases its reservation
        w2 = fsg.World(c, 128, 128, 0)          # and the next one succeeds
        w2.close()
        w.close()
        assert c.usage() == 0
    finally:
        c.close()


def test_invalid_cells_rejected(ctx):
    with fsg.World(ctx, 64, 64, 0) as w:
        bad = np.zeros((64, 64), dtype=np.uint64)
        bad[3, 3] = 9   # no material 9 in the 8-material table
        with pytest.raises(fsg.FsgError) as e:
            w.upload(bad)
        assert e.value.status == -1
        cfg = fsg.default_config(cell_iter=4)   # the fused tick supports cell_iter <= 3
        with pytest.raises(fsg.FsgError):
            fsg.World(ctx, 64, 64, 0, cfg)


@pytest.mark.parametrize("tuning", [(1, -1, 5), (4, 1, 5), (8, 0, 4), (32, 1, 4), (16, 0, 5)])
def test_every_tuning_is_bit_exact(ctx, tuning):
    """Graph block size, skipping override and kernel variant change speed only."""
    T, skip, occ = tuning
    worlds = uneven_worlds(ctx, 6, 700, skip=1, lo=64, hi=320)
    alone = uneven_worlds(ctx, 6, 700, skip=1, lo=64, hi=320)
    ctx.set_tuning(graph_ticks=T, chunk_skip=skip, occupancy=occ)
    try:
        fsg.batch_step(ctx, [w for w, _, _ in worlds], 41)
        fsg.batch_step(ctx, [w for w, _, _ in worlds[:3]], 7)   # another world set, then back
        fsg.batch_step(ctx, [w for w, _, _ in worlds], 5)
    finally:
        ctx.set_tuning()
    for i, (w, _, _) in enumerate(alone):
        w.tick(53 if i < 3 else 46)
    for (wb, _, _), (wa, _, _) in zip(worlds, alone):
        assert np.array_equal(wb.download(), wa.download())
    # a world stepped alone afterwards (its own skipping) still matches
    worlds[0][0].tick(9)
    alone[0][0].tick(9)
    assert np.array_equal(worlds[0][0].download(), alone[0][0].download())
    for w, _, _ in worlds + alone:
        w.close()


def test_autotune_leaves_worlds_untouched_and_caches(ctx):
    worlds = uneven_worlds(ctx, 5, 900, skip=1, lo=64, hi=256)
    before = [w.download() for w, _, _ in worlds]
    base = ctx.usage()
    rep = ctx.autotune([w for w, _, _ in worlds], sample_ticks=8)
    assert not rep["from_cache"]
    assert len(rep["candidates"]) == 20
    assert all(c["ms"] > 0 for c in rep["candidates"])
    best = min(rep["candidates"], key=lambda c: c["ms"])
    assert rep["best"] == {k: best[k] for k in ("graph_ticks", "chunk_skip", "occupancy")}
    assert ctx.tuning() == rep["best"]
    assert ctx.usage() == base                      # the copies are gone
    for (w, _, _), b in zip(worlds, before):
        assert w.tick_count == 0 and np.array_equal(w.download(), b)
    again = ctx.autotune([w for w, _, _ in worlds], sample_ticks=8)
    assert again["from_cache"] and again["best"] == rep["best"]
    ctx.set_tuning()
    for w, _, _ in worlds:
        w.close()
