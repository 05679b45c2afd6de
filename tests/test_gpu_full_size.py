"""The tick at the full sizes of the benchmark worlds.

At 4096^2 (the basin world with lava and oil: reactions, temperature and
dirty-chunk skipping on) the device's grid after two dozen ticks equals the
oracle's, cell for cell.  At 16384^2 (the all-active striped world, 268M
cells) the oracle would take minutes, so the checks are properties that do
not depend on size: with reactions off every material's count is kept, and
the same ticks give the same grid with chunk skipping on or off and through
the batch kernel (a different instantiation of the chunk update, launched
over a work list) as through the single-world kernel."""
import numpy as np
import pytest

import falling_sand_engine_amd as fsg

pytestmark = pytest.mark.gpu


def test_4096_basin_with_reactions_matches_oracle(ctx, orc):
    side, seed, ticks = 4096, 2, 24
    cells = fsg.generate(2, side, side, seed)
    cfg = orc.default_config(chunk_skip=1)
    print("4096^2 basin world vs the oracle", flush=True)
    with fsg.World(ctx, side, side, seed, cfg) as w:
        w.upload(cells)
        w.tick(ticks)
        got = w.download()
        st = w.stats()
    ref = orc.run(cells, cfg, seed, 0, ticks)  # in place
    assert np.array_equal(got, ref)
    assert st.tiles_processed < st.tiles_offered  # chunks were skipped


def test_16384_all_active_world_properties(ctx, orc):
    side, seed, ticks = 16384, 3, 20
    sums = {}
    print("16384^2 all-active world: no skip, skip, batch", flush=True)
    for name, skip, batched in (("plain", 0, False), ("skip", 1, False), ("batch", 0, True)):
        with fsg.World(ctx, side, side, seed, orc.default_config(reactions=0, chunk_skip=skip)) as w:
            w.generate(3, seed)
            before = w.stats()
            if batched:
                fsg.batch_step(ctx, [w], ticks)
            else:
                w.tick(ticks)
            after = w.stats()
        assert list(after.count) == list(before.count), name  # no reactions: counts kept
        assert after.checksum != before.checksum, name
        sums[name] = after.checksum
    assert sums["plain"] == sums["skip"] == sums["batch"]
