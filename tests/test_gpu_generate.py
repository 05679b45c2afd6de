"""Device world generator (fsg_world_generate) vs the host generator
(fsg_generate): the same cells, bit for bit."""
import numpy as np
import pytest

import falling_sand_engine_amd as fsg
from falling_sand_engine_amd import _abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,H", [(64, 64), (100, 37), (1, 1), (4096, 64), (513, 1025), (2048, 2048), (70, 3000)])
@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_device_generator_matches_host(ctx, W, H, kind):
    seed = 1000 * kind + W + 7 * H
    want = fsg.generate(kind, W, H, seed)
    print(f"generate kind={kind} {W}x{H} seed={seed}", flush=True)
    with fsg.World(ctx, W, H, 1, fsg.default_config()) as w:
        w.generate(kind, seed)
        got = w.download()
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{len(bad)} cells differ; first (x={bad[0][1]}, y={bad[0][0]})"


def test_generated_world_ticks_like_an_uploaded_one(ctx):
    W, H = 300, 200
    cfg = fsg.default_config(chunk_skip=1)
    with fsg.World(ctx, W, H, 9, cfg) as a, fsg.World(ctx, W, H, 9, cfg) as b:
        a.generate(4, 77)
        b.upload(fsg.generate(4, W, H, 77))
        a.tick(40)
        b.tick(40)
        assert np.array_equal(a.download(), b.download())
        assert fsg.lib().fsg_world_generate(a.handle, 9, 1) == _abi.FSG_ERR_INVALID
