"""The device block rule, compiled for the host through the fsg_rule_eval
test hook, against the oracle's block rule on random blocks: both the staged
form and the table-driven form the tick kernel runs must agree bit for bit."""
import ctypes as C

import numpy as np
import pytest

import falling_sand_engine_amd as fsg


def rand_blocks(rng, n, mats, hot_wood=False):
    m = rng.choice(mats, size=(n, 4))
    temp = rng.integers(0, 600 * 16, size=(n, 4)) if hot_wood else np.full((n, 4), 320)
    moved = rng.integers(0, 2, size=(n, 4))
    rgba = rng.integers(0, 2**32, size=(n, 4), dtype=np.uint64)
    return (m.astype(np.uint64) | (moved.astype(np.uint64) << np.uint64(8)) |
            (temp.astype(np.uint64) << np.uint64(16)) | (rgba << np.uint64(32)))


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("mats,reactions,hot", [
    ([0, 1, 2, 3, 4, 5, 6, 7], 1, True),
    ([0, 2, 3, 5], 0, False),          # the all-active world's materials
    ([0, 2, 2, 2, 1], 1, False),        # sand piles (friction paths)
    ([0, 3, 4, 6, 5], 1, False),        # liquids and lava
    ([0, 1, 5, 5, 2], 0, False),        # gas
    ([0, 1, 2, 3, 4, 5, 6, 7, 15], 1, True),  # with the world wall
])
def test_device_rule_matches_oracle(orc, form, mats, reactions, hot):
    rng = np.random.default_rng(hash((form, tuple(mats), reactions)) % 2**32)
    cfg = orc.default_config(reactions=reactions)
    lib = fsg.lib()
    blocks = rand_blocks(rng, 40000, mats, hot)
    rs = rng.integers(0, 2**32, size=(len(blocks), 2), dtype=np.uint64)
    bad = 0
    for i in range(len(blocks)):
        it = int(i % cfg.cell_iter)
        q = [int(v) for v in blocks[i]]
        r, s = int(rs[i, 0]), int(rs[i, 1])
        want_ch, want = orc.block_rule(cfg, it, r, s, q)
        arr = (C.c_uint64 * 4)(*q)
        got_ch = lib.fsg_rule_eval(C.byref(cfg), it, r, s, arr, form)
        assert got_ch in (0, 1)
        if list(arr) != want:
            bad += 1
            assert bad < 1, (form, it, [hex(v) for v in q], hex(r), hex(s),
                             [hex(v) for v in arr], [hex(v) for v in want])
        if form == 0:
            assert bool(got_ch) == want_ch


def test_rule_eval_rejects_bad_arguments():
    lib = fsg.lib()
    cfg = fsg.default_config()
    arr = (C.c_uint64 * 4)(0, 0, 0, 0)
    assert lib.fsg_rule_eval(C.byref(cfg), 3, 0, 0, arr, 0) == -1
    assert lib.fsg_rule_eval(C.byref(cfg), 0, 0, 0, arr, 7) == -1
