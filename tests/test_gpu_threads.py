"""Calls on different worlds of one context from several threads at once give
what the same calls give one after another: a batched step, body
stamp/tick/unstamp cycles on two worlds (their claim rounds share the
context's open-pixel list), chunk packing, timed steps and world create /
destroy churn.  ctypes drops the GIL for each library call, so the threads'
calls overlap in the library."""
import threading

import numpy as np
import pytest

import falling_sand_engine_amd as fsg
from bodies import random_bodies

pytestmark = pytest.mark.gpu


def jobs(ctx):
    """Fresh worlds and one job per thread; each job returns what it saw."""
    made = []

    def world(W, H, seed, kind):
        w = fsg.World(ctx, W, H, seed)
        w.upload(fsg.generate(kind, W, H, seed))
        made.append(w)
        return w

    batch = [world(64 * (3 + i % 4), 64 * (2 + i % 3), 300 + i, 4) for i in range(8)]

    def run_batch():
        for _ in range(6):
            fsg.batch_step(ctx, batch, 16)
        return [w.download() for w in batch]

    def bodies_job(W, H, seed, nb, side):
        w = world(W, H, seed, 1)
        bodies, pixels = random_bodies(np.random.default_rng(seed), W, H, nb, max_side=side)

        def run():
            out = []
            for _ in range(6):
                res = w.stamp_bodies(bodies, pixels)
                out.append((res["placed"], res["displaced"], res["hits_sand"], res["hits_soup"]))
                w.tick(2)
                out.append(w.unstamp_bodies(pixels.size))
            out.append(w.download())
            return out
        return run

    wp = world(512, 512, 77, 2)

    def run_pack():
        out = []
        for _ in range(5):
            wp.tick(3)
            out.append(wp.pack(0, 0, 4, 4))
        return out

    def run_churn():
        out = []
        for i in range(12):
            with fsg.World(ctx, 128 + 64 * (i % 3), 128, 900 + i) as w:
                w.generate(4, 900 + i)
                assert w.step_timed(3) >= 0.0
                out.append(w.stats().checksum)
        return out

    return made, [run_batch, bodies_job(320, 256, 2, 120, 16), bodies_job(512, 512, 3, 300, 32),
                  run_pack, run_churn]


def same(a, b):
    if isinstance(a, np.ndarray):
        return np.array_equal(a, b)
    if isinstance(a, (list, tuple)):
        return len(a) == len(b) and all(same(x, y) for x, y in zip(a, b))
    return a == b


def test_threads_on_one_context_match_one_at_a_time(ctx):
    made, js = jobs(ctx)
    print("sequential", flush=True)
    seq = [j() for j in js]
    for w in made:
        w.close()

    made, js = jobs(ctx)
    got = [None] * len(js)
    errors = []

    def runner(k):
        try:
            got[k] = js[k]()
        except Exception as e:  # reported by the main thread
            errors.append((k, repr(e)))

    print("threaded", len(js), "jobs", flush=True)
    ts = [threading.Thread(target=runner, args=(k,)) for k in range(len(js))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in ts)
    for w in made:
        w.close()
    assert not errors, errors
    for k in range(len(js)):
        assert same(got[k], seq[k]), k
