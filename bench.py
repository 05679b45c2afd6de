#!/usr/bin/env python3
"""Falling-sand tick throughput on one MI355X.

Default workload: the 16384 x 16384 all-active world (268,435,456 cells,
2.1 GB per state buffer; rows alternate 50% sand / air, 10% water rows, 10%
steam rows, stone floor; reactions off; uniform 20 degrees), ticked in
place in HBM.  A step is one tick of the whole world; the timed region is
exactly --steps ticks between two stream synchronisations.

Prints ONE JSON line: Gcell-updates/s, the HBM roofline of the tick kernel
(16 algorithmic bytes per cell-update: one 8-byte cell read and one write),
the CPU baseline (the reference's sequential tick restated, timed on this
host's cores on a bounded sample), and secondary figures for the 4096^2
dirty-chunk world, the 256-world batch and the serving loop.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 16


PROFILE = os.path.join(ROOT, "profiles", "k_tick_world_16384_all_active.md")


def profile_field(side, key):
    """A `key: value` line of the committed rocprofv3 summary of the tick
    kernel on the 16384^2 world, or None."""
    if side != 16384 or not os.path.exists(PROFILE):
        return None
    for line in open(PROFILE):
        if line.startswith(key + ":"):
            return line.split(":", 1)[1].strip()
    return None


def profiled_traffic(side):
    """HBM bytes per tick launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE
    (copy-kernel calibrated) in the committed profile summary, or None."""
    v = profile_field(side, "traffic_bytes_per_launch")
    return None if v is None else int(v)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bench_c3(fsg, steps, warmup, side=16384, seed=3):
    cells = fsg.generate(3, side, side, seed)
    cfg = fsg.default_config(reactions=0, temp_every=1)
    ctx = fsg.Context(0)
    w = fsg.World(ctx, side, side, seed, cfg)
    w.upload(cells)
    del cells
    if warmup:
        w.tick(warmup)
    w.sync()
    t0 = time.perf_counter()
    ms_dev = w.step_timed(steps)  # HIP events on the ctx stream around the launches
    w.sync()
    wall = time.perf_counter() - t0
    st = w.stats()
    w.close()
    ctx.close()
    return {"wall_s": wall, "dev_ms": ms_dev, "cells": side * side, "counts": list(st.count)[:8]}


def bench_c1(fsg, ticks=1000, side=512, seed=1):
    """The reference's own CPU-runnable case: 512^2 basin world with sand
    discs, 1000 ticks on the GPU, and the same 1000 ticks of the sequential
    tick restatement on the host's cores."""
    from oracle import pyoracle
    cells = fsg.generate(1, side, side, seed)
    with fsg.Context(0) as ctx, fsg.World(ctx, side, side, seed, fsg.default_config()) as w:
        w.upload(cells)
        w.sync()
        t0 = time.perf_counter()
        ms = w.step_timed(ticks)
        wall = time.perf_counter() - t0
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    pyoracle.refseq(cells.copy(), pyoracle.default_config(), seed, ticks, threads)
    cpu_s = time.perf_counter() - t0
    cu = side * side * ticks
    return {"workload": f"basin world {side}x{side} with sand discs, {ticks} ticks",
            "gpu_gcell_s": round(cu / wall / 1e9, 2), "gpu_dev_ms": round(ms, 2),
            "cpu_refseq_gcell_s": round(cu / cpu_s / 1e9, 4), "cpu_threads": threads,
            "cpu_refseq_s": round(cpu_s, 2)}


def bench_c2(fsg, ticks=500, side=4096, seed=2):
    """4096^2 basin world with lava and oil, reactions on, temperature every
    tick, exact dirty-rect skipping on 64x64 chunks; also run without
    skipping to show the two are bit-identical."""
    cells = fsg.generate(2, side, side, seed)
    out = {}
    with fsg.Context(0) as ctx:
        for skip in (1, 0):
            w = fsg.World(ctx, side, side, seed, fsg.default_config(chunk_skip=skip))
            w.upload(cells)
            w.sync()
            w.reset_activity()
            t0 = time.perf_counter()
            ms = w.step_timed(ticks)
            wall = time.perf_counter() - t0
            st = w.stats()
            out[skip] = {"wall_s": wall, "dev_ms": ms, "checksum": st.checksum,
                         "active_fraction": st.tiles_processed / max(1, st.tiles_offered)}
            w.close()
    cu = side * side * ticks
    return {"workload": f"basin world {side}x{side}, lava+oil, reactions on, {ticks} ticks",
            "gcell_s_nominal": round(cu / out[1]["wall_s"] / 1e9, 2),
            "gcell_s_active": round(cu * out[1]["active_fraction"] / out[1]["wall_s"] / 1e9, 2),
            "active_fraction": round(out[1]["active_fraction"], 4),
            "gcell_s_no_skip": round(cu / out[0]["wall_s"] / 1e9, 2),
            "skip_speedup": round(out[0]["wall_s"] / out[1]["wall_s"], 3),
            "bit_exact_vs_no_skip": out[0]["checksum"] == out[1]["checksum"]}


def uneven_batch_sizes(n=256, seed=4):
    """Sides log-uniform in [256, 2048] rounded to 64, aspect ratio uniform
    in [0.5, 2]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = []
    for _ in range(n):
        side = float(np.exp(rng.uniform(np.log(256), np.log(2048))))
        a = float(rng.uniform(0.5, 2.0))
        w = max(64, int(round(side * np.sqrt(a) / 64)) * 64)
        h = max(64, int(round(side / np.sqrt(a) / 64)) * 64)
        sizes.append((w, h))
    return sizes


def bench_c4(fsg, ticks=256, n=256, tune_ticks=32):
    """256 worlds of uneven sizes: one launch per tick over all of their
    chunks (HIP graphs of ticks, launch shape picked by the autotuner on
    copies of the same worlds) vs one launch per world per tick.  Per-world
    launches run with the same dirty-chunk skipping as the tuned batch, so
    the headline ratio is batching alone; the batch with skipping off is
    timed against per-world launches without it as well."""
    sizes = uneven_batch_sizes(n)
    with fsg.Context(0) as ctx:
        def world_set(skip):
            ws = []
            for i, (W, H) in enumerate(sizes):
                w = fsg.World(ctx, W, H, 1000 + i, fsg.default_config(chunk_skip=skip))
                w.generate(4, 1000 + i)
                ws.append(w)
            return ws
        tune_set = world_set(0)
        t0 = time.perf_counter()
        tune = ctx.autotune(tune_set, sample_ticks=tune_ticks)
        tune_s = time.perf_counter() - t0
        for w in tune_set:
            w.close()
        best = tune["best"]
        skip = 1 if best["chunk_skip"] == 1 else 0
        runs = {}
        # batched with the tuned choice; batched with skipping off; per-world
        # with the batch's skipping; per-world without skipping
        for name, batched, sk in (("batched", True, None), ("batched_noskip", True, 0),
                                  ("perworld", False, skip), ("perworld_noskip", False, 0)):
            ws = world_set(skip if sk is None else sk)
            if batched:
                ctx.set_tuning(best["graph_ticks"], best["chunk_skip"] if sk is None else 0, best["occupancy"])
                fsg.batch_step(ctx, ws, 16)  # warm-up + graph capture
            else:
                for w in ws:
                    w.tick(16)
            ws[0].sync()
            t0 = time.perf_counter()
            ms = fsg.batch_step_timed(ctx, ws, ticks) if batched else fsg.perworld_step_timed(ctx, ws, ticks)
            wall = time.perf_counter() - t0
            runs[name] = {"wall": wall, "ms": ms, "sums": [w.stats().checksum for w in ws]}
            for w in ws:
                w.close()
        ctx.set_tuning(best["graph_ticks"], best["chunk_skip"], best["occupancy"])
    exact = all(r["sums"] == runs["batched"]["sums"] for r in runs.values())
    cells = sum(W * H for W, H in sizes)
    cand = sorted(tune["candidates"], key=lambda c: c["ms"])
    g = {k: round(cells * ticks / r["wall"] / 1e9, 2) for k, r in runs.items()}
    return {"workload": f"{n} worlds, sides log-uniform [256,2048], aspect [0.5,2], {ticks} ticks",
            "cells": cells,
            "batched_gcell_s": g["batched"], "perworld_gcell_s": g["perworld"],
            "ratio": round(runs["perworld"]["wall"] / runs["batched"]["wall"], 3),
            "ratio_note": f"per-world launches with the batch's chunk skipping ({skip}): batching alone",
            "batched_noskip_gcell_s": g["batched_noskip"], "perworld_noskip_gcell_s": g["perworld_noskip"],
            "ratio_noskip": round(runs["perworld_noskip"]["wall"] / runs["batched_noskip"]["wall"], 3),
            "ratio_tuned_vs_perworld_noskip": round(runs["perworld_noskip"]["wall"] / runs["batched"]["wall"], 3),
            "dev_ms": {k: round(r["ms"], 2) for k, r in runs.items()},
            "bit_exact": exact,
            "tuning": tune["best"],
            "autotune": {"sample_ticks": tune_ticks, "seconds": round(tune_s, 2),
                         "ms_per_tick": {f"T{c['graph_ticks']}/skip{c['chunk_skip']}/occ{c['occupancy']}":
                                         round(c["ms"] / tune_ticks, 3) for c in cand}}}


def bench_c5(fsg, rates=(30.0, 120.0, 400.0, 1000.0), duration_s=6.0, budget=64 << 30, seed=5):
    """Serving: Poisson arrivals at each offered rate for duration_s.  Each
    request creates a world (side uniform in {512..4096 step 64}, config-2
    mix scaled to its area, 0-200 rigid bodies) or steps a resident one,
    for k ~ U[1,64] ticks, under a 64 GB HBM state budget."""
    import numpy as np
    out = []
    for rate in rates:
        rng = np.random.default_rng(seed + int(rate))
        with fsg.Context(0, hbm_budget=budget) as ctx, fsg.Server(ctx) as srv:
            t0 = time.perf_counter()
            created, n, t = [], 0, 0.0
            while True:
                t += float(rng.exponential(1.0 / rate))
                if t > duration_s:
                    break
                delay = t0 + t - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                k = int(rng.integers(1, 65))
                if created and rng.random() < 0.5:
                    wid, W, H = created[int(rng.integers(len(created)))]
                    srv.submit(wid, k, kind=4, width=W, height=H, seed=wid)
                else:
                    side = int(rng.integers(512 // 64, 4096 // 64 + 1)) * 64
                    wid = 10_000 + n
                    created.append((wid, side, side))
                    srv.submit(wid, k, kind=4, width=side, height=side, seed=wid,
                               n_bodies=int(rng.integers(0, 201)))
                n += 1
            rc = srv.drain(600)
            wall = time.perf_counter() - t0
            res = srv.poll(1 << 16)
            st = srv.stats()
        ok = [r for r in res if r["status"] == 0]
        lat = np.array([(r["t_done_ns"] - r["t_submit_ns"]) / 1e6 for r in ok])
        cu = sum(r["cells"] * r["ticks_done"] for r in ok)
        out.append({"offered_req_s": rate, "requests": n, "completed": len(ok), "drained": rc == 0,
                    "completed_req_s": round(len(ok) / wall, 2),
                    "gcell_s": round(cu / wall / 1e9, 2),
                    "p50_ms": round(float(np.percentile(lat, 50)), 2) if len(lat) else None,
                    "p99_ms": round(float(np.percentile(lat, 99)), 2) if len(lat) else None,
                    "steps": st["steps"], "evictions": st["evictions"]})
    return {"workload": f"Poisson request stream, sides U{{512..4096}}, k~U[1,64] ticks, 0-200 bodies, "
                        f"{budget >> 30} GB state budget, {duration_s:.0f} s per rate",
            "rates": out}


def bench_render(fsg, side=16384, seed=3, reps=20):
    """Texture pass (game.cpp:1994-2067) over the whole all-active world after
    10 ticks, device time from HIP events on the ctx stream: the cells layer
    alone (8 B read + 4 B written per cell), the three layers (8 B read, 8 B
    written, 4 B per fire-layer texel that changes: AIR and fire cells), and
    the three layers with the movingTiles census (8 B more read; each call
    also waits for the 17 counters)."""
    cells = fsg.generate(3, side, side, seed)
    cfg = fsg.default_config(reactions=0, temp_every=1)
    n = side * side
    allm = fsg.LAYER_CELLS | fsg.LAYER_EMISSION | fsg.LAYER_FIRE
    fire = fsg.MAT_WATER  # stands in for the reference's FIRE material
    out = {"workload": f"all-active striped world {side}x{side} after 10 ticks, whole-world render, "
                       f"fire layer drawn for water", "reps": reps}
    with fsg.Context(0) as ctx, fsg.World(ctx, side, side, seed, cfg) as w:
        w.upload(cells)
        del cells
        w.tick(10)
        st = w.stats()
        fire_texels = int(st.count[fsg.MAT_AIR] + st.count[fire])
        out["fire_layer_texels"] = fire_texels
        cases = [("cells_layer", fsg.LAYER_CELLS, False, 12 * n),
                 ("three_layers", allm, False, 16 * n + 4 * fire_texels),
                 ("three_layers_census", allm, True, 24 * n + 4 * fire_texels)]
        for name, layers, census, nbytes in cases:
            w.render(layers, fire_mat=fire, stats=census)  # allocation + warm-up
            ctx.timer_start()
            for _ in range(reps):
                w.render(layers, fire_mat=fire, stats=census)
            ms = ctx.timer_stop() / reps
            gbs = nbytes / (ms / 1e3) / 1e9
            out[name] = {"ms": round(ms, 4), "gcell_s": round(n / ms / 1e6, 1),
                         "algorithmic_bytes_per_cell": round(nbytes / n, 3),
                         "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                                      "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}}
    return out


def fragments(side, seed, p=0.5):
    """A world of stone fragments near the percolation threshold (islands of
    every size up to one spanning component) among air, sand and water."""
    import numpy as np
    import falling_sand_engine_amd as fsg
    rng = np.random.default_rng(seed)
    mats = np.where(rng.random((side, side)) < p, fsg.MAT_STONE,
                    rng.choice([fsg.MAT_AIR, fsg.MAT_SAND, fsg.MAT_WATER], size=(side, side)))
    return mats.astype(np.uint64) | (np.uint64(320) << np.uint64(16)) | (np.uint64(0xFF808080) << np.uint64(32))


def bench_islands(fsg, side=4096, nseeds=4096, seed=9):
    """world::physicsCheck for a batch of random seed cells on a world of stone
    fragments: device time of the whole call (search launch, host dedupe,
    image/clear launch, copies).  BFS levels and barriers bound it, not HBM."""
    import numpy as np
    cells = fragments(side, seed)
    rng = np.random.default_rng(seed)
    seeds = [(int(x), int(y)) for x, y in zip(rng.integers(0, side, nseeds), rng.integers(0, side, nseeds))]
    with fsg.Context(0) as ctx, fsg.World(ctx, side, side, seed, fsg.default_config()) as w:
        w.upload(cells)
        w.physicsCheck(seeds)  # warm-up at full size (the world's scratch grows here, not in the timed call)
        w.upload(cells)        # the untouched world again
        arr = (fsg._abi.Island * nseeds)()
        for i, (x, y) in enumerate(seeds):
            arr[i].x, arr[i].y = x, y
        need = fsg.C.c_size_t()
        ctx.timer_start()  # the C call alone: search, host dedupe, images and clearing
        fsg.check(fsg.lib().fsg_world_physics_check(w.handle, arr, nseeds, fsg.C.byref(need)))
        ms = ctx.timer_stop()
    st = [arr[i].status for i in range(nseeds)]
    return {"workload": f"{side}x{side} world of stone fragments (p=0.5), {nseeds} random seed cells",
            "ms": round(ms, 3), "seeds_per_s": round(nseeds / (ms / 1e3)),
            "bodies": st.count(3), "removed_small": st.count(2), "attached": st.count(1), "none": st.count(0)}


def bench_mesh(fsg, side=4096, seed=12, reps=3):
    """Outlines of the SOLID cells of every 128x128 chunk that ends before
    the world's last row and column (world::updateChunkMesh for a whole mesh
    zone, world.cpp:722-998) on a basin world and on a world of stone
    fragments; the whole synchronous call (launch, pools, host ordering,
    download of the loop and vertex arrays).  The oracle's restatement of
    the same step in C, one thread, on 64 chunks of each world is the CPU
    figure."""
    import numpy as np
    from oracle import pyoracle
    out = {"workload": f"every 128x128 chunk of {side}x{side} worlds (basin kind 2; stone fragments p=0.5)"}
    n = side // 128
    regions = [(cx * 128, cy * 128, 128, 128) for cy in range(n) for cx in range(n)]
    cfg = fsg.default_config()
    for name, cells in (("basin", fsg.generate(2, side, side, seed)), ("fragments", fragments(side, seed))):
        with fsg.Context(0) as ctx, fsg.World(ctx, side, side, seed, cfg) as w:
            w.upload(cells)
            w.mesh(regions, raw=True)  # pools sized
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                regs, lps, vs = w.mesh(regions, raw=True)
                ts.append(time.perf_counter() - t0)
        loops = sum(regs[i].nloops for i in range(len(regions)))
        meshed = sum(1 for i in range(len(regions)) if regs[i].status == 0)
        types = np.array([cfg.materials[i].type for i in range(cfg.n_materials)] + [0] * 16, np.uint8)
        solid = (types[(cells & np.uint64(15)).astype(np.int64)] == fsg.SOLID).astype(np.uint8)
        sample = [r for r in regions if r[0] + 128 < side and r[1] + 128 < side][:64]
        masks = [np.ascontiguousarray(solid[y0:y0 + 128, x0:x0 + 128]) for x0, y0, _w, _h in sample]
        t0 = time.perf_counter()
        for m in masks:
            pyoracle.mesh_mask(m, raw=True)
        cpu = (time.perf_counter() - t0) / len(sample)
        ms = min(ts) * 1e3
        out[name] = {"chunks_meshed": meshed, "loops": loops, "ms": round(ms, 2),
                     "chunks_per_s": round(meshed / (ms / 1e3)),
                     "cpu_oracle_1_thread_chunks_per_s": round(1.0 / cpu, 1)}
    return out


def bench_particles(fsg, side=4096, n=100000, passes=20, seed=21):
    """world::tickCells over a spray of n particles (cells carried by bodies'
    displaced cells in the reference) above a basin world: device time per
    pass of the whole synchronous call; the oracle's restatement, one thread,
    on the same list is the CPU figure."""
    import numpy as np
    from oracle import pyoracle
    rng = np.random.default_rng(seed)
    cfg = fsg.default_config()
    cells = fsg.generate(2, side, side, seed)
    ps = [fsg.make_particle(fsg.make_cell(int(rng.choice([2, 3, 4])), 320, 0xFF808080),
                            float(rng.uniform(1, side - 1)), float(rng.uniform(1, side * 0.5)),
                            float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1)), 0.0, 0.1) for _ in range(n)]
    with fsg.Context(0) as ctx, fsg.World(ctx, side, side, seed, cfg) as w:
        w.upload(cells)
        w.particles_add(ps)
        w.particles_tick()
        ts, landed = [], 0
        for _ in range(passes):
            t0 = time.perf_counter()
            alive, nl = w.particles_tick()
            ts.append(time.perf_counter() - t0)
            landed += nl
    ref = cells.copy()
    t0 = time.perf_counter()
    pyoracle.particles_tick(ref, pyoracle.default_config(), ps)
    cpu = time.perf_counter() - t0
    ms = float(np.median(ts)) * 1e3
    return {"workload": f"{n} particles over a {side}x{side} basin world, {passes} passes",
            "ms_per_pass": round(ms, 3), "particles_per_s": round(n / (ms / 1e3)), "landed": landed,
            "alive_after": alive, "cpu_oracle_1_thread_particles_per_s": round(n / cpu)}


PACK_RAW = 393216 + 65536  # .pack payload per chunk: tiles + layer2 records, background


def bench_pack(fsg, side=16384, kind=2, seed=41, reps=3):
    """Chunk .pack images of every 128x128 chunk of a side^2 basin world
    (noisy terrain colours, water, lava and oil discs; no layer2 or
    background, as the benchmark worlds have none): device time of
    fsg_world_pack (encode, offsets, assembly, the size copy) and of
    fsg_world_unpack from the device-resident images (decode, checks, cell
    writes, the status copy); the round trip is checked by the grid checksum."""
    import numpy as np
    n = (side // 128) ** 2
    sizes = np.zeros(n, dtype=np.uintp)
    status = np.zeros(n, dtype=np.int32)
    L = fsg.lib()
    with fsg.Context(0) as ctx, fsg.World(ctx, side, side, seed, fsg.default_config()) as w:
        w.generate(kind, seed)
        before = w.stats().checksum
        pms, ums = [], []
        for _ in range(reps + 1):
            ctx.timer_start()
            fsg.check(L.fsg_world_pack(w.handle, 0, 0, side // 128, side // 128, None, None, None,
                                       sizes.ctypes.data), "fsg_world_pack")
            pms.append(ctx.timer_stop())
            ctx.timer_start()
            fsg.check(L.fsg_world_unpack(w.handle, 0, 0, side // 128, side // 128, None, None, None, None, None,
                                         status.ctypes.data), "fsg_world_unpack")
            ums.append(ctx.timer_stop())
        ok = bool((status == 0).all()) and w.stats().checksum == before
        # through host memory, as a caller writing and reading files sees it: pack + download of the
        # images, and unpack from host images (header parse, upload, decode)
        total = int(sizes.sum())
        host = np.empty(total, dtype=np.uint8)
        t0 = time.perf_counter()
        fsg.check(L.fsg_world_pack(w.handle, 0, 0, side // 128, side // 128, None, None, None, sizes.ctypes.data))
        fsg.check(L.fsg_world_pack_download(w.handle, host.ctypes.data, total))
        hp = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        fsg.check(L.fsg_world_unpack(w.handle, 0, 0, side // 128, side // 128, host.ctypes.data, sizes.ctypes.data,
                                     None, None, None, status.ctypes.data))
        hu = (time.perf_counter() - t0) * 1e3
        ok = ok and bool((status == 0).all()) and w.stats().checksum == before
    pm, um = float(np.median(pms[1:])), float(np.median(ums[1:]))
    return {"workload": f"{side}x{side} basin world (kind {kind}), {n} chunks of 128x128, "
                        "no layer2/background",
            "pack_ms": round(pm, 3), "unpack_ms": round(um, 3),
            "pack_chunks_per_s": round(n / (pm / 1e3)), "unpack_chunks_per_s": round(n / (um / 1e3)),
            "pack_raw_GBs": round(n * PACK_RAW / (pm / 1e3) / 1e9, 1),
            "unpack_raw_GBs": round(n * PACK_RAW / (um / 1e3) / 1e9, 1),
            "image_bytes": total, "ratio": round(n * PACK_RAW / total, 3), "round_trip_exact": ok,
            # algorithmic HBM bytes per call: the 8-byte cells read and the images written (pack), the
            # images read and the cells written (unpack), over the whole call's device time
            "roofline": {"bound": "issue (serial parse per record row / per LZ4 block)", "peak": HBM_PEAK_GBS,
                         "unit": "GB/s",
                         "pack_achieved": round((side * side * 8 + total) / (pm / 1e3) / 1e9, 1),
                         "unpack_achieved": round((side * side * 8 + total) / (um / 1e3) / 1e9, 1),
                         "pack_frac": round((side * side * 8 + total) / (pm / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                         "unpack_frac": round((side * side * 8 + total) / (um / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
            "host_pack_download_ms": round(hp, 1), "host_unpack_from_host_ms": round(hu, 1),
            "host_note": "wall time through pageable host memory (PCIe copies of the images included)"}


def cpu_pack_baseline(side=2048, kind=2, seed=41, nchunks=256):
    """ChunkWrite / ChunkRead on the host, 1 thread, on every chunk of a
    smaller world of the same kind (the same shares of terrain, water and
    air as the device workload): the records loop (chunk.cpp:238-242, the oracle's
    restatement) plus the reference's own LZ4_compress_fast(..., 10) and
    LZ4_decompress_safe (oracle/_ref, built from the reference's lz4.c)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import lz4ref
    import falling_sand_engine_amd as fsg
    from oracle import pyoracle
    if lz4ref.lib() is None:
        return None
    cells = fsg.generate(kind, side, side, seed)
    per = side // 128
    chunks = [np.ascontiguousarray(cells[(k // per) * 128:(k // per + 1) * 128, (k % per) * 128:(k % per + 1) * 128])
              for k in range(nchunks)]
    t0 = time.perf_counter()
    imgs = []
    for t in chunks:
        ra, rb = pyoracle.chunk_records(t)
        imgs.append((lz4ref.compress_fast(ra, 10), lz4ref.compress_fast(rb, 10)))
    dp = time.perf_counter() - t0
    t0 = time.perf_counter()
    for za, zb in imgs:
        ra = lz4ref.decompress_safe(za, 393216)
        rb = lz4ref.decompress_safe(zb, 65536)
        assert ra[0] == 393216 and rb[0] == 65536
    du = time.perf_counter() - t0
    zbytes = sum(len(a) + len(b) + 17 for a, b in imgs)
    return {"pack_ref_1_thread_chunks_per_s": round(nchunks / dp, 1),
            "unpack_ref_1_thread_chunks_per_s": round(nchunks / du, 1),
            "pack_ref_ratio": round(nchunks * PACK_RAW / zbytes, 3),
            "pack_sample": f"{nchunks} chunks of a {side}x{side} basin world (kind {kind}); records loop "
                           "restated, LZ4 from the reference's lz4.c; decode excludes the cell copy"}


def cpu_baseline(threads=None, side_w=8192, side_h=4096, ticks=64):
    """refseq (the reference's sequential chunk-checkerboard tick, restated)
    on a bounded sample of the all-active world, on the host's cores; also
    its 1-thread rate and the deterministic oracle's (1 thread) on smaller
    samples of the same world."""
    import falling_sand_engine_amd as fsg
    from oracle import pyoracle
    hw = os.cpu_count() or 1
    threads = threads or min(16, hw)
    cells = fsg.generate(3, side_w, side_h, 3)
    cfg = pyoracle.default_config(reactions=0, temp_every=1)
    t0 = time.perf_counter()
    pyoracle.refseq(cells, cfg, 3, ticks, threads)
    dt = time.perf_counter() - t0
    small = fsg.generate(3, 2048, 2048, 3)
    t0 = time.perf_counter()
    pyoracle.refseq(small.copy(), cfg, 3, 16, 1)
    dt1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    pyoracle.run(small.copy(), cfg, 3, 0, 16)
    dto = time.perf_counter() - t0
    # texture pass, restated per cell as game.cpp:1994-2067 does it (1 thread)
    rp = fsg.RenderParams()
    rp.layers = fsg.LAYER_CELLS | fsg.LAYER_EMISSION | fsg.LAYER_FIRE
    rp.fire_mat = fsg.MAT_WATER
    rp.count_moved = 1
    for m, e in enumerate(fsg.EMIT_DEFAULT):
        rp.emit[m] = e
    prev = fsg.generate(3, 4096, 4096, 3)
    cur = pyoracle.run(prev.copy(), cfg, 3, 0, 1)
    t0 = time.perf_counter()
    pyoracle.render(cur, cfg, rp, prev=prev)
    dtr = time.perf_counter() - t0
    # island check as the reference runs it: one seed at a time, a W*H visited
    # array cleared per call, recursive fill (restated iteratively)
    import numpy as np
    frag = fragments(4096, 9)
    rng = np.random.default_rng(9)
    iseeds = [(int(x), int(y)) for x, y in zip(rng.integers(0, 4096, 64), rng.integers(0, 4096, 64))]
    t0 = time.perf_counter()
    pyoracle.physics_check(frag, pyoracle.default_config(), iseeds)
    dti = time.perf_counter() - t0
    packs = cpu_pack_baseline() or {}
    return {"value": side_w * side_h * ticks / dt / 1e9, "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"refseq restatement of world::tick, all-active striped world {side_w}x{side_h}, "
                      f"{ticks} ticks, {threads} threads, {dt:.2f} s",
            "host_threads": hw,
            "refseq_1_thread": round(2048 * 2048 * 16 / dt1 / 1e9, 4),
            "det_oracle_1_thread": round(2048 * 2048 * 16 / dto / 1e9, 4),
            "small_sample": "2048x2048 striped world, 16 ticks",
            "render_oracle_1_thread_gcell_s": round(4096 * 4096 / dtr / 1e9, 4),
            "render_sample": "texture pass, three layers + census, 4096x4096 striped world, 1 thread",
            "islands_oracle_1_thread_seeds_per_s": round(64 / dti, 1),
            "islands_sample": "physicsCheck restated, 64 seeds one at a time, 4096x4096 fragment world; "
                              "the oracle makes two passes (a dry run sizes the body images)",
            **packs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--side", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--mode", default="all", choices=["all", "c1", "c2", "c3", "c4", "copy", "c5", "render", "islands", "pack",
                                                      "mesh", "particles"],
                    help="all: the JSON line; c3: only the timed tick loop (for profiling); "
                         "copy: the HBM copy calibration kernel; render: only the texture pass")
    ap.add_argument("--copy-bytes", type=int, default=2 << 30)
    ap.add_argument("--rate", type=float, default=0.0, help="c5 mode: one offered rate (req/s) instead of the set")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1 "
                                   "(the multi-GPU sharding is a design, not built)"}))
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return
    import falling_sand_engine_amd as fsg

    if args.mode == "copy":
        with fsg.Context(0) as ctx:
            ms = ctx.copy_timed(args.copy_bytes, max(1, args.steps))
        gbs = 2 * args.copy_bytes / (ms / 1e3) / 1e9
        print(json.dumps({"mode": "copy", "bytes": args.copy_bytes, "ms": ms, "GBs_read_plus_write": gbs}))
        return
    if args.mode == "c5":
        print(json.dumps(bench_c5(fsg, rates=(args.rate,)) if args.rate > 0 else bench_c5(fsg)))
        return
    if args.mode in ("c1", "c2", "c4"):
        print(json.dumps({"c1": bench_c1, "c2": bench_c2, "c4": bench_c4}[args.mode](fsg)))
        return
    if args.mode == "islands":
        print(json.dumps(bench_islands(fsg)))
        return
    if args.mode == "mesh":
        print(json.dumps(bench_mesh(fsg)))
        return
    if args.mode == "particles":
        print(json.dumps(bench_particles(fsg)))
        return
    if args.mode == "pack":
        r = bench_pack(fsg, side=args.side)
        if not args.no_cpu:
            r["cpu"] = cpu_pack_baseline()
        print(json.dumps(r))
        return
    if args.mode == "render":
        print(json.dumps(bench_render(fsg, side=args.side, reps=max(1, args.steps))))
        return
    if args.mode == "c3":
        r = bench_c3(fsg, args.steps, args.warmup, args.side)
        print(json.dumps({"mode": "c3", "steps": args.steps, "dev_ms": r["dev_ms"], "wall_s": r["wall_s"]}))
        return

    log(f"[bench] C3 {args.side}^2 warmup={args.warmup} steps={args.steps}")
    r = bench_c3(fsg, args.steps, args.warmup, args.side)
    cell_updates = r["cells"] * args.steps
    value = cell_updates / r["wall_s"] / 1e9
    kernel_s = r["dev_ms"] / 1e3 / args.steps
    achieved = BYTES_PER_CELL_UPDATE * r["cells"] / kernel_s / 1e9
    log(f"[bench] wall {r['wall_s']:.4f}s  device {r['dev_ms']:.2f}ms  {value:.1f} Gcell/s")

    secondary = {}
    if not args.no_secondary:
        log("[bench] C1 512^2 world")
        secondary["c1_small_world"] = bench_c1(fsg)
        log(f"[bench] {secondary['c1_small_world']}")
        log("[bench] C2 dirty-chunk world")
        secondary["c2_dirty_chunks"] = bench_c2(fsg)
        log(f"[bench] {secondary['c2_dirty_chunks']}")
        log("[bench] C4 uneven batch")
        secondary["c4_batch"] = bench_c4(fsg)
        log(f"[bench] {secondary['c4_batch']}")
        log("[bench] C5 serving")
        secondary["c5_serving"] = bench_c5(fsg)
        log(f"[bench] {secondary['c5_serving']}")
        log("[bench] texture pass")
        secondary["render"] = bench_render(fsg)
        log(f"[bench] {secondary['render']}")
        log("[bench] island check")
        secondary["islands"] = bench_islands(fsg)
        log(f"[bench] {secondary['islands']}")
        log("[bench] chunk outlines")
        secondary["mesh"] = bench_mesh(fsg)
        log(f"[bench] {secondary['mesh']}")
        log("[bench] particles")
        secondary["particles"] = bench_particles(fsg)
        log(f"[bench] {secondary['particles']}")
        log("[bench] chunk .pack codec")
        secondary["pack"] = bench_pack(fsg)
        log(f"[bench] {secondary['pack']}")
        with fsg.Context(0) as ctx:
            cms = ctx.copy_timed(2 << 30, 10)
        secondary["copy_kernel_GBs"] = round(2 * (2 << 30) / (cms / 1e3) / 1e9, 1)

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline()
        log(f"[bench] cpu {cpu['value']:.4f} Gcell/s")

    out = {
        "metric": "Gcell-updates/s (16384^2 all-active world)",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall_s"] * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"all-active striped world {args.side}x{args.side}, reactions off, "
                               f"temperature on, {args.steps} ticks",
                   "cells": r["cells"], "cell_iter": 3, "chunk": "64x64", "parallelism": "none (1 GPU)"},
        # bound: what the kernel's counters say limits it (the summary's "bound" line: "valu" while
        # VALU issue takes most of the launch); achieved / peak are against the HBM roofline
        "roofline": {"bound": profile_field(args.side, "bound") or "hbm", "roofline": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(args.side),
                     "traffic_source": "profiles/k_tick_world_16384_all_active.md (rocprofv3 FETCH_SIZE x2.00 "
                                       "+ WRITE_SIZE x1.00, bytes per launch)",
                     "kernel": "k_tick_world", "kernel_avg_ms": round(kernel_s * 1e3, 4)},
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
