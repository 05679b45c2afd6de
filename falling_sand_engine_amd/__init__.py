"""MI355X-native falling-sand tick: Python host mirror of the C-ABI.

The product is ``lib/libfsg.so`` (hand-written HIP kernels for gfx950 behind
the C-ABI in ``include/fsg.h``).  This module is a thin ctypes layer with the
reference's ``ME::world`` method names (source/engine/world.hpp:148-192):
``World.tick``, ``World.tickTemperature``, ``World.getTile``, ``World.setTile``
plus bulk upload/download, statistics, rigid-body stamping and batching.

Cells are 8-byte values, exposed to numpy as ``uint64``:
bits 0-7 material id, bit 8 MOVED, bits 16-31 temperature (int16, 1/16 degree),
bits 32-63 RGBA colour.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import (AIR, GAS, LAYER_CELLS, LAYER_EMISSION, LAYER_FIRE, MOVED, SAND, SOLID, SOUP,  # noqa: F401
                   Body, Cell, CellEdit, Config, FsgError, PackMap, RenderParams, RenderStats, Stats,
                   StampResult, check, lib)

__all__ = ["Context", "World", "default_config", "generate", "make_cell", "cell_mat",
           "cell_temp", "cell_rgba", "batch_step", "FsgError", "Body", "Config"]

MAT_AIR, MAT_STONE, MAT_SAND, MAT_WATER, MAT_OIL, MAT_STEAM, MAT_LAVA, MAT_WOOD = range(8)
MATERIAL_NAMES = ["air", "stone", "sand", "water", "oil", "steam", "lava", "wood"]

# Material::emitColor of the benchmark materials, 0xAARRGGBB: water and lava
# glow (game_datastruct.cpp:91-92), the others emit nothing.
EMIT_DEFAULT = [0, 0, 0, 0x3000AFB5, 0, 0, 0xFFFF6900, 0] + [0] * 8
NO_FIRE = 255


def default_config(**overrides) -> Config:
    """The 8-material benchmark table (fsg_default_config) with field overrides."""
    cfg = Config()
    check(lib().fsg_default_config(C.byref(cfg)), "fsg_default_config")
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


# Material ids of the reference's own material list, in registration order
# (game_datastruct.cpp:72-110): the indices its .pack records hold.
REFERENCE_MATERIAL_IDS = {
    "GENERIC_AIR": 0, "GENERIC_SOLID": 1, "GENERIC_SAND": 2, "GENERIC_LIQUID": 3, "GENERIC_GAS": 4,
    "GENERIC_PASSABLE": 5, "GENERIC_OBJECT": 6, "STONE": 7, "GRASS": 8, "DIRT": 9, "SMOOTH_STONE": 10,
    "COBBLE_STONE": 11, "SMOOTH_DIRT": 12, "COBBLE_DIRT": 13, "SOFT_DIRT": 14, "WATER": 15, "LAVA": 16,
    "CLOUD": 17, "GOLD_ORE": 18, "GOLD_MOLTEN": 19, "GOLD_SOLID": 20, "IRON_ORE": 21, "OBSIDIAN": 22,
    "STEAM": 23, "SOFT_DIRT_SAND": 24, "FIRE": 25, "FLAT_COBBLE_STONE": 26, "FLAT_COBBLE_DIRT": 27,
}


def pack_map_default(cfg: Config = None) -> PackMap:
    """index = material id (fsg_pack_map_default)."""
    m = PackMap()
    check(lib().fsg_pack_map_default(C.byref(cfg if cfg is not None else default_config()), C.byref(m)),
          "fsg_pack_map_default")
    return m


def reference_pack_map(keep_moved: bool = True) -> PackMap:
    """Map between the benchmark materials and the reference's material
    list, so .pack files carry the reference's indices.  Written: air ->
    GENERIC_AIR, stone -> STONE, sand -> GENERIC_SAND, water -> WATER, oil ->
    GENERIC_LIQUID (the reference has no oil), steam -> STEAM, lava -> LAVA,
    wood -> GENERIC_SOLID (no wood either).  Read: those indices back, and
    every other reference material by its PhysicsType (game_datastruct.cpp:
    72-110): SOLID and OBJECT kinds as stone (OBSIDIAN is the stone the
    water+lava reaction makes here), SAND kinds as sand, GOLD_MOLTEN as lava,
    GENERIC_GAS as steam, PASSABLE kinds (FIRE) as air."""
    R = REFERENCE_MATERIAL_IDS
    m = PackMap()
    write = {MAT_AIR: "GENERIC_AIR", MAT_STONE: "STONE", MAT_SAND: "GENERIC_SAND", MAT_WATER: "WATER",
             MAT_OIL: "GENERIC_LIQUID", MAT_STEAM: "STEAM", MAT_LAVA: "LAVA", MAT_WOOD: "GENERIC_SOLID"}
    for i in range(256):
        m.from_file[i] = -1
    for mid, name in write.items():
        m.to_file[mid] = R[name]
        m.from_file[R[name]] = mid
    for name in ("SMOOTH_STONE", "COBBLE_STONE", "SMOOTH_DIRT", "COBBLE_DIRT", "SOFT_DIRT", "CLOUD",
                 "GOLD_SOLID", "OBSIDIAN", "FLAT_COBBLE_STONE", "FLAT_COBBLE_DIRT", "GENERIC_OBJECT"):
        m.from_file[R[name]] = MAT_STONE
    for name in ("GRASS", "DIRT", "GOLD_ORE", "IRON_ORE", "SOFT_DIRT_SAND"):
        m.from_file[R[name]] = MAT_SAND
    m.from_file[R["GOLD_MOLTEN"]] = MAT_LAVA
    m.from_file[R["GENERIC_GAS"]] = MAT_STEAM
    m.from_file[R["GENERIC_PASSABLE"]] = MAT_AIR
    m.from_file[R["FIRE"]] = MAT_AIR
    m.flags = 0 if keep_moved else _abi.PACK_NO_MOVED
    return m


def make_particle(cell, x, y, vx=0.0, vy=0.0, ax=0.0, ay=0.0, lifetime=0, flags=0, target=None) -> _abi.Particle:
    """A CellData (game_utils/cells.h:15-36) carrying `cell` (a uint64 cell)."""
    p = _abi.Particle()
    p.x, p.y, p.vx, p.vy, p.ax, p.ay = x, y, vx, vy, ax, ay
    if target is not None:
        p.target_x, p.target_y, p.target_force = target
    p.lifetime, p.flags = lifetime, flags
    c = int(cell)
    p.tile.mat, p.tile.temp, p.tile.rgba = c & 0xFFFF, ((c >> 16) & 0xFFFF) - (0x10000 if (c >> 31) & 1 else 0), c >> 32
    return p


def particles_from_displaced(displaced, seed: int, tick: int):
    """Particles for the cells a body stamp displaced, as game.cpp:1789-1791
    ejects them: at (x, y - 3) with vx = (r % 10 - 5) / 10, vy = -(r' % 5 + 5)
    / 10 and gravity 0.1, r and r' from a counter hash of (seed, tick, index)
    in place of rand()."""
    out = []
    key = lib().fsg_phase_key(seed, tick)
    for k, (x, y, cell) in enumerate(displaced):
        r = lib().fsg_splitmix64(key ^ (k + 1))
        vx = ((r & 0xFFFFFFFF) % 10 - 5) / 10.0
        vy = -(((r >> 32) % 5) + 5) / 10.0
        out.append(make_particle(cell, float(x), float(y - 3), vx, vy, 0.0, 0.1))
    return out


def make_cell(mat, temp=320, rgba=0):
    return (int(mat) & 0xFFFF) | ((int(temp) & 0xFFFF) << 16) | ((int(rgba) & 0xFFFFFFFF) << 32)


def cell_mat(a):
    return (np.asarray(a, dtype=np.uint64) & np.uint64(0xFF)).astype(np.uint8)


def cell_temp(a):
    return ((np.asarray(a, dtype=np.uint64) >> np.uint64(16)) & np.uint64(0xFFFF)).astype(np.uint16).view(np.int16)


def cell_rgba(a):
    return (np.asarray(a, dtype=np.uint64) >> np.uint64(32)).astype(np.uint32)


def generate(kind: int, width: int, height: int, seed: int) -> np.ndarray:
    """Benchmark world generator (host code in libfsg): returns (H, W) uint64."""
    out = np.empty((height, width), dtype=np.uint64)
    check(lib().fsg_generate(kind, width, height, seed, out.ctypes.data), "fsg_generate")
    return out


class Context:
    """One device, one HIP stream, an HBM budget (replaces WorldSystem's pools)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        cfg = _abi.CtxConfig(device, hbm_budget)
        h = C.c_void_p()
        check(lib().fsg_ctx_create(C.byref(cfg), C.byref(h)), "fsg_ctx_create")
        self.handle = h

    def usage(self) -> int:
        v = C.c_uint64()
        check(lib().fsg_ctx_usage(self.handle, C.byref(v)), "fsg_ctx_usage")
        return v.value

    def timer_start(self):
        """HIP event on the ctx stream; timer_stop() returns the device ms since."""
        check(lib().fsg_ctx_timer_start(self.handle), "fsg_ctx_timer_start")

    def timer_stop(self) -> float:
        ms = C.c_double()
        check(lib().fsg_ctx_timer_stop(self.handle, C.byref(ms)), "fsg_ctx_timer_stop")
        return ms.value

    def inject_alloc_failures(self, n: int):
        check(lib().fsg_ctx_inject_alloc_failures(self.handle, n))

    def tuning(self) -> dict:
        t = _abi.Tuning()
        check(lib().fsg_ctx_get_tuning(self.handle, C.byref(t)), "fsg_ctx_get_tuning")
        return {"graph_ticks": t.graph_ticks, "chunk_skip": t.chunk_skip, "occupancy": t.occupancy}

    def set_tuning(self, graph_ticks: int = 16, chunk_skip: int = -1, occupancy: int = 5):
        t = _abi.Tuning(graph_ticks, chunk_skip, occupancy, 0)
        check(lib().fsg_ctx_set_tuning(self.handle, C.byref(t)), "fsg_ctx_set_tuning")

    def autotune(self, worlds, sample_ticks: int = 32) -> dict:
        """Times every batch tuning on copies of `worlds` (which do not
        advance) and keeps the fastest; returns the report."""
        rep = _abi.TuneReport()
        check(lib().fsg_autotune_batch(self.handle, _handles(worlds), len(worlds), sample_ticks, C.byref(rep)),
              "fsg_autotune_batch")

        def td(t):
            return {"graph_ticks": t.graph_ticks, "chunk_skip": t.chunk_skip, "occupancy": t.occupancy}
        return {"best": td(rep.best), "from_cache": bool(rep.from_cache), "size_class": rep.size_class,
                "candidates": [dict(td(rep.candidate[i]), ms=rep.ms[i]) for i in range(rep.n_candidates)]}

    def copy_timed(self, nbytes: int, reps: int = 10) -> float:
        """Mean ms of a plain device copy of nbytes (HBM calibration)."""
        ms = C.c_double()
        check(lib().fsg_ctx_copy_timed(self.handle, nbytes, reps, C.byref(ms)), "fsg_ctx_copy_timed")
        return ms.value

    def close(self):
        if self.handle:
            check(lib().fsg_ctx_destroy(self.handle), "fsg_ctx_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


MIRROR_PAGE = 64  # side of the getTile mirror's pages


class World:
    """Device-resident world with the reference's world API (world.hpp:63-193)."""

    def __init__(self, ctx: Context, width: int, height: int, seed: int = 0, config: Config = None):
        self.ctx = ctx
        self.width, self.height, self.seed = width, height, seed
        self.config = config if config is not None else default_config()
        h = C.c_void_p()
        check(lib().fsg_world_create(ctx.handle, width, height, seed, C.byref(self.config), C.byref(h)),
              "fsg_world_create")
        self.handle = h
        self._edits = []
        self._pages = {}  # host mirror for getTile: (px, py) -> MIRROR_PAGE^2 cells

    # --- reference method names -------------------------------------------
    def tick(self, n: int = 1):
        """world::tick (world.cpp:1036) n times, temperature by cadence."""
        self._drop_mirror()
        self._flush()
        check(lib().fsg_world_step(self.handle, n), "fsg_world_step")

    def tickTemperature(self):  # noqa: N802 (reference name)
        """world::tickTemperature (world.cpp:1950)."""
        self._drop_mirror()
        self._flush()
        check(lib().fsg_world_tick_temperature(self.handle), "fsg_world_tick_temperature")

    @property
    def tickZone(self):  # noqa: N802 (reference name)
        """World::tickZone (world.hpp:138) as (x, y, w, h): the rectangle that
        ticks.  A block acts iff one of its cells lies inside; only zone cells
        take the temperature step and only particles in it move."""
        v = [C.c_int32() for _ in range(4)]
        check(lib().fsg_world_get_tick_zone(self.handle, *[C.byref(x) for x in v]), "fsg_world_get_tick_zone")
        return tuple(x.value for x in v)

    @tickZone.setter
    def tickZone(self, zone):  # noqa: N802
        x, y, w, h = (int(v) for v in zone)
        self._flush()
        check(lib().fsg_world_set_tick_zone(self.handle, x, y, w, h), "fsg_world_set_tick_zone")

    def getTile(self, x: int, y: int) -> int:  # noqa: N802
        """world::getTile (world.cpp:999): out of bounds reads as a solid wall.
        Reads come from a host mirror of MIRROR_PAGE^2-cell pages: a page is
        downloaded on its first read (one stream sync) and kept until a call
        changes the grid on the device (tick, stamp, upload, ...), so the
        reference's per-cell readers do not wait on the device per cell."""
        if x < 0 or y < 0 or x >= self.width or y >= self.height:
            return 15
        key = (x // MIRROR_PAGE, y // MIRROR_PAGE)
        page = self._pages.get(key)
        if page is None:
            x0, y0 = key[0] * MIRROR_PAGE, key[1] * MIRROR_PAGE
            page = self.download(x0, y0, min(MIRROR_PAGE, self.width - x0), min(MIRROR_PAGE, self.height - y0))
            self._pages[key] = page
        return int(page[y % MIRROR_PAGE, x % MIRROR_PAGE])

    def setTile(self, x: int, y: int, cell: int):  # noqa: N802
        """world::setTile (world.cpp:1004): queued, applied before the next
        device call; a mirrored page takes the write at once, so getTile sees
        it as the reference's does.  Out of bounds is ignored."""
        if x < 0 or y < 0 or x >= self.width or y >= self.height:
            return
        self._edits.append((x, y, int(cell)))
        page = self._pages.get((x // MIRROR_PAGE, y // MIRROR_PAGE))
        if page is not None:
            page[y % MIRROR_PAGE, x % MIRROR_PAGE] = np.uint64(int(cell) & 0xFFFFFFFFFFFFFFFF)

    def prefetch(self, x0: int = 0, y0: int = 0, w: int = None, h: int = None):
        """Fills the getTile mirror for the pages that meet the rectangle, in
        one download."""
        w = self.width - x0 if w is None else w
        h = self.height - y0 if h is None else h
        if w <= 0 or h <= 0:
            return
        P = MIRROR_PAGE
        px0, py0 = max(x0, 0) // P, max(y0, 0) // P
        px1, py1 = (min(x0 + w, self.width) - 1) // P, (min(y0 + h, self.height) - 1) // P
        ax, ay = px0 * P, py0 * P
        bw, bh = min((px1 + 1) * P, self.width) - ax, min((py1 + 1) * P, self.height) - ay
        blk = self.download(ax, ay, bw, bh)
        for py in range(py0, py1 + 1):
            for px in range(px0, px1 + 1):
                self._pages[(px, py)] = blk[py * P - ay:min((py + 1) * P, self.height) - ay,
                                            px * P - ax:min((px + 1) * P, self.width) - ax].copy()

    def _drop_mirror(self):
        self._pages.clear()

    # --- bulk / state -----------------------------------------------------------
    def _flush(self):
        if not self._edits:
            return
        arr = (CellEdit * len(self._edits))()
        for i, (x, y, v) in enumerate(self._edits):
            arr[i].x, arr[i].y = x, y
            arr[i].cell.mat = v & 0xFFFF
            arr[i].cell.temp = C.c_int16(v >> 16 & 0xFFFF).value
            arr[i].cell.rgba = (v >> 32) & 0xFFFFFFFF
        self._edits = []
        check(lib().fsg_world_set_cells(self.handle, arr, len(arr)), "fsg_world_set_cells")

    def upload(self, cells: np.ndarray):
        self._drop_mirror()
        a = np.ascontiguousarray(cells, dtype=np.uint64)
        assert a.size == self.width * self.height
        check(lib().fsg_world_upload(self.handle, a.ctypes.data, a.size), "fsg_world_upload")

    def generate(self, kind: int, seed: int):
        """Fills the grid with generate(kind, width, height, seed), on the device."""
        self._drop_mirror()
        self._edits = []
        check(lib().fsg_world_generate(self.handle, kind, seed), "fsg_world_generate")

    def download(self, x0=0, y0=0, w=None, h=None) -> np.ndarray:
        self._flush()
        w = self.width if w is None else w
        h = self.height if h is None else h
        out = np.empty((h, w), dtype=np.uint64)
        check(lib().fsg_world_download(self.handle, out.ctypes.data, x0, y0, w, h), "fsg_world_download")
        return out

    def step_timed(self, n: int) -> float:
        self._drop_mirror()
        self._flush()
        ms = C.c_double()
        check(lib().fsg_world_step_timed(self.handle, n, C.byref(ms)), "fsg_world_step_timed")
        return ms.value

    def sync(self):
        check(lib().fsg_world_sync(self.handle))

    @property
    def tick_count(self) -> int:
        t = C.c_uint64()
        check(lib().fsg_world_info(self.handle, None, None, C.byref(t)))
        return t.value

    def set_tick(self, t: int):
        self._drop_mirror()
        check(lib().fsg_world_set_tick(self.handle, t))

    def stats(self) -> Stats:
        s = Stats()
        check(lib().fsg_world_stats(self.handle, C.byref(s)), "fsg_world_stats")
        return s

    def reset_activity(self):
        check(lib().fsg_world_reset_activity(self.handle))

    # --- rigid bodies -----------------------------------------------------------
    def stamp_bodies(self, bodies, pixels: np.ndarray, displaced_cap: int = 0):
        self._drop_mirror()
        nb = len(bodies)
        barr = (Body * max(nb, 1))(*bodies)
        px = np.ascontiguousarray(pixels, dtype=np.uint64)
        res = StampResult()
        disp = (CellEdit * max(displaced_cap, 1))()
        hs = (C.c_uint32 * max(nb, 1))()
        hw = (C.c_uint32 * max(nb, 1))()
        res.displaced_cap = displaced_cap
        res.displaced_out = C.cast(disp, C.POINTER(CellEdit))
        res.hits_sand = C.cast(hs, C.POINTER(C.c_uint32))
        res.hits_soup = C.cast(hw, C.POINTER(C.c_uint32))
        check(lib().fsg_world_stamp_bodies(self.handle, barr, nb, px.ctypes.data, px.size, C.byref(res)),
              "fsg_world_stamp_bodies")
        n = min(res.displaced, displaced_cap)
        displaced = [(disp[i].x, disp[i].y,
                      make_cell(disp[i].cell.mat, disp[i].cell.temp, disp[i].cell.rgba)) for i in range(n)]
        return {"placed": res.placed, "unplaced": res.unplaced, "displaced": res.displaced,
                "displaced_cells": displaced, "hits_sand": list(hs)[:nb], "hits_soup": list(hw)[:nb]}

    def unstamp_bodies(self, npixels: int) -> np.ndarray:
        self._drop_mirror()
        out = np.empty(npixels, dtype=np.uint64)
        check(lib().fsg_world_unstamp_bodies(self.handle, out.ctypes.data, npixels), "fsg_world_unstamp_bodies")
        return out

    # --- island check: world::physicsCheck (world.cpp:3330-3429) ---------------
    def physicsCheck(self, seeds):  # noqa: N802
        """Runs the island check for seed cells [(x, y), ...] in order.  Returns
        one dict per seed: status (ISLAND_*), count, box (min_x, min_y, w, h) and,
        for a new body, its (h, w) uint64 pixel image (air outside the island)."""
        self._drop_mirror()
        self._flush()
        n = len(seeds)
        arr = (_abi.Island * max(n, 1))()
        for i, (x, y) in enumerate(seeds):
            arr[i].x, arr[i].y = int(x), int(y)
        need = C.c_size_t()
        check(lib().fsg_world_physics_check(self.handle, arr, n, C.byref(need)), "fsg_world_physics_check")
        self._islands = (arr, n)
        pix = np.empty(need.value, dtype=np.uint64)
        if need.value:
            check(lib().fsg_world_island_pixels(self.handle, pix.ctypes.data, pix.size), "fsg_world_island_pixels")
        out = []
        for i in range(n):
            r = arr[i]
            d = {"status": r.status, "count": r.count, "box": (r.min_x, r.min_y, r.w, r.h)}
            if r.status == _abi.ISLAND_BODY:
                d["pixels"] = pix[r.pixel_offset:r.pixel_offset + r.w * r.h].reshape(r.h, r.w).copy()
            out.append(d)
        return out

    # --- texture pass (game.cpp:1994-2067) -----------------------------------
    def render(self, layers=LAYER_CELLS, rect=None, fire_mat=NO_FIRE, emit=None, stats=False,
               count_moved=True):
        """Updates the RGBA8 layers over rect = (x0, y0, w, h) (whole world
        by default).  With stats, waits and returns the movingTiles census."""
        self._flush()
        p = RenderParams()
        p.layers = layers
        p.x0, p.y0, p.w, p.h = rect if rect is not None else (0, 0, 0, 0)
        p.fire_mat = fire_mat
        p.count_moved = 1 if count_moved else 0
        for m, e in enumerate(EMIT_DEFAULT if emit is None else emit):
            p.emit[m] = e
        st = RenderStats()
        check(lib().fsg_world_render(self.handle, C.byref(p), C.byref(st) if stats else None),
              "fsg_world_render")
        if stats:
            return {"moved": [int(v) for v in st.moved], "fire_cells": int(st.fire_cells),
                    "had_dirty": bool(st.had_dirty)}
        return None

    def pack(self, x0, y0, ncx, ncy, phase=None, layer2=None, background=None, download=True):
        """Chunk .pack images (Chunk::ChunkWrite, chunk.cpp:219-330) of the
        ncx x ncy chunks whose top-left cell is (x0, y0).  layer2 (uint64) and
        background (uint32) are region-shaped arrays or None; phase is one int8
        per chunk or None.  Returns a list of bytes objects (download=True) or
        the image sizes (the images stay on the device for unpack(None))."""
        self._flush()
        n = ncx * ncy
        sizes = np.zeros(n, dtype=np.uintp)
        ph = None if phase is None else np.ascontiguousarray(phase, dtype=np.int8)
        l2 = None if layer2 is None else np.ascontiguousarray(layer2, dtype=np.uint64)
        bg = None if background is None else np.ascontiguousarray(background, dtype=np.uint32)
        shape = (ncy * _abi.CHUNK, ncx * _abi.CHUNK)
        assert ph is None or ph.size == n
        assert l2 is None or l2.shape == shape
        assert bg is None or bg.shape == shape
        check(lib().fsg_world_pack(self.handle, x0, y0, ncx, ncy, None if ph is None else ph.ctypes.data,
                                   None if l2 is None else l2.ctypes.data, None if bg is None else bg.ctypes.data,
                                   sizes.ctypes.data), "fsg_world_pack")
        if not download:
            return [int(v) for v in sizes]
        buf = np.empty(max(int(sizes.sum()), 1), dtype=np.uint8)
        check(lib().fsg_world_pack_download(self.handle, buf.ctypes.data, buf.size), "fsg_world_pack_download")
        out, o = [], 0
        for v in sizes:
            out.append(buf[o:o + int(v)].tobytes())
            o += int(v)
        return out

    def _mesh_result(self, regs, nl, nv):
        loops = (_abi.MeshLoop * max(nl.value, 1))()
        verts = np.zeros(max(2 * nv.value, 2), np.float32)
        check(lib().fsg_world_mesh_download(self.handle, loops, nl.value, verts.ctypes.data, nv.value),
              "fsg_world_mesh_download")
        out = []
        for R in regs:
            ls = []
            for k in range(R.first_loop, R.first_loop + R.nloops):
                L = loops[k]
                v = verts[2 * L.first:2 * (L.first + L.count)].reshape(-1, 2)
                ls.append((L.start_x, L.start_y, bool(L.hole), [(float(a), float(b)) for a, b in v]))
            out.append((R.status, ls))
        return out

    def mesh(self, regions, raw=False):
        """Outlines of the SOLID cells of grid regions [(x0, y0, w, h), ...]
        (world::updateChunkMesh's marching squares, world.cpp:722-895): per
        region (status, [(start_x, start_y, hole, [(x, y), ...]), ...]).
        raw=True returns the C-ABI's arrays instead: (regions, loops,
        vertices as an (n, 2) float32 array), with no per-vertex objects."""
        self._flush()
        regs = (_abi.MeshRegion * max(len(regions), 1))()
        for i, (x0, y0, w, h) in enumerate(regions):
            regs[i].x0, regs[i].y0, regs[i].w, regs[i].h = x0, y0, w, h
        nl, nv = C.c_uint32(), C.c_uint32()
        check(lib().fsg_world_mesh(self.handle, regs, len(regions), C.byref(nl), C.byref(nv)), "fsg_world_mesh")
        if raw:
            loops = (_abi.MeshLoop * max(nl.value, 1))()
            verts = np.zeros((max(nv.value, 1), 2), np.float32)
            check(lib().fsg_world_mesh_download(self.handle, loops, nl.value, verts.ctypes.data, nv.value),
                  "fsg_world_mesh_download")
            return regs, loops, verts[:nv.value]
        return self._mesh_result(list(regs)[:len(regions)], nl, nv)

    def mesh_islands(self):
        """Outlines of the body images of the last physicsCheck (the marching
        squares of updateRigidBodyHitbox, world.cpp:288-510), one entry per
        seed as mesh() returns them."""
        islands, n = self._islands
        regs = (_abi.MeshRegion * max(n, 1))()
        nl, nv = C.c_uint32(), C.c_uint32()
        check(lib().fsg_world_mesh_islands(self.handle, islands, n, regs, C.byref(nl), C.byref(nv)),
              "fsg_world_mesh_islands")
        return self._mesh_result(list(regs)[:n], nl, nv)

    # --- free-fall particles (CellData, world::tickCells) --------------------
    def particles_add(self, particles):
        """Appends Particle structs (see make_particle) to the world's list."""
        self._flush()
        arr = (_abi.Particle * max(len(particles), 1))(*particles)
        check(lib().fsg_world_particles_add(self.handle, arr, len(particles)), "fsg_world_particles_add")

    def particles_tick(self):
        """One world::tickCells pass; returns (particles left, cells landed)."""
        self._drop_mirror()
        self._flush()
        alive, landed = C.c_size_t(), C.c_size_t()
        check(lib().fsg_world_particles_tick(self.handle, C.byref(alive), C.byref(landed)),
              "fsg_world_particles_tick")
        return alive.value, landed.value

    def particles(self):
        n = C.c_size_t()
        check(lib().fsg_world_particles_get(self.handle, None, 0, C.byref(n)), "fsg_world_particles_get")
        arr = (_abi.Particle * max(n.value, 1))()
        check(lib().fsg_world_particles_get(self.handle, arr, n.value, C.byref(n)), "fsg_world_particles_get")
        return [arr[i] for i in range(n.value)]

    def set_pack_map(self, pack_map=None):
        """Material index map of later pack()/unpack() calls (a PackMap, e.g.
        reference_pack_map(); None restores index = material id)."""
        check(lib().fsg_world_set_pack_map(self.handle, None if pack_map is None else C.byref(pack_map)),
              "fsg_world_set_pack_map")

    def unpack(self, x0, y0, ncx, ncy, images=None, want_layer2=False, want_background=False):
        """Chunk::ChunkRead (chunk.cpp:74-217) of ncx x ncy images into the grid
        at (x0, y0); images None reads the device-resident images of the last
        pack().  Returns (status per chunk, phases, layer2 or None,
        background or None)."""
        self._drop_mirror()
        self._flush()
        n = ncx * ncy
        status = np.zeros(n, dtype=np.int32)
        phase = np.zeros(n, dtype=np.int8)
        shape = (ncy * _abi.CHUNK, ncx * _abi.CHUNK)
        l2 = np.zeros(shape, dtype=np.uint64) if want_layer2 else None
        bg = np.zeros(shape, dtype=np.uint32) if want_background else None
        src = sizes = None
        if images is not None:
            assert len(images) == n
            sizes = np.array([len(b) for b in images], dtype=np.uintp)
            src = np.frombuffer(b"".join(images) or b"\0", dtype=np.uint8)
        check(lib().fsg_world_unpack(self.handle, x0, y0, ncx, ncy, None if src is None else src.ctypes.data,
                                     None if sizes is None else sizes.ctypes.data, phase.ctypes.data,
                                     None if l2 is None else l2.ctypes.data, None if bg is None else bg.ctypes.data,
                                     status.ctypes.data), "fsg_world_unpack")
        return status, phase, l2, bg

    def layer(self, layer, x0=0, y0=0, w=None, h=None) -> np.ndarray:
        """Host copy of a layer rectangle as (h, w) uint32 texels (bytes R,G,B,A)."""
        w = self.width if w is None else w
        h = self.height if h is None else h
        out = np.empty((h, w), dtype=np.uint32)
        check(lib().fsg_world_layer_download(self.handle, layer, x0, y0, w, h, out.ctypes.data),
              "fsg_world_layer_download")
        return out

    def layer_ptr(self, layer) -> int:
        ptr = C.c_void_p()
        check(lib().fsg_world_layer(self.handle, layer, C.byref(ptr)), "fsg_world_layer")
        return ptr.value or 0

    def close(self):
        if self.handle:
            check(lib().fsg_world_destroy(self.handle), "fsg_world_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _handles(worlds):
    arr = (C.c_void_p * len(worlds))()
    for i, w in enumerate(worlds):
        w._flush()
        w._drop_mirror()
        arr[i] = w.handle
    return arr


def batch_step(ctx: Context, worlds, ticks: int):
    """All worlds advance `ticks` ticks in shared launches (bit-identical to
    stepping each alone)."""
    check(lib().fsg_batch_step(ctx.handle, _handles(worlds), len(worlds), ticks), "fsg_batch_step")


def batch_step_timed(ctx: Context, worlds, ticks: int) -> float:
    ms = C.c_double()
    check(lib().fsg_batch_step_timed(ctx.handle, _handles(worlds), len(worlds), ticks, C.byref(ms)),
          "fsg_batch_step_timed")
    return ms.value


def lz4_decode_blocks(ctx: Context, blocks, guard: int = 0, fill: int = 0xA5):
    """Decodes raw LZ4 blocks [(bytes, capacity), ...] on the device with the
    unpack's decode kernel (fsg_lz4_decode_blocks).  Each block's output is
    followed by `guard` bytes; all start as `fill`.  Returns [(size or -1,
    output bytes incl. the guard), ...]."""
    n = len(blocks)
    src = b"".join(b for b, _ in blocks)
    so = np.zeros(max(n, 1), np.int64)
    sb = np.array([len(b) for b, _ in blocks] or [0], np.int32)
    cap = np.array([c for _, c in blocks] or [0], np.int32)
    so[1:n] = np.cumsum(sb[:n])[:-1] if n > 1 else so[1:n]
    oo = np.zeros(max(n, 1), np.int64)
    oo[1:n] = np.cumsum(cap[:n].astype(np.int64) + guard)[:-1] if n > 1 else oo[1:n]
    total = int(cap[:n].astype(np.int64).sum()) + guard * n
    out = np.full(max(total, 1), fill, np.uint8)
    res = np.zeros(max(n, 1), np.int32)
    sbuf = np.frombuffer(src, np.uint8) if src else np.zeros(1, np.uint8)
    check(lib().fsg_lz4_decode_blocks(ctx.handle, sbuf.ctypes.data, len(src), so.ctypes.data, sb.ctypes.data,
                                      cap.ctypes.data, n, out.ctypes.data, total, oo.ctypes.data, res.ctypes.data),
          "fsg_lz4_decode_blocks")
    return [(int(res[k]), out[oo[k]:oo[k] + cap[k] + guard].tobytes()) for k in range(n)]


def perworld_step_timed(ctx: Context, worlds, ticks: int) -> float:
    ms = C.c_double()
    check(lib().fsg_perworld_step_timed(ctx.handle, _handles(worlds), len(worlds), ticks, C.byref(ms)),
          "fsg_perworld_step_timed")
    return ms.value


class Server:
    """Continuous-batching request loop (fsg_serve_*): worlds stay resident in
    HBM between requests, admitted under the context's HBM budget."""

    def __init__(self, ctx: Context, max_active: int = 0, block_ticks: int = 16, verify_resident: bool = False):
        cfg = _abi.ServeConfig(max_active, block_ticks, 1 if verify_resident else 0, 0)
        h = C.c_void_p()
        check(lib().fsg_serve_create(ctx.handle, C.byref(cfg), C.byref(h)), "fsg_serve_create")
        self.handle = h

    def submit(self, world_id: int, ticks: int, kind: int = 4, width: int = 512, height: int = 512,
               seed: int = 0, n_bodies: int = 0, deadline_ns: int = 0) -> int:
        r = _abi.ServeRequest(world_id, kind, width, height, ticks, seed, n_bodies, 0, deadline_ns)
        rid = C.c_uint64()
        check(lib().fsg_serve_submit(self.handle, C.byref(r), C.byref(rid)), "fsg_serve_submit")
        return rid.value

    def cancel(self, request_id: int) -> int:
        return lib().fsg_serve_cancel(self.handle, request_id)

    def poll(self, cap: int = 4096):
        arr = (_abi.ServeResult * cap)()
        n = C.c_int()
        check(lib().fsg_serve_poll(self.handle, arr, cap, C.byref(n)), "fsg_serve_poll")
        return [{f: getattr(arr[i], f) for f, _ in _abi.ServeResult._fields_} for i in range(n.value)]

    def drain(self, timeout_s: float = 60.0) -> int:
        return lib().fsg_serve_drain(self.handle, int(timeout_s * 1e9))

    def release(self, world_id: int):
        check(lib().fsg_serve_release(self.handle, world_id))

    def corrupt(self, world_id: int):
        check(lib().fsg_serve_corrupt(self.handle, world_id))

    def stats(self) -> dict:
        s = _abi.ServeStats()
        check(lib().fsg_serve_stats_get(self.handle, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _abi.ServeStats._fields_}

    def close(self):
        if self.handle:
            check(lib().fsg_serve_destroy(self.handle), "fsg_serve_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def now_ns() -> int:
    return lib().fsg_now_ns()


def body_from_pose(x: float, y: float, angle: float, mat_w: int, mat_h: int, pixel_offset: int,
                   enabled: bool = True) -> Body:
    """Fixed-point pose from a float pose (b2Body::GetPosition/GetAngle), done
    once per body per tick on the host."""
    import math
    return Body(int(round(x * 65536)), int(round(y * 65536)), int(round(math.cos(angle) * 65536)),
                int(round(math.sin(angle) * 65536)), mat_w, mat_h, pixel_offset, 1 if enabled else 0)
