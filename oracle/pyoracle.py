"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from falling_sand_engine_amd._abi import (Body, CellEdit, Config, Island, MeshLoop, PackMap, Particle, RenderParams,
                                          RenderStats)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

_L = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u64, u32, i32, P = C.c_uint64, C.c_uint32, C.c_int, C.c_void_p
        cfgp = C.POINTER(Config)
        sig = {
            "orc_default_config": (None, [cfgp]),
            "orc_splitmix64": (u64, [u64]),
            "orc_lowbias32": (u32, [u32]),
            "orc_lowbias32_inv": (u32, [u32]),
            "orc_phase_key": (u64, [u64, u64]),
            "orc_block_hash": (None, [u64, i32, i32, C.POINTER(u32), C.POINTER(u32)]),
            "orc_block_rule": (i32, [cfgp, i32, u32, u32, P]),
            "orc_tick": (i32, [P, i32, i32, cfgp, u64, u64]),
            "orc_run": (i32, [P, i32, i32, cfgp, u64, u64, i32]),
            "orc_run_zone": (i32, [P, i32, i32, cfgp, u64, u64, i32, P]),
            "orc_temperature": (i32, [P, i32, i32, cfgp]),
            "orc_temperature_zone": (i32, [P, i32, i32, cfgp, P]),
            "orc_region_unstable": (i32, [P, i32, i32, cfgp] + [i32] * 8),
            "orc_stamp": (i32, [P, i32, i32, cfgp, P, i32, P, C.c_size_t, P, C.POINTER(u32),
                                C.POINTER(u32), P, u32, P, P]),
            "orc_unstamp": (i32, [P, i32, i32, cfgp, P, i32, P, C.c_size_t]),
            "orc_body_pixel_xy": (i32, [P, i32, i32, C.POINTER(i32), C.POINTER(i32)]),
            "orc_refstamp": (i32, [P, i32, i32, cfgp, P, i32, P, C.c_size_t, P, C.POINTER(u32),
                                   C.POINTER(u32), P, u32]),
            "orc_refunstamp": (i32, [P, i32, i32, cfgp, P, i32, P, C.c_size_t]),
            "orc_spiral": (i32, [P]),
            "orc_particles_tick": (i32, [P, i32, i32, cfgp, P, C.c_size_t, C.POINTER(C.c_size_t),
                                         C.POINTER(C.c_size_t)]),
            "orc_particles_tick_zone": (i32, [P, i32, i32, cfgp, P, C.c_size_t, C.POINTER(C.c_size_t),
                                              C.POINTER(C.c_size_t), P]),
            "orc_ms_value": (i32, [P, i32, i32, i32, i32]),
            "orc_find_perimeter": (i32, [P, i32, i32, i32, i32, P, i32]),
            "orc_simplify": (C.c_size_t, [P, C.c_size_t, C.c_float, P]),
            "orc_mesh_mask": (i32, [P, i32, i32, P, u32, P, u32, C.POINTER(u32), C.POINTER(u32)]),
            "orc_physics_check": (i32, [P, i32, i32, cfgp, P, i32, P, C.c_size_t, C.POINTER(C.c_size_t)]),
            "orc_render": (i32, [P, P, i32, i32, cfgp, C.POINTER(RenderParams), P, P, P,
                                 C.POINTER(RenderStats)]),
            "orc_counts": (None, [P, C.c_size_t, P]),
            "orc_checksum": (u64, [P, C.c_size_t]),
            "refseq_run": (i32, [P, i32, i32, cfgp, u64, i32, i32]),
            "orc_chunk_records": (None, [P, C.c_size_t, P, C.c_size_t, P, C.c_size_t, P, P,
                                         C.POINTER(PackMap)]),
            "orc_lz4_decode": (i32, [P, i32, P, i32]),
            "orc_lz4_encode_rows": (i32, [P, i32, i32, P, i32]),
            "orc_pack_chunk": (C.c_size_t, [P, P, C.c_int8, P]),
            "orc_unpack_chunk": (i32, [P, C.c_size_t, u32, P, C.c_size_t, P, C.c_size_t, P, C.c_size_t,
                                       C.POINTER(C.c_int8), C.POINTER(PackMap)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _L = L
    return _L


def default_config(**overrides) -> Config:
    cfg = Config()
    lib().orc_default_config(C.byref(cfg))
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


def _grid(cells):
    assert cells.dtype == np.uint64 and cells.flags.c_contiguous and cells.ndim == 2
    return cells.ctypes.data, cells.shape[1], cells.shape[0]


def _zone(zone):
    if zone is None:
        return None, None
    z = np.ascontiguousarray(zone, dtype=np.int32)
    assert z.shape == (4,)
    return z, z.ctypes.data


def run(cells: np.ndarray, cfg: Config, seed: int, tick0: int, nticks: int, zone=None) -> np.ndarray:
    """Det-mode ticks tick0 .. tick0+nticks-1, in place on an (H, W) uint64
    grid; zone = (x, y, w, h) limits them to a tick zone."""
    p, W, H = _grid(cells)
    z, zp = _zone(zone)
    rc = lib().orc_run_zone(p, W, H, C.byref(cfg), seed, tick0, nticks, zp)
    assert rc == 0, rc
    return cells


def temperature(cells: np.ndarray, cfg: Config, zone=None) -> np.ndarray:
    p, W, H = _grid(cells)
    z, zp = _zone(zone)
    rc = lib().orc_temperature_zone(p, W, H, C.byref(cfg), zp)
    assert rc == 0, rc
    return cells


def refseq(cells: np.ndarray, cfg: Config, seed: int, nticks: int, threads: int = 1) -> np.ndarray:
    p, W, H = _grid(cells)
    rc = lib().refseq_run(p, W, H, C.byref(cfg), seed, nticks, threads)
    assert rc == 0, rc
    return cells


def region_unstable(cells, cfg, x0, y0, x1, y1, cx0, cy0, cx1, cy1) -> bool:
    p, W, H = _grid(cells)
    return bool(lib().orc_region_unstable(p, W, H, C.byref(cfg), x0, y0, x1, y1, cx0, cy0, cx1, cy1))


def block_rule(cfg, it, r, s, q):
    arr = (C.c_uint64 * 4)(*q)
    ch = lib().orc_block_rule(C.byref(cfg), it, r, s, arr)
    return bool(ch), list(arr)


def splitmix64(x):
    return lib().orc_splitmix64(x)


def lowbias32(x):
    return lib().orc_lowbias32(x)


def lowbias32_inv(x):
    return lib().orc_lowbias32_inv(x)


def phase_key(seed, g):
    return lib().orc_phase_key(seed, g)


def block_hash(key, x0, y0):
    r, s = C.c_uint32(), C.c_uint32()
    lib().orc_block_hash(key, x0, y0, C.byref(r), C.byref(s))
    return r.value, s.value


def counts(cells) -> np.ndarray:
    out = np.zeros(16, dtype=np.uint64)
    a = np.ascontiguousarray(cells, dtype=np.uint64)
    lib().orc_counts(a.ctypes.data, a.size, out.ctypes.data)
    return out


def checksum(cells) -> int:
    a = np.ascontiguousarray(cells, dtype=np.uint64)
    return lib().orc_checksum(a.ctypes.data, a.size)


def render(cells, cfg, params: RenderParams, layers=None, prev=None):
    """Texture pass over params' rectangle.  layers: dict LAYER_* -> (H, W)
    uint32 arrays updated in place (missing ones start at zero).  Returns
    (layers, RenderStats)."""
    ptr, W, H = _grid(cells)
    layers = {} if layers is None else layers
    for bit in (1, 2, 4):
        if bit not in layers:
            layers[bit] = np.zeros((H, W), dtype=np.uint32)
        assert layers[bit].dtype == np.uint32 and layers[bit].shape == (H, W) and layers[bit].flags.c_contiguous
    st = RenderStats()
    pv = None
    if prev is not None:
        assert prev.shape == cells.shape and prev.dtype == np.uint64 and prev.flags.c_contiguous
        pv = prev.ctypes.data
    rc = lib().orc_render(ptr, pv, W, H, C.byref(cfg), C.byref(params), layers[1].ctypes.data,
                          layers[2].ctypes.data, layers[4].ctypes.data, C.byref(st))
    assert rc == 0
    return layers, st


def body_pixel_xy(body: Body, tx, ty):
    x, y = C.c_int(), C.c_int()
    lib().orc_body_pixel_xy(C.byref(body), tx, ty, C.byref(x), C.byref(y))
    return x.value, y.value


def stamp(cells, cfg, bodies, pixels, displaced_cap=0):
    p, W, H = _grid(cells)
    nb = len(bodies)
    barr = (Body * max(nb, 1))(*bodies)
    px = np.ascontiguousarray(pixels, dtype=np.uint64)
    placed = np.zeros(max(px.size, 1), dtype=np.uint8)
    npl, ndp = C.c_uint32(), C.c_uint32()
    disp = (CellEdit * max(displaced_cap, 1))()
    hs = (C.c_uint32 * max(nb, 1))()
    hw = (C.c_uint32 * max(nb, 1))()
    lib().orc_stamp(p, W, H, C.byref(cfg), barr, nb, px.ctypes.data, px.size, placed.ctypes.data,
                    C.byref(npl), C.byref(ndp), disp, displaced_cap, hs, hw)
    n = min(ndp.value, displaced_cap)
    displaced = [(disp[i].x, disp[i].y,
                  (disp[i].cell.mat & 0xFFFF) | ((disp[i].cell.temp & 0xFFFF) << 16) |
                  ((disp[i].cell.rgba & 0xFFFFFFFF) << 32)) for i in range(n)]
    return {"placed": npl.value, "displaced": ndp.value, "displaced_cells": displaced,
            "hits_sand": list(hs)[:nb], "hits_soup": list(hw)[:nb], "placed_mask": placed[:px.size]}


def unstamp(cells, cfg, bodies, pixels):
    p, W, H = _grid(cells)
    nb = len(bodies)
    barr = (Body * max(nb, 1))(*bodies)
    px = np.array(pixels, dtype=np.uint64, copy=True)
    lib().orc_unstamp(p, W, H, C.byref(cfg), barr, nb, px.ctypes.data, px.size)
    return px


def refstamp(cells, cfg, bodies, pixels, displaced_cap=0):
    """The reference's own pixel-major stamp loop (game.cpp:1768-1812), in
    place on cells; displaced cells in ejection order."""
    p, W, H = _grid(cells)
    nb = len(bodies)
    barr = (Body * max(nb, 1))(*bodies)
    px = np.ascontiguousarray(pixels, dtype=np.uint64)
    placed = np.zeros(max(px.size, 1), dtype=np.uint8)
    npl, ndp = C.c_uint32(), C.c_uint32()
    disp = (CellEdit * max(displaced_cap, 1))()
    lib().orc_refstamp(p, W, H, C.byref(cfg), barr, nb, px.ctypes.data, px.size, placed.ctypes.data,
                       C.byref(npl), C.byref(ndp), disp, displaced_cap)
    n = min(ndp.value, displaced_cap)
    displaced = [(disp[i].x, disp[i].y,
                  (disp[i].cell.mat & 0xFFFF) | ((disp[i].cell.temp & 0xFFFF) << 16) |
                  ((disp[i].cell.rgba & 0xFFFFFFFF) << 32)) for i in range(n)]
    return {"placed": npl.value, "displaced": ndp.value, "displaced_cells": displaced,
            "placed_mask": placed[:px.size]}


def refunstamp(cells, cfg, bodies, pixels):
    """The reference's own pixel-major unstamp loop (game.cpp:1908-1965)."""
    p, W, H = _grid(cells)
    nb = len(bodies)
    barr = (Body * max(nb, 1))(*bodies)
    px = np.array(pixels, dtype=np.uint64, copy=True)
    lib().orc_refunstamp(p, W, H, C.byref(cfg), barr, nb, px.ctypes.data, px.size)
    return px


def physics_check(cells, cfg, seeds):
    """world::physicsCheck for seeds [(x, y), ...] in order, in place on cells.
    Returns per-seed dicts like World.physicsCheck (pixels for bodies)."""
    ptr, W, H = _grid(cells)
    n = len(seeds)
    arr = (Island * max(n, 1))()
    for i, (x, y) in enumerate(seeds):
        arr[i].x, arr[i].y = int(x), int(y)
    need = C.c_size_t()
    rc = lib().orc_physics_check(ptr, W, H, C.byref(cfg), arr, n, None, 0, C.byref(need))
    pix = np.empty(max(need.value, 1), dtype=np.uint64)
    if rc != 0:
        rc = lib().orc_physics_check(ptr, W, H, C.byref(cfg), arr, n, pix.ctypes.data, pix.size, C.byref(need))
    assert rc == 0, rc
    out = []
    for i in range(n):
        r = arr[i]
        d = {"status": r.status, "count": r.count, "box": (r.min_x, r.min_y, r.w, r.h)}
        if r.status == 3:
            d["pixels"] = pix[r.pixel_offset:r.pixel_offset + r.w * r.h].reshape(r.h, r.w).copy()
        out.append(d)
    return out


# --- chunk .pack images (chunk.cpp:74-330) ---------------------------------
PACK_NA, PACK_NB, PACK_HDR = 393216, 65536, 17


def lz4_bound(n: int) -> int:
    return n + n // 255 + 16


def _ptr(a):
    return None if a is None else a.ctypes.data


def chunk_records(tiles, layer2=None, bg=None, pack_map=None):
    """(raw_a, raw_b) records of one chunk; tiles/layer2 (uint64) and bg
    (uint32) are 128x128 windows of any row stride; pack_map a PackMap or
    None (index = material id)."""
    ra = np.zeros(PACK_NA, np.uint8)
    rb = np.zeros(PACK_NB, np.uint8)

    def win(a, dt):
        if a is None:
            return None, 0
        assert a.dtype == dt and a.shape[0] >= 128 and a.shape[1] >= 128 and a.strides[1] == a.itemsize
        return a, a.strides[0] // a.itemsize

    t, ts = win(tiles, np.uint64)
    l2, ls = win(layer2, np.uint64)
    b, bs = win(bg, np.uint32)
    lib().orc_chunk_records(_ptr(t), ts, _ptr(l2), ls, _ptr(b), bs, ra.ctypes.data, rb.ctypes.data,
                            None if pack_map is None else C.byref(pack_map))
    return ra, rb


def lz4_decode(src: bytes, cap: int):
    """Decoded bytes, or None when the block is rejected or not cap long."""
    s = np.frombuffer(bytes(src), np.uint8)
    out = np.zeros(max(cap, 1), np.uint8)
    r = lib().orc_lz4_decode(s.ctypes.data if len(s) else None, len(s), out.ctypes.data, cap)
    return None if r != cap else out[:cap].tobytes()


def lz4_encode_rows(raw, rec: int) -> bytes:
    raw = np.ascontiguousarray(raw, np.uint8)
    out = np.zeros(lz4_bound(len(raw)), np.uint8)
    r = lib().orc_lz4_encode_rows(raw.ctypes.data, len(raw), rec, out.ctypes.data, len(out))
    assert r >= 0, r
    return out[:r].tobytes()


def pack_chunk(raw_a, raw_b, phase: int = 0) -> bytes:
    img = np.zeros(PACK_HDR + lz4_bound(PACK_NA) + lz4_bound(PACK_NB), np.uint8)
    n = lib().orc_pack_chunk(raw_a.ctypes.data, raw_b.ctypes.data, phase, img.ctypes.data)
    assert n > 0
    return img[:n].tobytes()


def unpack_chunk(img: bytes, n_materials: int, pack_map=None):
    """(status, tiles, layer2, bg, phase) of one image (arrays 128x128)."""
    s = np.frombuffer(bytes(img), np.uint8)
    tiles = np.zeros((128, 128), np.uint64)
    l2 = np.zeros((128, 128), np.uint64)
    bg = np.zeros((128, 128), np.uint32)
    ph = C.c_int8(0)
    rc = lib().orc_unpack_chunk(s.ctypes.data if len(s) else None, len(s), n_materials, tiles.ctypes.data, 128,
                                l2.ctypes.data, 128, bg.ctypes.data, 128, C.byref(ph),
                                None if pack_map is None else C.byref(pack_map))
    return rc, tiles, l2, bg, ph.value


def _mask(m):
    m = np.ascontiguousarray(m, dtype=np.uint8)
    assert m.ndim == 2
    return m, m.shape[1], m.shape[0]


def ms_value(mask, x, y) -> int:
    """MarchingSquares::value at vertex (x, y) of a 0/1 mask (H, W)."""
    m, W, H = _mask(mask)
    return lib().orc_ms_value(m.ctypes.data, W, H, x, y)


def find_perimeter(mask, x0, y0):
    """FindPerimeter's runs [(dx, dy), ...] from vertex (x0, y0); [] off a perimeter."""
    m, W, H = _mask(mask)
    cap = 4 * (W + 1) * (H + 1) + 8
    d = np.zeros(2 * cap, np.int32)
    n = lib().orc_find_perimeter(m.ctypes.data, W, H, x0, y0, d.ctypes.data, cap)
    assert n >= 0, n
    return [(int(d[2 * i]), int(d[2 * i + 1])) for i in range(n)]


def simplify(points, tol=1.0):
    """simplify (Ramer-Douglas-Peucker in f32) of [(x, y), ...]."""
    p = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 2))
    out = np.zeros_like(p)
    n = lib().orc_simplify(p.ctypes.data, len(p), tol, out.ctypes.data)
    return [(float(a), float(b)) for a, b in out[:n]]


def mesh_mask(mask, raw=False):
    """updateChunkMesh's outline of a 0/1 mask: [(start_x, start_y, hole, [(x, y), ...]), ...]
    (raw=True: the loop count and vertex count only, with no Python objects)."""
    m, W, H = _mask(mask)
    lcap, vcap = W * H + 8, 4 * (W + 1) * (H + 1) + 8
    loops = (MeshLoop * lcap)()
    verts = np.zeros(2 * vcap, np.float32)
    nl, nv = C.c_uint32(), C.c_uint32()
    rc = lib().orc_mesh_mask(m.ctypes.data, W, H, loops, lcap, verts.ctypes.data, vcap, C.byref(nl), C.byref(nv))
    assert rc == 0, rc
    if raw:
        return nl.value, nv.value
    out = []
    for i in range(nl.value):
        L = loops[i]
        v = verts[2 * L.first:2 * (L.first + L.count)].reshape(-1, 2)
        out.append((L.start_x, L.start_y, bool(L.hole), [(float(a), float(b)) for a, b in v]))
    return out


def spiral():
    """tickCells' landing spiral offsets [(x, y), ...] in search order."""
    o = np.zeros(2 * 1024, np.int32)
    n = lib().orc_spiral(o.ctypes.data)
    return [(int(o[2 * i]), int(o[2 * i + 1])) for i in range(n)]


def particles_tick(cells, cfg, particles, zone=None):
    """world::tickCells over a list of Particle structs, in place on cells;
    returns (surviving particles, cells landed).  zone = (x, y, w, h) is the
    tick zone."""
    ptr, W, H = _grid(cells)
    arr = (Particle * max(len(particles), 1))(*particles)
    alive, landed = C.c_size_t(), C.c_size_t()
    z, zp = _zone(zone)
    rc = lib().orc_particles_tick_zone(ptr, W, H, C.byref(cfg), arr, len(particles), C.byref(alive),
                                       C.byref(landed), zp)
    assert rc == 0
    return [arr[i] for i in range(alive.value)], landed.value
