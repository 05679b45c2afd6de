"""C-ABI library: loads without a GPU, exports every declared symbol, host
entry points behave; generator invariants."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import falling_sand_engine_amd as fsg
from falling_sand_engine_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(fsg_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = fsg.lib()
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_signature_table_covers_header():
    missing = declared_symbols() - set(_abi.SIGNATURES)
    assert not missing, missing


def test_struct_sizes():
    assert C.sizeof(_abi.Cell) == 8
    assert C.sizeof(_abi.Material) == 28
    assert C.sizeof(_abi.Config) == 456
    assert C.sizeof(_abi.Body) == 32
    assert C.sizeof(_abi.Stats) == 168
    assert C.sizeof(_abi.CellEdit) == 16
    assert C.sizeof(_abi.Tuning) == 16
    assert C.sizeof(_abi.TuneReport) == 800
    assert C.sizeof(_abi.RenderParams) == 92
    assert C.sizeof(_abi.RenderStats) == 144


def test_default_config_matches_oracle_table(orc):
    assert bytes(fsg.default_config()) == bytes(orc.default_config())


def test_status_strings():
    lib = fsg.lib()
    assert lib.fsg_status_string(0) == b"ok"
    assert lib.fsg_status_string(-4) == b"HBM state budget exceeded"
    assert lib.fsg_status_string(12345) == b"unknown status"


def test_invalid_arguments_are_rejected_without_device():
    lib = fsg.lib()
    h = C.c_void_p()
    assert lib.fsg_world_create(None, 64, 64, 0, None, C.byref(h)) == _abi.FSG_ERR_INVALID
    assert lib.fsg_world_step(None, 1) == _abi.FSG_ERR_INVALID
    assert lib.fsg_generate(9, 16, 16, 0, None) == _abi.FSG_ERR_INVALID
    out = np.empty((4, 4), dtype=np.uint64)
    assert lib.fsg_generate(9, 4, 4, 0, out.ctypes.data) == _abi.FSG_ERR_INVALID
    assert lib.fsg_generate(1, 0, 4, 0, out.ctypes.data) == _abi.FSG_ERR_INVALID
    t = _abi.Tuning(16, -1, 5, 0)
    assert lib.fsg_ctx_get_tuning(None, C.byref(t)) == _abi.FSG_ERR_INVALID
    assert lib.fsg_ctx_set_tuning(None, C.byref(t)) == _abi.FSG_ERR_INVALID
    assert lib.fsg_autotune_batch(None, None, 0, 8, None) == _abi.FSG_ERR_INVALID
    assert lib.fsg_ctx_timer_start(None) == _abi.FSG_ERR_INVALID
    assert lib.fsg_ctx_timer_stop(None, None) == _abi.FSG_ERR_INVALID
    assert lib.fsg_world_physics_check(None, None, 0, None) == _abi.FSG_ERR_INVALID
    assert lib.fsg_world_island_pixels(None, None, 0) == _abi.FSG_ERR_INVALID
    assert C.sizeof(_abi.Island) == 40
    rp = _abi.RenderParams()
    assert lib.fsg_world_render(None, C.byref(rp), None) == _abi.FSG_ERR_INVALID
    ptr = C.c_void_p()
    assert lib.fsg_world_layer(None, _abi.LAYER_CELLS, C.byref(ptr)) == _abi.FSG_ERR_INVALID
    tex = np.empty((4, 4), dtype=np.uint32)
    assert lib.fsg_world_layer_download(None, _abi.LAYER_CELLS, 0, 0, 4, 4, tex.ctypes.data) == _abi.FSG_ERR_INVALID


def test_ctx_create_without_gpu_reports_device_error():
    # this container has no GPU: the library must say so, not crash
    import subprocess
    import sys
    code = ("import ctypes as C, falling_sand_engine_amd as f; from falling_sand_engine_amd import _abi;"
            "h=C.c_void_p(); cfg=_abi.CtxConfig(0,0);"
            "print(f.lib().fsg_ctx_create(C.byref(cfg), C.byref(h)))")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.stdout.strip() == str(_abi.FSG_ERR_DEVICE), r.stdout + r.stderr


# --- generator ---------------------------------------------------------------------

def test_generator_basin():
    W, H = 512, 512
    g = fsg.generate(1, W, H, 1)
    m = fsg.cell_mat(g)
    t = fsg.cell_temp(g)
    assert set(np.unique(m)) <= {0, 1, 2, 3}
    assert (t == 320).all()
    for x in range(0, W, 7):
        col = m[:, x]
        stone = np.nonzero(col == 1)[0]
        if stone.size:  # stone is one run down to the bottom row
            assert stone.max() == H - 1 and stone.size == stone.max() - stone.min() + 1
            assert stone.size <= 0.4 * H + 1
    water_rows = np.nonzero((m == 3).any(axis=1))[0]
    assert water_rows.min() >= H - (H * 55) // 100
    assert (m == 2).sum() > 0
    assert (fsg.cell_rgba(g)[m == 0] == 0).all()
    assert np.array_equal(g, fsg.generate(1, W, H, 1))  # deterministic
    assert not np.array_equal(g, fsg.generate(1, W, H, 2))


def test_generator_lava_and_oil():
    g = fsg.generate(2, 1024, 1024, 2)
    m = fsg.cell_mat(g)
    t = fsg.cell_temp(g)
    assert (m == 6).sum() > 0 and (m == 4).sum() > 0
    assert (t[m == 6] == 1200 * 16).all()
    assert (t[m != 6] == 320).all()


def test_generator_striped():
    W, H = 256, 200
    g = fsg.generate(3, W, H, 3)
    m = fsg.cell_mat(g)
    assert (m[H - 1] == 1).all()
    assert (m[H // 10 + 1] == 3).all() and (m[8 * H // 10 + 1] == 5).all()
    sand_rows = [y for y in range(H - 1) if not (H <= 10 * y < 2 * H or 8 * H <= 10 * y < 9 * H)]
    frac = np.mean([np.mean(m[y] == 2) for y in sand_rows if y % 2 == 0])
    assert 0.45 < frac < 0.55
    assert all((m[y] == 0).all() for y in sand_rows if y % 2 == 1)
    assert abs((m == 3).sum() / (W * H) - 0.1) < 0.01 and abs((m == 5).sum() / (W * H) - 0.1) < 0.01
