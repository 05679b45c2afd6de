"""ctypes mirror of include/fsg.h (structs, status codes, function signatures)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfsg.so")

FSG_OK = 0
STATUS = {
    0: "FSG_OK", -1: "FSG_ERR_INVALID", -2: "FSG_ERR_NOMEM", -3: "FSG_ERR_DEVICE",
    -4: "FSG_ERR_BUDGET", -5: "FSG_ERR_STATE", -6: "FSG_ERR_CHECKSUM", -7: "FSG_ERR_TIMEOUT",
    -8: "FSG_ERR_CANCELLED", -9: "FSG_ERR_NOTFOUND", -10: "FSG_ERR_AGAIN",
}
FSG_ERR_INVALID, FSG_ERR_NOMEM, FSG_ERR_DEVICE, FSG_ERR_BUDGET, FSG_ERR_STATE = -1, -2, -3, -4, -5
FSG_ERR_CHECKSUM, FSG_ERR_TIMEOUT, FSG_ERR_CANCELLED, FSG_ERR_NOTFOUND, FSG_ERR_AGAIN = -6, -7, -8, -9, -10

AIR, SOLID, SAND, SOUP, GAS = 0, 1, 2, 3, 4
MOVED = 0x100


class Material(C.Structure):
    _fields_ = [("name", C.c_char * 16), ("type", C.c_uint8), ("rank", C.c_uint8),
                ("iterations", C.c_uint8), ("slipperyness", C.c_uint8), ("add_temp", C.c_int16),
                ("reserved", C.c_uint16), ("rgba", C.c_uint32)]


class Config(C.Structure):
    _fields_ = [("cell_iter", C.c_uint32), ("reactions", C.c_uint32), ("temp_every", C.c_uint32),
                ("temp_phase", C.c_uint32), ("chunk_skip", C.c_uint32), ("wood_burn_temp", C.c_int32),
                ("mat_air", C.c_uint8), ("mat_water", C.c_uint8), ("mat_lava", C.c_uint8),
                ("mat_steam", C.c_uint8), ("mat_stone", C.c_uint8), ("mat_wood", C.c_uint8),
                ("pad", C.c_uint8 * 2), ("n_materials", C.c_uint32), ("materials", Material * 15)]


class CtxConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_uint64)]


class Cell(C.Structure):
    _fields_ = [("mat", C.c_uint16), ("temp", C.c_int16), ("rgba", C.c_uint32)]


class CellEdit(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("cell", Cell)]


class Stats(C.Structure):
    _fields_ = [("tick", C.c_uint64), ("count", C.c_uint64 * 16), ("checksum", C.c_uint64),
                ("tiles_total", C.c_uint64), ("tiles_processed", C.c_uint64),
                ("tiles_offered", C.c_uint64)]


class Body(C.Structure):
    _fields_ = [("pos_x_q16", C.c_int32), ("pos_y_q16", C.c_int32), ("cos_q16", C.c_int32),
                ("sin_q16", C.c_int32), ("mat_w", C.c_int32), ("mat_h", C.c_int32),
                ("pixel_offset", C.c_uint32), ("enabled", C.c_uint32)]


class StampResult(C.Structure):
    _fields_ = [("placed", C.c_uint32), ("unplaced", C.c_uint32), ("displaced", C.c_uint32),
                ("displaced_cap", C.c_uint32), ("displaced_out", C.POINTER(CellEdit)),
                ("hits_sand", C.POINTER(C.c_uint32)), ("hits_soup", C.POINTER(C.c_uint32))]


class ServeConfig(C.Structure):
    _fields_ = [("max_active", C.c_int32), ("block_ticks", C.c_int32), ("verify_resident", C.c_int32),
                ("gen_threads", C.c_int32)]


class ServeRequest(C.Structure):
    _fields_ = [("world_id", C.c_uint64), ("kind", C.c_int32), ("width", C.c_int32),
                ("height", C.c_int32), ("ticks", C.c_int32), ("seed", C.c_uint64),
                ("n_bodies", C.c_int32), ("reserved", C.c_int32), ("deadline_ns", C.c_uint64)]


class ServeResult(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("world_id", C.c_uint64), ("status", C.c_int32),
                ("ticks_done", C.c_int32), ("tick", C.c_uint64), ("checksum", C.c_uint64),
                ("cells", C.c_uint64), ("t_submit_ns", C.c_uint64), ("t_admit_ns", C.c_uint64),
                ("t_done_ns", C.c_uint64)]


class ServeStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "completed", "failed", "queued", "active",
                                          "resident_worlds", "resident_bytes", "steps", "cell_updates",
                                          "evictions")]


class Tuning(C.Structure):
    _fields_ = [("graph_ticks", C.c_int32), ("chunk_skip", C.c_int32), ("occupancy", C.c_int32),
                ("reserved", C.c_int32)]


TUNE_MAX_CANDIDATES = 32


class TuneReport(C.Structure):
    _fields_ = [("best", Tuning), ("n_candidates", C.c_int32), ("from_cache", C.c_int32),
                ("size_class", C.c_uint64), ("candidate", Tuning * TUNE_MAX_CANDIDATES),
                ("ms", C.c_double * TUNE_MAX_CANDIDATES)]


LAYER_CELLS, LAYER_EMISSION, LAYER_FIRE = 1, 2, 4
ISLAND_NONE, ISLAND_ATTACHED, ISLAND_REMOVED, ISLAND_BODY = 0, 1, 2, 3
ISLAND_MAX_CELLS, ISLAND_MIN_BODY = 1000, 11


class Island(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("status", C.c_int32), ("count", C.c_uint32),
                ("min_x", C.c_int32), ("min_y", C.c_int32), ("w", C.c_int32), ("h", C.c_int32),
                ("pixel_offset", C.c_uint64)]


class RenderParams(C.Structure):
    _fields_ = [("layers", C.c_uint32), ("x0", C.c_int32), ("y0", C.c_int32), ("w", C.c_int32),
                ("h", C.c_int32), ("fire_mat", C.c_uint32), ("count_moved", C.c_uint32),
                ("emit", C.c_uint32 * 16)]


class RenderStats(C.Structure):
    _fields_ = [("moved", C.c_uint64 * 16), ("fire_cells", C.c_uint64), ("had_dirty", C.c_uint32),
                ("reserved", C.c_uint32)]


class MeshRegion(C.Structure):
    _fields_ = [("x0", C.c_int32), ("y0", C.c_int32), ("w", C.c_int32), ("h", C.c_int32), ("status", C.c_int32),
                ("first_loop", C.c_uint32), ("nloops", C.c_uint32)]


class MeshLoop(C.Structure):
    _fields_ = [("first", C.c_uint32), ("count", C.c_uint32), ("hole", C.c_int32), ("start_x", C.c_int32),
                ("start_y", C.c_int32), ("region", C.c_uint32)]


class Particle(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("x", "y", "vx", "vy", "ax", "ay", "target_x", "target_y",
                                         "target_force")] + \
               [("lifetime", C.c_int32), ("flags", C.c_uint32), ("in_object_state", C.c_uint32), ("tile", Cell)]


class PackMap(C.Structure):
    _fields_ = [("to_file", C.c_uint32 * 16), ("from_file", C.c_int16 * 256), ("flags", C.c_uint32)]


P = C.c_void_p
u64p = C.POINTER(C.c_uint64)

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "fsg_status_string": (C.c_char_p, [C.c_int]),
    "fsg_default_config": (C.c_int, [C.POINTER(Config)]),
    "fsg_ctx_create": (C.c_int, [C.POINTER(CtxConfig), C.POINTER(P)]),
    "fsg_ctx_destroy": (C.c_int, [P]),
    "fsg_ctx_usage": (C.c_int, [P, u64p]),
    "fsg_ctx_inject_alloc_failures": (C.c_int, [P, C.c_int]),
    "fsg_ctx_copy_timed": (C.c_int, [P, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "fsg_ctx_timer_start": (C.c_int, [P]),
    "fsg_ctx_timer_stop": (C.c_int, [P, C.POINTER(C.c_double)]),
    "fsg_world_create": (C.c_int, [P, C.c_int, C.c_int, C.c_uint64, C.POINTER(Config), C.POINTER(P)]),
    "fsg_world_destroy": (C.c_int, [P]),
    "fsg_world_info": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), u64p]),
    "fsg_world_set_tick": (C.c_int, [P, C.c_uint64]),
    "fsg_world_upload": (C.c_int, [P, P, C.c_size_t]),
    "fsg_world_download": (C.c_int, [P, P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "fsg_world_generate": (C.c_int, [P, C.c_int, C.c_uint64]),
    "fsg_world_pack": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P, P, P, P]),
    "fsg_world_pack_download": (C.c_int, [P, P, C.c_size_t]),
    "fsg_world_unpack": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P, P, P, P, P, P]),
    "fsg_world_set_pack_map": (C.c_int, [P, C.POINTER(PackMap)]),
    "fsg_pack_map_default": (C.c_int, [C.POINTER(Config), C.POINTER(PackMap)]),
    "fsg_lz4_decode_blocks": (C.c_int, [P, P, C.c_size_t, P, P, P, C.c_int32, P, C.c_size_t, P, P]),
    "fsg_world_set_cells": (C.c_int, [P, P, C.c_size_t]),
    "fsg_world_set_cells_noalloc": (C.c_int, [P, P, C.c_size_t]),
    "fsg_world_step": (C.c_int, [P, C.c_int]),
    "fsg_world_step_timed": (C.c_int, [P, C.c_int, C.POINTER(C.c_double)]),
    "fsg_world_tick_temperature": (C.c_int, [P]),
    "fsg_world_set_tick_zone": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "fsg_world_get_tick_zone": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32)]),
    "fsg_world_sync": (C.c_int, [P]),
    "fsg_world_stats": (C.c_int, [P, C.POINTER(Stats)]),
    "fsg_world_reset_activity": (C.c_int, [P]),
    "fsg_world_stamp_bodies": (C.c_int, [P, P, C.c_int, P, C.c_size_t, C.POINTER(StampResult)]),
    "fsg_world_unstamp_bodies": (C.c_int, [P, P, C.c_size_t]),
    "fsg_world_physics_check": (C.c_int, [P, P, C.c_int, C.POINTER(C.c_size_t)]),
    "fsg_world_island_pixels": (C.c_int, [P, P, C.c_size_t]),
    "fsg_world_mesh": (C.c_int, [P, P, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "fsg_world_mesh_islands": (C.c_int, [P, P, C.c_int, P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "fsg_world_mesh_download": (C.c_int, [P, P, C.c_uint32, P, C.c_uint32]),
    "fsg_world_particles_add": (C.c_int, [P, P, C.c_size_t]),
    "fsg_world_particles_tick": (C.c_int, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "fsg_world_particles_get": (C.c_int, [P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fsg_world_render": (C.c_int, [P, C.POINTER(RenderParams), C.POINTER(RenderStats)]),
    "fsg_world_layer": (C.c_int, [P, C.c_uint32, C.POINTER(P)]),
    "fsg_world_layer_download": (C.c_int, [P, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int, P]),
    "fsg_batch_step": (C.c_int, [P, P, C.c_int, C.c_int]),
    "fsg_batch_step_timed": (C.c_int, [P, P, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "fsg_perworld_step_timed": (C.c_int, [P, P, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "fsg_ctx_get_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
    "fsg_ctx_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
    "fsg_autotune_batch": (C.c_int, [P, P, C.c_int, C.c_int, C.POINTER(TuneReport)]),
    "fsg_generate": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, P]),
    "fsg_rule_eval": (C.c_int, [C.POINTER(Config), C.c_int, C.c_uint32, C.c_uint32, P, C.c_int]),
    "fsg_now_ns": (C.c_uint64, []),
    "fsg_serve_create": (C.c_int, [P, C.POINTER(ServeConfig), C.POINTER(P)]),
    "fsg_serve_destroy": (C.c_int, [P]),
    "fsg_serve_submit": (C.c_int, [P, C.POINTER(ServeRequest), C.POINTER(C.c_uint64)]),
    "fsg_serve_cancel": (C.c_int, [P, C.c_uint64]),
    "fsg_serve_poll": (C.c_int, [P, C.POINTER(ServeResult), C.c_int, C.POINTER(C.c_int)]),
    "fsg_serve_drain": (C.c_int, [P, C.c_uint64]),
    "fsg_serve_release": (C.c_int, [P, C.c_uint64]),
    "fsg_serve_stats_get": (C.c_int, [P, C.POINTER(ServeStats)]),
    "fsg_serve_corrupt": (C.c_int, [P, C.c_uint64]),
    "fsg_splitmix64": (C.c_uint64, [C.c_uint64]),
    "fsg_phase_key": (C.c_uint64, [C.c_uint64, C.c_uint64]),
}

_LIB = None


def lib():
    """The loaded libfsg.so.  Raises if it has not been built: there is no
    fallback implementation of the tick."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m falling_sand_engine_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(L, name):
                continue
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class FsgError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__(f"{what}: {STATUS.get(status, status)}")


def check(status, what=""):
    if status != FSG_OK:
        raise FsgError(status, what)
    return status


CHUNK = 128  # FSG_CHUNK: chunk side of the .pack images (const.h:19-20)
PACK_NO_MOVED = 1  # FSG_PACK_NO_MOVED
MESH_MAX_CELLS = 16384  # FSG_MESH_MAX_CELLS
MESH_EDGE = 1  # FSG_MESH_EDGE
PARTICLE_PHASE, PARTICLE_TEMPORARY = 1, 2  # FSG_PARTICLE_*
