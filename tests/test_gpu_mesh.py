"""Device outlines (fsg_world_mesh, fsg_world_mesh_islands) against the
oracle's restatement of updateChunkMesh's marching squares: the same loops
(start cell, hole flag, every vertex) in the same order, region by region."""
import numpy as np
import pytest

import falling_sand_engine_amd as fsg
from falling_sand_engine_amd import _abi

pytestmark = pytest.mark.gpu


def fragments(W, H, seed, p):
    rng = np.random.default_rng(seed)
    mats = np.where(rng.random((H, W)) < p, fsg.MAT_STONE,
                    rng.choice([fsg.MAT_AIR, fsg.MAT_SAND, fsg.MAT_WATER], size=(H, W)))
    return mats.astype(np.uint64) | (np.uint64(320) << np.uint64(16)) | (np.uint64(0xFF808080) << np.uint64(32))


def solid_mask(cells, cfg):
    types = np.array([cfg.materials[i].type for i in range(cfg.n_materials)] + [0] * 16, np.uint8)
    return (types[(cells & np.uint64(15)).astype(np.int64)] == fsg.SOLID).astype(np.uint8)


def check_regions(orc, cells, cfg, regions, got):
    H, W = cells.shape
    solid = solid_mask(cells, cfg)
    n = 0
    for (x0, y0, w, h), (status, loops) in zip(regions, got):
        if x0 < 0 or y0 < 0 or x0 + w >= W or y0 + h >= H:
            assert status == _abi.MESH_EDGE and loops == []
            continue
        assert status == 0
        want = orc.mesh_mask(solid[y0:y0 + h, x0:x0 + w])
        assert loops == want, (x0, y0, w, h, len(loops), len(want))
        n += len(loops)
    return n


@pytest.mark.parametrize("kind,p,seed", [("fragments", 0.45, 1), ("fragments", 0.6, 2), ("basin", 0.0, 3)])
def test_chunk_meshes_match_oracle(ctx, orc, kind, p, seed):
    W, H = 640, 512
    cfg = orc.default_config()
    cells = fragments(W, H, seed, p) if kind == "fragments" else fsg.generate(1, W, H, seed)
    regions = [(cx * 128, cy * 128, 128, 128) for cy in range(H // 128) for cx in range(W // 128)]
    regions += [(13, 7, 37, 91), (200, 100, 128, 128), (5, 5, 1, 50), (300, 20, 64, 1), (0, 0, 3, 3),
                (-5, 10, 20, 20), (600, 10, 40, 20), (400, 300, 100, 160)]
    print(f"mesh {kind} p={p} {W}x{H}, {len(regions)} regions", flush=True)
    with fsg.World(ctx, W, H, seed, cfg) as w:
        w.upload(cells)
        got = w.mesh(regions)
        again = w.mesh(regions[::-1])
    n = check_regions(orc, cells, cfg, regions, got)
    assert n > (20 if kind == "fragments" else 3)  # basin terrain: a few smooth outlines
    assert again == got[::-1]  # region order in, region order out


def test_island_body_meshes_match_oracle(ctx, orc):
    W, H = 512, 512
    cfg = orc.default_config()
    cells = fragments(W, H, 9, 0.5)
    rng = np.random.default_rng(9)
    seeds = [(int(x), int(y)) for x, y in zip(rng.integers(0, W, 600), rng.integers(0, H, 600))]
    print("mesh island images", flush=True)
    with fsg.World(ctx, W, H, 9, cfg) as w:
        w.upload(cells)
        res = w.physicsCheck(seeds)
        got = w.mesh_islands()
    bodies = 0
    for r, (status, loops) in zip(res, got):
        mw, mh = r["box"][2], r["box"][3]
        if r["status"] != _abi.ISLAND_BODY or mw <= 1 or mh <= 1:
            assert status == _abi.MESH_EDGE and loops == []
            continue
        bodies += 1
        m = (fsg.cell_mat(r["pixels"]) != fsg.MAT_AIR).astype(np.uint8)
        assert status == 0 and loops == orc.mesh_mask(m)
    assert bodies > 20


def test_mesh_arguments(ctx, orc):
    with fsg.World(ctx, 256, 256, 1, orc.default_config()) as w:
        w.upload(fragments(256, 256, 3, 0.5))
        for bad in [(0, 0, 0, 10), (0, 0, 200, 200), (0, 0, -1, 5)]:
            with pytest.raises(fsg.FsgError):
                w.mesh([bad])
        assert w.mesh([]) == []
