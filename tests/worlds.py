"""Small hand-built worlds for rule tests (ASCII art -> uint64 grids)."""
import numpy as np

CHARS = {".": 0, "#": 1, "s": 2, "w": 3, "o": 4, "~": 5, "l": 6, "W": 7}
AIR, STONE, SAND, WATER, OIL, STEAM, LAVA, WOOD = range(8)


def cell(mat, temp=320, rgba=0, moved=False):
    return (mat | (0x100 if moved else 0)) | ((temp & 0xFFFF) << 16) | ((rgba & 0xFFFFFFFF) << 32)


def parse(art, temp=320):
    rows = [r for r in art.strip("\n").split("\n")]
    H, W = len(rows), len(rows[0])
    g = np.zeros((H, W), dtype=np.uint64)
    for y, r in enumerate(rows):
        assert len(r) == W
        for x, ch in enumerate(r):
            g[y, x] = cell(CHARS[ch], temp)
    return g


def mats(g):
    return (g & np.uint64(0xFF)).astype(np.int64)


def temps(g):
    return ((g >> np.uint64(16)) & np.uint64(0xFFFF)).astype(np.uint16).view(np.int16).astype(np.int64)


def render(g):
    inv = {v: k for k, v in CHARS.items()}
    return "\n".join("".join(inv[int(v) & 0xFF] for v in row) for row in g)
