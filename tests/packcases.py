"""Chunk contents shared by the .pack tests and their golden-image script:
deterministic layer2 and background patterns (the benchmark worlds have
neither)."""
import numpy as np

STONE = 1


def layer2_pattern(h: int, w: int, seed: int) -> np.ndarray:
    """Stone cells on a sparse diagonal lattice, air elsewhere (uint64 cells)."""
    y, x = np.mgrid[0:h, 0:w]
    on = ((x * 7 + y * 3 + seed) % 11) == 0
    rgba = (0xFF707070 + ((x + seed) % 5)).astype(np.uint64)
    cell = np.uint64(STONE) | (np.uint64(320) << np.uint64(16)) | (rgba << np.uint64(32))
    return np.where(on, cell, np.uint64(0)).astype(np.uint64)


def background_pattern(h: int, w: int, seed: int) -> np.ndarray:
    """Opaque colour blocks of 8x8 cells with a seed-dependent blue (uint32)."""
    y, x = np.mgrid[0:h, 0:w]
    return (0xFF000000 | ((y // 8) % 256) << 16 | ((x // 8) % 256) << 8 | (seed % 256)).astype(np.uint32)


def chunk_of(cells: np.ndarray, cx: int, cy: int) -> np.ndarray:
    return np.ascontiguousarray(cells[cy * 128:(cy + 1) * 128, cx * 128:(cx + 1) * 128])
