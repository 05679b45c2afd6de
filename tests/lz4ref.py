"""ctypes binding of the reference's own LZ4 (oracle/_ref/liblz4ref.so, built
from source/libs/lz4/lz4.c by `make -C oracle ref`).  Test infrastructure:
the decoder every .pack block must satisfy and the compressor chunk.cpp
calls.  lib() is None when the library has not been built."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "liblz4ref.so")
_L = False


def lib():
    global _L
    if _L is False:
        _L = None
        if os.path.exists(PATH):
            L = C.CDLL(PATH)
            P, i = C.c_void_p, C.c_int
            L.LZ4_compress_fast.restype = i
            L.LZ4_compress_fast.argtypes = [P, P, i, i, i]
            L.LZ4_decompress_safe.restype = i
            L.LZ4_decompress_safe.argtypes = [P, P, i, i]
            L.LZ4_compressBound.restype = i
            L.LZ4_compressBound.argtypes = [i]
            L.LZ4_versionNumber.restype = i
            L.LZ4_versionNumber.argtypes = []
            _L = L
    return _L


def compress_fast(raw, accel: int = 10) -> bytes:
    """LZ4_compress_fast(src, dst, n, bound, accel), as chunk.cpp:251 calls it."""
    raw = np.ascontiguousarray(np.frombuffer(bytes(raw), np.uint8))
    cap = lib().LZ4_compressBound(len(raw))
    out = np.zeros(cap, np.uint8)
    r = lib().LZ4_compress_fast(raw.ctypes.data, out.ctypes.data, len(raw), cap, accel)
    assert r > 0, r
    return out[:r].tobytes()


def decompress_safe(src: bytes, cap: int):
    """LZ4_decompress_safe(src, dst, n, cap) -> (return value, output bytes)."""
    s = np.frombuffer(bytes(src), np.uint8)
    out = np.zeros(max(cap, 1), np.uint8)
    r = lib().LZ4_decompress_safe(s.ctypes.data if len(s) else None, out.ctypes.data, len(s), cap)
    return r, (out[:r].tobytes() if r >= 0 else None)
