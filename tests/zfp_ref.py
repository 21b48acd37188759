"""ctypes access to ZFP 0.5.5 compiled from the reference's vendored source
(oracle/_ref/zfp/libzfp.so, built by oracle/build.py) -- TEST
INFRASTRUCTURE ONLY: the checker this library's own ZFP stream codec is
compared with.  Calls follow SRC/compression_tools.c:33-70 / :94-132
(zfp_field_2d over the flat array, headerless stream)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "zfp", "libzfp.so")
ZFP_TYPE_FLOAT = 3          # zfp.h: zfp_type_float
ZFP_EXEC_OMP = 1            # zfp.h: zfp_exec_omp

_lib = None


def available() -> bool:
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB)
        vp, sz, u = C.c_void_p, C.c_size_t, C.c_uint
        for name, res, args in (
                ("zfp_field_2d", vp, [vp, C.c_int, u, u]),
                ("zfp_field_free", None, [vp]),
                ("zfp_stream_open", vp, [vp]),
                ("zfp_stream_close", None, [vp]),
                ("zfp_stream_set_accuracy", C.c_double, [vp, C.c_double]),
                ("zfp_stream_set_precision", u, [vp, u]),
                ("zfp_stream_set_rate", C.c_double, [vp, C.c_double, C.c_int, u, C.c_int]),
                ("zfp_stream_maximum_size", sz, [vp, vp]),
                ("zfp_stream_set_bit_stream", None, [vp, vp]),
                ("zfp_stream_rewind", None, [vp]),
                ("zfp_stream_set_execution", C.c_int, [vp, C.c_int]),
                ("zfp_compress", sz, [vp, vp]),
                ("zfp_decompress", sz, [vp, vp]),
                ("stream_open", vp, [vp, sz]),
                ("stream_close", None, [vp])):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _set_mode(L, zs, mode, value):
    if mode == "accuracy":
        L.zfp_stream_set_accuracy(zs, float(value))
    elif mode == "precision":
        L.zfp_stream_set_precision(zs, int(value))
    elif mode == "rate":
        L.zfp_stream_set_rate(zs, float(value), ZFP_TYPE_FLOAT, 2, 0)
    else:
        raise ValueError(mode)


def compress(a: np.ndarray, mode: str, value, omp: bool = False) -> bytes:
    """a[ny][nx] float32 (x fastest) -> stream bytes."""
    L = lib()
    a = np.ascontiguousarray(a, np.float32)
    ny, nx = a.shape
    field = L.zfp_field_2d(a.ctypes.data, ZFP_TYPE_FLOAT, nx, ny)
    zs = L.zfp_stream_open(None)
    _set_mode(L, zs, mode, value)
    cap = L.zfp_stream_maximum_size(zs, field)
    buf = C.create_string_buffer(cap)
    bs = L.stream_open(C.cast(buf, C.c_void_p), cap)
    L.zfp_stream_set_bit_stream(zs, bs)
    L.zfp_stream_rewind(zs)
    if omp:
        L.zfp_stream_set_execution(zs, ZFP_EXEC_OMP)
    n = L.zfp_compress(zs, field)
    L.stream_close(bs)
    L.zfp_stream_close(zs)
    L.zfp_field_free(field)
    return buf.raw[:n]


def max_size(nx: int, ny: int, mode: str, value) -> int:
    L = lib()
    field = L.zfp_field_2d(None, ZFP_TYPE_FLOAT, nx, ny)
    zs = L.zfp_stream_open(None)
    _set_mode(L, zs, mode, value)
    n = L.zfp_stream_maximum_size(zs, field)
    L.zfp_stream_close(zs)
    L.zfp_field_free(field)
    return int(n)


def decompress(stream: bytes, nx: int, ny: int, mode: str, value) -> np.ndarray:
    L = lib()
    out = np.zeros((ny, nx), np.float32)
    field = L.zfp_field_2d(out.ctypes.data, ZFP_TYPE_FLOAT, nx, ny)
    zs = L.zfp_stream_open(None)
    _set_mode(L, zs, mode, value)
    cap = max(L.zfp_stream_maximum_size(zs, field), len(stream))
    buf = C.create_string_buffer(bytes(stream) + b"\0" * (cap - len(stream)), cap)
    bs = L.stream_open(C.cast(buf, C.c_void_p), cap)
    L.zfp_stream_set_bit_stream(zs, bs)
    L.zfp_stream_rewind(zs)
    L.zfp_decompress(zs, field)
    L.stream_close(bs)
    L.zfp_stream_close(zs)
    L.zfp_field_free(field)
    return out
