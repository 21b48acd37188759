"""rtmuq2d_amd -- MI355X-native drop-in for rtmUq2D's reverse-time-migration path.

The product is ``lib/librtm_hip.so`` (C ABI declared in ``include/rtm_hip.h``):
hand-written HIP kernels for gfx950 plus a host C layer exporting the
reference's own entry points.  This module is a thin ctypes binding over that
ABI for Python callers (tests, the benchmark, the serving loop).  There is no
Python or CPU fallback for any computation: when the library is missing the
import of a function raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "librtm_hip.so")
PATH_LEN = 512

SNAP_FP32, SNAP_FP16, SNAP_BF16, SNAP_BOUNDARY, SNAP_I16 = 0, 1, 2, 3, 4
REAL_F32, REAL_F64 = 0, 1
INTERIOR_TILES, INTERIOR_COLUMNS, INTERIOR_MIXED, INTERIOR_TILES8, INTERIOR_PAIRS = 0, 1, 2, 3, 4


class ZfpMode(C.Structure):
    """Mirror of ``rtm_zfp_mode`` (include/rtm_hip.h)."""
    _fields_ = [("minbits", C.c_uint), ("maxbits", C.c_uint), ("maxprec", C.c_uint), ("minexp", C.c_int)]


class RtmParams(C.Structure):
    """Mirror of ``rtm_params`` (include/rtm_hip.h); field order is the ABI."""

    _fields_ = [
        ("velocity_file", C.c_char * PATH_LEN),
        ("wavelet_file", C.c_char * PATH_LEN),
        ("seismogram_file", C.c_char * PATH_LEN),
        ("stacked_file", C.c_char * PATH_LEN),
        ("stacked_laplacian_file", C.c_char * PATH_LEN),
        ("zfp_tolerance", C.c_float),
        ("spacing_grid", C.c_float),
        ("sampling", C.c_float),
        ("number_of_simulations", C.c_int),
        ("nx", C.c_int), ("nz", C.c_int),
        ("first_location_shot", C.c_int), ("depth_shot", C.c_int),
        ("spacing_shot", C.c_int), ("number_of_shots", C.c_int),
        ("first_location_receiver", C.c_int), ("depth_receiver", C.c_int),
        ("spacing_receiver", C.c_int), ("number_of_receivers", C.c_int),
        ("number_of_time_steps", C.c_int),
        ("sampling_wavelet", C.c_int),
        ("cut_frequency", C.c_float),
        ("thickness_up", C.c_int), ("thickness_down", C.c_int),
        ("thickness_left", C.c_int), ("thickness_right", C.c_int),
        ("time_slices", C.c_int),
        ("snapshot_mode", C.c_int),
        ("image_kahan", C.c_int),
        ("device", C.c_int),
        ("normalize_illumination", C.c_int),
    ]

    # geometry helpers
    @property
    def X(self) -> int:
        return self.thickness_left + self.nx + self.thickness_right

    @property
    def Z(self) -> int:
        return self.thickness_up + self.nz + self.thickness_down

    def copy(self) -> "RtmParams":
        q = RtmParams()
        C.pointer(q)[0] = self
        return q

    def update(self, **kw) -> "RtmParams":
        for k, v in kw.items():
            if isinstance(getattr(self, k), bytes):
                v = v.encode() if isinstance(v, str) else v
            setattr(self, k, v)
        return self


_lib = None


def lib() -> C.CDLL:
    """Load librtm_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    P = C.POINTER
    f32p, f64p, i32p = P(C.c_float), P(C.c_double), P(C.c_int)
    plan_p = C.c_void_p
    sig = {
        "rtm_params_default": (None, [P(RtmParams)]),
        "rtm_set_params": (C.c_int, [P(RtmParams)]),
        "rtm_get_params": (C.c_int, [P(RtmParams)]),
        "rtm_last_error": (C.c_char_p, []),
        "rtm_device_count": (C.c_int, []),
        "rtm_routine": (None, [C.c_int]),
        "isotropicAcousticModeling": (None, [C.c_int, i32p, f32p, f32p]),
        "adjointModeling": (C.c_int, [C.c_int, i32p, f32p, f32p, f32p, f32p, f32p]),
        "iso_acoustic_wave_equation": (None, [f32p, f32p, f32p]),
        "attenuation_profile_cpml": (None, [f32p] * 8),
        "aiso_attenuation_cpml": (None, [f32p] * 19),
        "cross_correlation": (None, [f32p] * 3),
        "normalized_cross_correlation": (None, [f32p] * 4),
        "zfp2d_lossyCompression": (C.c_int, [P(C.c_void_p), f32p, C.c_int, C.c_int, C.c_float]),
        "zfp2d_lossyDecompression": (C.c_int, [C.c_char_p, f32p, C.c_int, C.c_int, C.c_float]),
        "rtm_zfp_mode_accuracy": (C.c_double, [P(ZfpMode), C.c_double]),
        "rtm_zfp_mode_precision": (C.c_uint, [P(ZfpMode), C.c_uint]),
        "rtm_zfp_mode_rate": (C.c_double, [P(ZfpMode), C.c_double]),
        "rtm_zfp2f_max_bytes": (C.c_size_t, [P(ZfpMode), C.c_uint, C.c_uint]),
        "rtm_zfp2f_encode": (C.c_size_t, [P(ZfpMode), f32p, C.c_uint, C.c_uint, C.c_void_p, C.c_size_t,
                                          C.c_int]),
        "rtm_zfp2f_decode": (C.c_size_t, [P(ZfpMode), C.c_void_p, C.c_size_t, C.c_uint, C.c_uint, f32p]),
        "reading_file": (C.c_int, [C.c_char_p, f32p, C.c_int, C.c_int]),
        "recording_file": (C.c_int, [C.c_char_p, f32p, C.c_int, C.c_int]),
        "extend_matrix": (None, [f32p, f32p]),
        "laplacian_filter2D": (None, [f32p, f32p]),
        "rtm_plan_create": (C.c_int, [P(RtmParams), C.c_int, P(plan_p)]),
        "rtm_plan_destroy": (C.c_int, [plan_p]),
        "rtm_plan_footprint": (C.c_size_t, [P(RtmParams), C.c_int]),
        "rtm_plan_set_tiling": (C.c_int, [plan_p, C.c_int, C.c_int, C.c_int, C.c_int]),
        "rtm_plan_get_tiling": (C.c_int, [plan_p, i32p, i32p, i32p, i32p]),
        "rtm_plan_autotune": (C.c_int, [plan_p, C.c_int, C.c_int, i32p, f64p, i32p]),
        "rtm_batch_get_tiling": (C.c_int, [plan_p, i32p, i32p, i32p, i32p]),
        "rtm_batch_autotune": (C.c_int, [plan_p, C.c_int, C.c_int, i32p, f64p, i32p]),
        "rtm_batch_set_group": (C.c_int, [plan_p, C.c_int]),
        "rtm_batch_get_group": (C.c_int, [plan_p, i32p, i32p]),
        "rtm_batch_autotune_group": (C.c_int, [plan_p, C.c_int, C.c_int, i32p, f64p, i32p]),
        "rtm_batch_set_lanes": (C.c_int, [plan_p, C.c_int]),
        "rtm_batch_get_lanes": (C.c_int, [plan_p, i32p]),
        "rtm_batch_autotune_lanes": (C.c_int, [plan_p, C.c_int, C.c_int, i32p, i32p, f64p, i32p]),
        "rtm_plan_set_velocity": (C.c_int, [plan_p, f32p]),
        "rtm_plan_set_wavelet": (C.c_int, [plan_p, f32p, C.c_int]),
        "rtm_plan_set_seismogram": (C.c_int, [plan_p, f32p]),
        "rtm_plan_zero_image": (C.c_int, [plan_p]),
        "rtm_plan_forward": (C.c_int, [plan_p, C.c_int, C.c_int, C.c_int]),
        "rtm_plan_backward": (C.c_int, [plan_p]),
        "rtm_plan_migrate": (C.c_int, [plan_p, C.c_int, C.c_int]),
        "rtm_plan_get_seismogram": (C.c_int, [plan_p, f32p]),
        "rtm_plan_get_image": (C.c_int, [plan_p, f32p, f32p]),
        "rtm_plan_get_image_f64": (C.c_int, [plan_p, f64p, f64p]),
        "rtm_plan_get_fields": (C.c_int, [plan_p, f32p, f32p]),
        "rtm_plan_get_fields_f64": (C.c_int, [plan_p, f64p, f64p]),
        "rtm_plan_synchronize": (C.c_int, [plan_p]),
        "rtm_plan_stack_image": (C.c_int, [plan_p]),
        "rtm_plan_zero_stack": (C.c_int, [plan_p]),
        "rtm_plan_get_stack": (C.c_int, [plan_p, C.c_int, f32p, f32p]),
        "rtm_plan_last_timing": (C.c_int, [plan_p, P(C.c_double), P(C.c_double), i32p]),
        "rtm_batch_create": (C.c_int, [P(RtmParams), C.c_int, C.c_int, P(plan_p)]),
        "rtm_batch_destroy": (C.c_int, [plan_p]),
        "rtm_batch_footprint": (C.c_size_t, [P(RtmParams), C.c_int, C.c_int]),
        "rtm_batch_size": (C.c_int, [plan_p]),
        "rtm_batch_set_tiling": (C.c_int, [plan_p, C.c_int, C.c_int, C.c_int, C.c_int]),
        "rtm_batch_set_velocity": (C.c_int, [plan_p, C.c_int, f32p]),
        "rtm_batch_set_wavelet": (C.c_int, [plan_p, f32p, C.c_int]),
        "rtm_batch_set_seismogram": (C.c_int, [plan_p, f32p]),
        "rtm_batch_zero_images": (C.c_int, [plan_p]),
        "rtm_batch_migrate": (C.c_int, [plan_p, C.c_int, C.c_int]),
        "rtm_batch_get_image": (C.c_int, [plan_p, C.c_int, f32p, f32p]),
        "rtm_batch_get_image_f64": (C.c_int, [plan_p, C.c_int, f64p, f64p]),
        "rtm_batch_welford": (C.c_int, [plan_p, f64p, f64p]),
        "rtm_batch_welford_noalloc": (C.c_int, [plan_p, f64p, f64p]),
        "rtm_batch_last_timing": (C.c_int, [plan_p, P(C.c_double), P(C.c_double), i32p]),
        "rtm_server_create": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.c_int, P(plan_p)]),
        "rtm_server_destroy": (C.c_int, [plan_p]),
        "rtm_server_set_tiling": (C.c_int, [plan_p, C.c_int, C.c_int, C.c_int, C.c_int]),
        "rtm_server_request_footprint": (C.c_size_t, [P(RtmParams)]),
        "rtm_server_submit": (C.c_int, [plan_p, P(RtmParams), f32p, f32p, f32p, C.c_int, C.c_int]),
        "rtm_server_tick": (C.c_int, [plan_p]),
        "rtm_server_next_done": (C.c_int, [plan_p, i32p]),
        "rtm_server_take_image": (C.c_int, [plan_p, C.c_int, f32p, C.c_size_t]),
        "rtm_server_stats": (C.c_int, [plan_p, i32p, i32p, P(C.c_size_t), P(C.c_longlong)]),
        "rtm_server_set_group_bytes": (C.c_int, [plan_p, C.c_size_t]),
        "rtm_server_get_group_bytes": (C.c_int, [plan_p, P(C.c_size_t), i32p]),
        "rtm_server_set_lanes": (C.c_int, [plan_p, C.c_int]),
        "rtm_server_get_lanes": (C.c_int, [plan_p, i32p]),
        "rtm_server_failures": (C.c_int, [plan_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class RtmError(RuntimeError):
    pass


def _autotune(fn, h, steps):
    cap = 32
    shapes = np.zeros((cap, 4), np.int32)
    ms = np.zeros(cap, np.float64)
    n = C.c_int()
    _check(fn(h, int(steps), cap, iptr(shapes), dptr(ms), C.byref(n)), "autotune")
    return [dict(rows_per_thread=int(a), graph_block=int(b), interior=int(c), seg_rows=int(d), ms_per_step=float(t))
            for (a, b, c, d), t in zip(shapes[:n.value], ms[:n.value])]


def _get_tiling(fn, h):
    v = [C.c_int() for _ in range(4)]
    _check(fn(h, *[C.byref(x) for x in v]), "get_tiling")
    return tuple(x.value for x in v)


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().rtm_last_error()
        raise RtmError(f"{what}: {msg.decode() if msg else 'error'} (rc={rc})")


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def default_params() -> RtmParams:
    """The reference's shipped configuration (SRC/fixed_parameters.h:8-70)."""
    p = RtmParams()
    lib().rtm_params_default(C.byref(p))
    return p


def make_params(**kw) -> RtmParams:
    return default_params().update(**kw)


def set_params(p: RtmParams) -> None:
    _check(lib().rtm_set_params(C.byref(p)), "rtm_set_params")


def footprint(p: RtmParams, real: int = REAL_F32) -> int:
    return int(lib().rtm_plan_footprint(C.byref(p), real))


class Plan:
    """One simulation's device-resident state (rtm_plan)."""

    def __init__(self, p: RtmParams, real: int = REAL_F32):
        self.p = p.copy()
        self.real = real
        h = C.c_void_p()
        _check(lib().rtm_plan_create(C.byref(self.p), real, C.byref(h)), "rtm_plan_create")
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().rtm_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_tiling(self, rows_per_thread: int = 4, graph_block: int = 100, interior: int = 0,
                   seg_rows: int = 0) -> None:
        """Kernel shape (rtm_plan_set_tiling): tile rows per thread, steps per
        graph block, interior kernel (INTERIOR_TILES / _COLUMNS / _MIXED / _TILES8 / _PAIRS) and
        rows per column segment."""
        _check(lib().rtm_plan_set_tiling(self.h, int(rows_per_thread), int(graph_block), int(interior),
                                         int(seg_rows)), "set_tiling")

    def tiling(self):
        """(rows_per_thread, graph_block, interior, seg_rows) in use."""
        return _get_tiling(lib().rtm_plan_get_tiling, self.h)

    def autotune(self, steps: int = 200):
        """rtm_plan_autotune: time every candidate shape, keep the fastest;
        returns the candidates with their ms per step."""
        return _autotune(lib().rtm_plan_autotune, self.h, steps)

    def set_velocity(self, vpad: np.ndarray) -> None:
        v = np.ascontiguousarray(vpad, dtype=np.float32)
        assert v.shape == (self.p.X, self.p.Z), v.shape
        _check(lib().rtm_plan_set_velocity(self.h, fptr(v)), "set_velocity")

    def set_wavelet(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(lib().rtm_plan_set_wavelet(self.h, fptr(w), int(w.size)), "set_wavelet")

    def set_seismogram(self, d: np.ndarray) -> None:
        d = np.ascontiguousarray(d, dtype=np.float32)
        assert d.shape == (self.p.number_of_receivers, self.p.number_of_time_steps), d.shape
        _check(lib().rtm_plan_set_seismogram(self.h, fptr(d)), "set_seismogram")

    def zero_image(self) -> None:
        _check(lib().rtm_plan_zero_image(self.h), "zero_image")

    def forward(self, shot_x: int, shot_z: int, record: bool = False) -> None:
        _check(lib().rtm_plan_forward(self.h, int(shot_x), int(shot_z), int(bool(record))), "forward")

    def backward(self) -> None:
        _check(lib().rtm_plan_backward(self.h), "backward")

    def migrate(self, shot_x: int, shot_z: int) -> None:
        _check(lib().rtm_plan_migrate(self.h, int(shot_x), int(shot_z)), "migrate")

    def seismogram(self) -> np.ndarray:
        d = np.empty((self.p.number_of_receivers, self.p.number_of_time_steps), np.float32)
        _check(lib().rtm_plan_get_seismogram(self.h, fptr(d)), "get_seismogram")
        return d

    def image(self, f64: bool = False):
        shape = (self.p.nx, self.p.nz)
        if f64:
            I, Is = np.empty(shape, np.float64), np.empty(shape, np.float64)
            _check(lib().rtm_plan_get_image_f64(self.h, dptr(I), dptr(Is)), "get_image_f64")
        else:
            I, Is = np.empty(shape, np.float32), np.empty(shape, np.float32)
            _check(lib().rtm_plan_get_image(self.h, fptr(I), fptr(Is)), "get_image")
        return I, Is

    def fields(self, f64: bool = False):
        shape = (self.p.X, self.p.Z)
        if f64:
            a, b = np.empty(shape, np.float64), np.empty(shape, np.float64)
            _check(lib().rtm_plan_get_fields_f64(self.h, dptr(a), dptr(b)), "get_fields_f64")
        else:
            a, b = np.empty(shape, np.float32), np.empty(shape, np.float32)
            _check(lib().rtm_plan_get_fields(self.h, fptr(a), fptr(b)), "get_fields")
        return a, b

    def synchronize(self) -> None:
        _check(lib().rtm_plan_synchronize(self.h), "synchronize")

    def stack_image(self) -> None:
        """rtm_plan_stack_image: add the current images into the device stacks."""
        _check(lib().rtm_plan_stack_image(self.h), "stack_image")

    def zero_stack(self) -> None:
        _check(lib().rtm_plan_zero_stack(self.h), "zero_stack")

    def stack(self, normalize: bool = False):
        """rtm_plan_get_stack: (stack, its Laplacian), [Nx][Nz] float32."""
        st = np.empty((self.p.nx, self.p.nz), np.float32)
        lp = np.empty((self.p.nx, self.p.nz), np.float32)
        _check(lib().rtm_plan_get_stack(self.h, int(bool(normalize)), fptr(st), fptr(lp)), "get_stack")
        return st, lp

    def last_timing(self):
        """(forward ms/step, backward ms/step, steps) of the last passes."""
        f, b = C.c_double(), C.c_double()
        n = C.c_int()
        _check(lib().rtm_plan_last_timing(self.h, C.byref(f), C.byref(b), C.byref(n)), "last_timing")
        return f.value, b.value, n.value


def extend(p: RtmParams, model: np.ndarray) -> np.ndarray:
    """extend_matrix for params p (SRC/generic_functions.c:5-48)."""
    set_params(p)
    m = np.ascontiguousarray(model, dtype=np.float32)
    out = np.zeros((p.X, p.Z), np.float32)
    lib().extend_matrix(fptr(out), fptr(m))
    return out


def zfp_compress(arr: np.ndarray, tol: float) -> bytes:
    """Headerless ZFP fixed-accuracy stream of a [d1][d2] float array with the
    reference's field convention (SRC/compression_tools.c:33)."""
    a = np.ascontiguousarray(arr, dtype=np.float32)
    d1, d2 = a.shape
    buf = C.c_void_p()
    n = lib().zfp2d_lossyCompression(C.byref(buf), fptr(a), d1, d2, C.c_float(tol))
    if n <= 0 or not buf.value:
        raise RtmError("zfp2d_lossyCompression failed: " + (lib().rtm_last_error() or b"").decode())
    out = C.string_at(buf.value, n)
    C.CDLL(None).free(buf)
    return out


def zfp_decompress_file(path: str, d1: int, d2: int, tol: float) -> np.ndarray:
    a = np.zeros((d1, d2), np.float32)
    rc = lib().zfp2d_lossyDecompression(path.encode(), fptr(a), d1, d2, C.c_float(tol))
    if rc != 0:
        raise RtmError(f"zfp2d_lossyDecompression({path}) failed")
    return a


def zfp_mode(mode: str, value) -> ZfpMode:
    """A ZFP lossy mode: ("accuracy", tol), ("precision", bit planes) or
    ("rate", bits per value)."""
    m = ZfpMode()
    if mode == "accuracy":
        lib().rtm_zfp_mode_accuracy(C.byref(m), float(value))
    elif mode == "precision":
        lib().rtm_zfp_mode_precision(C.byref(m), int(value))
    elif mode == "rate":
        lib().rtm_zfp_mode_rate(C.byref(m), float(value))
    else:
        raise ValueError(mode)
    return m


def zfp_encode(a: np.ndarray, mode: str, value, threads: int = 1) -> bytes:
    """ZFP 0.5.5 stream of a[ny][nx] float32 (x fastest) in the given mode,
    by this library's codec (rtm_zfp2f_encode)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    ny, nx = a.shape
    m = zfp_mode(mode, value)
    cap = lib().rtm_zfp2f_max_bytes(C.byref(m), nx, ny)
    buf = C.create_string_buffer(cap)
    n = lib().rtm_zfp2f_encode(C.byref(m), fptr(a), nx, ny, C.cast(buf, C.c_void_p), cap, int(threads))
    if n == 0:
        raise RtmError("rtm_zfp2f_encode: buffer too small")
    return buf.raw[:n]


def zfp_decode(stream: bytes, nx: int, ny: int, mode: str, value) -> np.ndarray:
    """The inverse of zfp_encode; raises if the stream ends before the field."""
    m = zfp_mode(mode, value)
    out = np.zeros((ny, nx), np.float32)
    buf = C.create_string_buffer(bytes(stream), len(stream))
    n = lib().rtm_zfp2f_decode(C.byref(m), C.cast(buf, C.c_void_p), len(stream), nx, ny, fptr(out))
    if n == 0:
        raise RtmError("rtm_zfp2f_decode: the stream ends before the field")
    return out


def zfp_max_bytes(nx: int, ny: int, mode: str, value) -> int:
    return int(lib().rtm_zfp2f_max_bytes(C.byref(zfp_mode(mode, value)), nx, ny))


class Batch:
    """An ensemble of realizations of one geometry on the device (rtm_batch)."""

    def __init__(self, p: RtmParams, nsims: int, real: int = REAL_F32):
        self.p = p.copy()
        self.n = int(nsims)
        self.real = real
        h = C.c_void_p()
        _check(lib().rtm_batch_create(C.byref(self.p), self.n, real, C.byref(h)), "rtm_batch_create")
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().rtm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_tiling(self, rows_per_thread: int = 4, graph_block: int = 100, interior: int = 0,
                   seg_rows: int = 0) -> None:
        _check(lib().rtm_batch_set_tiling(self.h, int(rows_per_thread), int(graph_block), int(interior),
                                          int(seg_rows)), "set_tiling")

    def tiling(self):
        return _get_tiling(lib().rtm_batch_get_tiling, self.h)

    def autotune(self, steps: int = 200):
        """rtm_batch_autotune: shapes timed on launch group 0 at the current
        group size (ms_per_step is that group's)."""
        return _autotune(lib().rtm_batch_autotune, self.h, steps)

    def set_group(self, sims_per_group: int) -> None:
        """rtm_batch_set_group: realizations per launch group (0: all)."""
        _check(lib().rtm_batch_set_group(self.h, int(sims_per_group)), "set_group")

    def group(self):
        """(realizations per launch group, number of groups)."""
        g, n = C.c_int(), C.c_int()
        _check(lib().rtm_batch_get_group(self.h, C.byref(g), C.byref(n)), "get_group")
        return g.value, n.value

    def autotune_group(self, steps: int = 200):
        """rtm_batch_autotune_group: shape-tune each candidate group size and
        keep the group with the least time per realization-step."""
        cap = 32
        gs = np.zeros(cap, np.int32)
        us = np.zeros(cap, np.float64)
        n = C.c_int()
        _check(lib().rtm_batch_autotune_group(self.h, int(steps), cap, iptr(gs), dptr(us), C.byref(n)),
               "autotune_group")
        return [dict(group=int(g), us_per_sim_step=float(t)) for g, t in zip(gs[:n.value], us[:n.value])]

    def set_lanes(self, lanes: int) -> None:
        """rtm_batch_set_lanes: launch groups in flight at once (1 to 4)."""
        _check(lib().rtm_batch_set_lanes(self.h, int(lanes)), "set_lanes")

    def lanes(self) -> int:
        n = C.c_int()
        _check(lib().rtm_batch_get_lanes(self.h, C.byref(n)), "get_lanes")
        return n.value

    def autotune_lanes(self, steps: int = 200):
        """rtm_batch_autotune_lanes: one lane at the current group size, then
        two lanes at smaller groups; keeps the fastest (2% rule)."""
        cap = 16
        ls = np.zeros(cap, np.int32)
        gs = np.zeros(cap, np.int32)
        us = np.zeros(cap, np.float64)
        n = C.c_int()
        _check(lib().rtm_batch_autotune_lanes(self.h, int(steps), cap, iptr(ls), iptr(gs), dptr(us), C.byref(n)),
               "autotune_lanes")
        return [dict(lanes=int(a), group=int(g), us_per_sim_step=float(t))
                for a, g, t in zip(ls[:n.value], gs[:n.value], us[:n.value])]

    def set_velocity(self, sim: int, vpad: np.ndarray) -> None:
        v = np.ascontiguousarray(vpad, dtype=np.float32)
        assert v.shape == (self.p.X, self.p.Z), v.shape
        _check(lib().rtm_batch_set_velocity(self.h, int(sim), fptr(v)), "set_velocity")

    def set_wavelet(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(lib().rtm_batch_set_wavelet(self.h, fptr(w), int(w.size)), "set_wavelet")

    def set_seismogram(self, d: np.ndarray) -> None:
        d = np.ascontiguousarray(d, dtype=np.float32)
        assert d.shape == (self.p.number_of_receivers, self.p.number_of_time_steps), d.shape
        _check(lib().rtm_batch_set_seismogram(self.h, fptr(d)), "set_seismogram")

    def zero_images(self) -> None:
        _check(lib().rtm_batch_zero_images(self.h), "zero_images")

    def migrate(self, shot_x: int, shot_z: int) -> None:
        _check(lib().rtm_batch_migrate(self.h, int(shot_x), int(shot_z)), "migrate")

    def image(self, sim: int, f64: bool = False):
        shape = (self.p.nx, self.p.nz)
        if f64:
            I, Is = np.empty(shape, np.float64), np.empty(shape, np.float64)
            _check(lib().rtm_batch_get_image_f64(self.h, int(sim), dptr(I), dptr(Is)), "get_image_f64")
        else:
            I, Is = np.empty(shape, np.float32), np.empty(shape, np.float32)
            _check(lib().rtm_batch_get_image(self.h, int(sim), fptr(I), fptr(Is)), "get_image")
        return I, Is

    def welford(self):
        shape = (self.p.nx, self.p.nz)
        m, v = np.empty(shape, np.float64), np.empty(shape, np.float64)
        _check(lib().rtm_batch_welford(self.h, dptr(m), dptr(v)), "welford")
        return m, v

    def welford_noalloc(self):
        """welford() through rtm_batch_welford_noalloc: no device allocation
        (a fixed scratch array of the library), the same results."""
        shape = (self.p.nx, self.p.nz)
        m, v = np.empty(shape, np.float64), np.empty(shape, np.float64)
        _check(lib().rtm_batch_welford_noalloc(self.h, dptr(m), dptr(v)), "welford_noalloc")
        return m, v

    def last_timing(self):
        f, b = C.c_double(), C.c_double()
        n = C.c_int()
        _check(lib().rtm_batch_last_timing(self.h, C.byref(f), C.byref(b), C.byref(n)), "last_timing")
        return f.value, b.value, n.value


def batch_footprint(p: RtmParams, nsims: int, real: int = REAL_F32) -> int:
    return int(lib().rtm_batch_footprint(C.byref(p), int(nsims), real))


def request_footprint(p: RtmParams) -> int:
    """HBM bytes a request with params p occupies in a Server while active."""
    return int(lib().rtm_server_request_footprint(C.byref(p)))


class Server:
    """The request server (rtm_server_*): migrations of any grid size are
    queued with submit(), admitted by HBM footprint and advanced together by
    tick() (one block of time steps per call, continuous batching); finished
    images are collected with done()."""

    def __init__(self, hbm_bytes: int, max_slots: int = 64, snapshot_mode: int = SNAP_FP16, device: int = 0):
        h = C.c_void_p()
        _check(lib().rtm_server_create(int(device), int(hbm_bytes), int(max_slots), int(snapshot_mode),
                                       C.byref(h)), "rtm_server_create")
        self.h = h
        self.snapshot_mode = snapshot_mode
        self.params = {}
        self.failed = {}        # request id -> reason, for requests that failed

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().rtm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_tiling(self, rows_per_thread: int = 4, graph_block: int = 100, interior: int = 0,
                   seg_rows: int = 0) -> None:
        _check(lib().rtm_server_set_tiling(self.h, int(rows_per_thread), int(graph_block), int(interior),
                                           int(seg_rows)), "set_tiling")

    def set_group_bytes(self, nbytes: int) -> None:
        """rtm_server_set_group_bytes: per-step working-set budget of a launch
        group (0: every active request in one group)."""
        _check(lib().rtm_server_set_group_bytes(self.h, int(nbytes)), "set_group_bytes")

    def group_bytes(self):
        """(launch-group budget in bytes, launch groups at the last tick)."""
        b, n = C.c_size_t(), C.c_int()
        _check(lib().rtm_server_get_group_bytes(self.h, C.byref(b), C.byref(n)), "get_group_bytes")
        return b.value, n.value

    def set_lanes(self, lanes: int) -> None:
        """rtm_server_set_lanes: launch groups in flight at once (1 to 4)."""
        _check(lib().rtm_server_set_lanes(self.h, int(lanes)), "set_lanes")

    def lanes(self) -> int:
        n = C.c_int()
        _check(lib().rtm_server_get_lanes(self.h, C.byref(n)), "get_lanes")
        return n.value

    def submit(self, p: RtmParams, vpad: np.ndarray, wavelet: np.ndarray, seis: np.ndarray, shot_x: int,
               shot_z: int) -> int:
        q = p.copy().update(snapshot_mode=self.snapshot_mode)
        v = np.ascontiguousarray(vpad, dtype=np.float32)
        assert v.shape == (q.X, q.Z), v.shape
        w = np.ascontiguousarray(wavelet, dtype=np.float32)
        assert w.size >= q.sampling_wavelet
        d = np.ascontiguousarray(seis, dtype=np.float32)
        assert d.shape == (q.number_of_receivers, q.number_of_time_steps), d.shape
        rid = lib().rtm_server_submit(self.h, C.byref(q), fptr(v), fptr(w), fptr(d), int(shot_x), int(shot_z))
        if rid < 0:
            _check(rid, "rtm_server_submit")
        self.params[rid] = (q.nx, q.nz)
        return rid

    def tick(self) -> int:
        n = lib().rtm_server_tick(self.h)
        if n < 0:
            _check(n, "rtm_server_tick")
        return n

    def done(self):
        """Retired requests not yet collected: [(id, image [Nx][Nz])], with
        None for a request that failed (its reason in self.failed[id])."""
        out = []
        rid = C.c_int()
        while lib().rtm_server_next_done(self.h, C.byref(rid)) == 1:
            nx, nz = self.params.pop(rid.value)
            I = np.empty((nx, nz), np.float32)
            rc = lib().rtm_server_take_image(self.h, rid.value, fptr(I), I.size)
            if rc == -2:
                self.failed[rid.value] = lib().rtm_last_error().decode()
                out.append((rid.value, None))
                continue
            _check(rc, "take_image")
            out.append((rid.value, I))
        return out

    def failures(self) -> int:
        """Requests failed since the server was created."""
        return int(lib().rtm_server_failures(self.h))

    def stats(self):
        a, q = C.c_int(), C.c_int()
        b = C.c_size_t()
        st = C.c_longlong()
        _check(lib().rtm_server_stats(self.h, C.byref(a), C.byref(q), C.byref(b), C.byref(st)), "stats")
        return {"active": a.value, "queued": q.value, "bytes_in_use": b.value, "steps": st.value}
