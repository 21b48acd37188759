"""This library's ZFP 0.5.5 stream codec (rtmuq2d_amd/csrc/rtm_zfp.c), pinned
three ways:

* ZFP's own known answers for small 2-D float arrays
  (ZFP_LIBRARY_0.5.5/tests/testzfp.cpp): the test array's checksum
  0x86609589 (:751-754, which pins the restated generator), the fixed-
  accuracy stream sizes 4936 / 11064 / 12520 B at tolerance 1e-3 / 2 eps / 0
  with the error within the tolerance (:846-860, :501-507), the fixed-
  precision sizes 592 / 1328 / 4384 B at 4 / 8 / 16 bit planes, and the
  fixed-rate error bounds 1.5 / 3.663e-3 / 0 at 2 / 8 / 32 bits per value;
* byte for byte against ZFP 0.5.5 compiled from the reference's source
  (oracle/_ref/zfp, a checker) on seeded fields of ragged sizes in every
  mode, encoding and decoding, serial and OpenMP;
* the reference's own streams: the golden fixtures (tests/golden, written by
  that ZFP through the reference's call) are reproduced by
  test_library.py / test_oracle_reference.py.
"""
from __future__ import annotations

import numpy as np
import pytest

import rtmuq2d_amd as R
import zfp_ref

# ---- ZFP's test array (testzfp.cpp:42-78, :133-176, :290-305) --------------


def _refine2d(f: np.ndarray, m: int) -> np.ndarray:
    """f[m][m] -> g[2m][2m]: even points copied, odd ones by the 4-point
    (-1, 9, 9, -1)/16 stencil per axis, periodic, integer division
    truncating toward zero as C's does."""
    w = np.array([-1, 9, 9, -1], np.int64)
    n = 2 * m
    g = np.zeros((n, n), np.int64)
    for y in range(n):
        ys = [(y // 2 + j - 1 + m) % m if y & 1 else y // 2 for j in range(4)]
        for x in range(n):
            xs = [(x // 2 + i - 1 + m) % m if x & 1 else x // 2 for i in range(4)]
            s = 0
            for j in range(4):
                for i in range(4):
                    s += int(w[i] * w[j]) * int(f[ys[j], xs[i]])
            q = abs(s) // 256
            g[y, x] = q if s >= 0 else -q
    return g


def zfp_test_array_2d(n: int = 64) -> np.ndarray:
    """The 'small' 2-D float test array: a 4x4 checkerboard of +-2^16 at
    (1..2, 1..2) refined to n x n, then scaled by 2^-12."""
    f = np.zeros((4, 4), np.int64)
    for y in (1, 2):
        for x in (1, 2):
            f[y, x] = 0x10000 * (1 - 2 * ((x ^ y) & 1))
    m = 4
    while m < n:
        f = _refine2d(f, m)
        m *= 2
    return np.ldexp(f.astype(np.float32), -12).astype(np.float32)


def jenkins_hash(b: bytes) -> int:
    """Jenkins one-at-a-time, 32-bit (testzfp.cpp:290-305)."""
    h = 0
    M = 0xFFFFFFFF
    for c in b:
        h = (h + c) & M
        h = (h + (h << 10)) & M
        h ^= h >> 6
    h = (h + (h << 3)) & M
    h ^= h >> 11
    h = (h + (h << 15)) & M
    return h


@pytest.fixture(scope="module")
def kat_array():
    a = zfp_test_array_2d(64)
    assert jenkins_hash(a.tobytes()) == 0x86609589   # testzfp.cpp:754, small 2-D float
    return a


@pytest.mark.parametrize("tol,size", [(1e-3, 4936), (2 * float(np.finfo(np.float32).eps), 11064), (0.0, 12520)])
def test_zfp_known_answers_fixed_accuracy(kat_array, tol, size):
    """testzfp.cpp:846-860 (sizes) and :501-507 (error within the tolerance
    ZFP reports using, 2^floor(log2 tol))."""
    used = R.zfp_mode("accuracy", tol)
    z = R.zfp_encode(kat_array, "accuracy", tol)
    assert len(z) == size
    back = R.zfp_decode(z, 64, 64, "accuracy", tol)
    tol_used = float(np.float32(np.ldexp(1.0, used.minexp))) if tol > 0 else 0.0
    err = float(np.max(np.abs(back - kat_array)))
    if tol > 0:
        assert err <= tol_used
    assert R.zfp_encode(kat_array, "accuracy", tol, threads=8) == z


@pytest.mark.parametrize("prec,size", [(4, 592), (8, 1328), (16, 4384)])
def test_zfp_known_answers_fixed_precision(kat_array, prec, size):
    """testzfp.cpp fixed-precision sizes, small 2-D float."""
    assert len(R.zfp_encode(kat_array, "precision", prec)) == size


@pytest.mark.parametrize("rate,emax", [(2, 1.500e+00), (8, 3.663e-03), (32, 0.0)])
def test_zfp_known_answers_fixed_rate(kat_array, rate, emax):
    """testzfp.cpp fixed-rate expected maximum errors, small 2-D float; the
    stream takes exactly rate x 16 bits per block."""
    z = R.zfp_encode(kat_array, "rate", rate)
    assert len(z) == (256 * 16 * rate + 63) // 64 * 8
    back = R.zfp_decode(z, 64, 64, "rate", rate)
    assert float(np.max(np.abs(back - kat_array))) <= np.float32(emax)


def test_zfp_decode_refuses_a_short_stream(kat_array):
    z = R.zfp_encode(kat_array, "accuracy", 1e-3)
    with pytest.raises(R.RtmError):
        R.zfp_decode(z[: len(z) // 2], 64, 64, "accuracy", 1e-3)


# ---- byte for byte against ZFP 0.5.5 from the reference's source -----------

needs_ref_zfp = pytest.mark.skipif(not zfp_ref.available(),
                                   reason="oracle/_ref/zfp not built (needs the reference tree)")

SHAPES = [(64, 64), (1, 1), (5, 3), (37, 29), (600, 41), (1701, 20)]
MODES = [("accuracy", 1e-4), ("accuracy", 1e-3), ("accuracy", 0.0), ("accuracy", 3.7e-7),
         ("precision", 3), ("precision", 13), ("precision", 0), ("rate", 1.5), ("rate", 7.25), ("rate", 24)]


def _field(shape, seed):
    ny, nx = shape
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((ny, nx)).astype(np.float32)
    a = (a * np.exp2(rng.integers(-20, 10, (ny, nx)))).astype(np.float32)   # wide exponent range
    a[rng.random((ny, nx)) < 0.1] = 0.0
    if ny * nx >= 64:
        a[: min(ny, 4), : min(nx, 8)] = 0.0                                    # an all-zero block
    return a


@needs_ref_zfp
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("mode,value", MODES)
def test_zfp_streams_equal_reference_zfp(shape, mode, value):
    ny, nx = shape
    a = _field(shape, hash((ny, nx)) % 1000)
    ref = zfp_ref.compress(a, mode, value)
    mine = R.zfp_encode(a, mode, value)
    assert mine == ref, (len(mine), len(ref))
    assert R.zfp_encode(a, mode, value, threads=8) == ref
    assert R.zfp_max_bytes(nx, ny, mode, value) == zfp_ref.max_size(nx, ny, mode, value)
    back_ref = zfp_ref.decompress(ref, nx, ny, mode, value)
    back = R.zfp_decode(ref, nx, ny, mode, value)
    assert np.array_equal(back.view(np.uint32), back_ref.view(np.uint32))


@needs_ref_zfp
def test_zfp_seismogram_streams_equal_reference_zfp_omp():
    """A c1-sized gather (1701 receivers x 4000 samples, the reference's
    field convention nx = receivers over the [receiver][time] array, tol 1e-4):
    this codec's serial and parallel streams equal ZFP's OpenMP stream (the
    reference compresses with zfp_exec_omp, SRC/compression_tools.c:20)."""
    rng = np.random.default_rng(3)
    t = np.arange(4000) * 0.0008
    d = (np.sin(2 * np.pi * 20 * t[None, :] - 0.003 * np.arange(1701)[:, None])
         * np.exp(-t[None, :]) * 1e-2 + 1e-5 * rng.standard_normal((1701, 4000))).astype(np.float32)
    flat = d.reshape(4000, 1701)          # the same memory seen as ny = 4000 rows of nx = 1701
    ref = zfp_ref.compress(flat, "accuracy", float(np.float32(1e-4)), omp=True)
    assert R.zfp_encode(flat, "accuracy", float(np.float32(1e-4)), threads=8) == ref
    assert R.zfp_compress(d, 1e-4) == ref
    back = R.zfp_decode(ref, 1701, 4000, "accuracy", float(np.float32(1e-4)))
    assert np.array_equal(back.view(np.uint32), zfp_ref.decompress(ref, 1701, 4000, "accuracy",
                                                                    float(np.float32(1e-4))).view(np.uint32))
